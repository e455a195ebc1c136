/*
 * mhd.h -- C ABI of the MI355X (gfx950) per-timestep compressible-MHD core.
 *
 * Drop-in for the reference solver's time-step path (p3d "VarElecCond",
 * Fortran 90).  The reference has no FFI: its step is a fixed sequence of
 * argument-less module subroutines acting on module globals, called from
 * `program pd3` (src/main.F90:340-747).  Each entry point below replaces one
 * of those subroutines (cited per function); the host keeps the reference's
 * loop structure, INPUT parameters and dump layout, and calls these through
 * iso_c_binding (Fortran) or ctypes (Python).
 *
 * Conventions
 *  - Host arrays are the reference's Fortran (nx,ny,nz) column-major FP64
 *    buffers: element (i,j,k), 0-based, at i + nx*(j + ny*k).  They are
 *    borrowed for the duration of the call only; device state is owned by
 *    the context.
 *  - Every function returns 0 on success.  Non-zero codes mirror the
 *    reference's iret / abort_comms codes (MHD_E_*); mhd_strerror() names them.
 *  - One context = one box on one GPU = one HIP stream.  Calls on one context
 *    are not thread-safe; distinct contexts are independent.
 */
#ifndef MHD_H
#define MHD_H

#ifdef __cplusplus
extern "C" {
#endif

/* return codes (src/fft.F90:112-119, src/main.F90:216-219, src/dbydz.F90
 * abort_comms(7), src/main.F90:589 abort_comms(12)) */
enum {
    MHD_OK = 0,
    MHD_E_ODD_NX = -1,       /* init_fft iret = -1 */
    MHD_E_ODD_NY = -2,       /* init_fft iret = -2 */
    MHD_E_ALLOC = -5,        /* allocation failure */
    MHD_E_NZ = 1,            /* fewer than six layers (abort_comms(1)) */
    MHD_E_FLAG = 7,          /* unknown derivative flag (abort_comms(7)) */
    MHD_E_DIVB = 12,         /* |div B| > 0.01 (abort_comms(12)) */
    MHD_E_ARG = -10,         /* null or inconsistent argument */
    MHD_E_SIZE = -20,        /* horizontal size has no transform kernel */
    MHD_E_DEVICE = -30,      /* HIP runtime error */
    MHD_E_STATE = -40,       /* call out of order (e.g. step before state) */
    MHD_E_SCRATCH = -50      /* box too large for a no-allocation entry point */
};

enum { MHD_PREC_FP64 = 0, MHD_PREC_MIXED = 1 };

/* Physical parameters after the INPUT flag-2 conversion (Ra, Q -> CK, F;
 * src/initial.F90:183-195), in src/params.f90 order, plus extension knobs. */
typedef struct mhd_params {
    double gamma, sigma, theta, ck, cm, f, tayl, psi, zeta, xmax, ymax, sf;
    double t_amp;   /* T noise amplitude of static() (reference hard-codes 0.1) */
    double b_amp;   /* Bz noise amplitude (extension; 0 = reference) */
    /* magnetic diffusivity eta(z) = zeta C_k [1 + (eta_c - 1)(1 + tanh((z - eta_z0)
     * / eta_w)) / 2] (extension; eta_c <= 1 keeps the reference's constant eta) */
    double eta_c, eta_z0, eta_w;
    int opts;       /* MHD_OPT_* switches of src/control.f90:23-34 (0 = reference defaults) */
} mhd_params_t;

/* The reference's physics switches (src/main.F90:109-116, all off there):
 * RADBC     radiative top T condition, dT/dz = theta T^4 (src/dbydz.F90:4037-4110)
 * CONSTFLUX fixed-flux bottom T condition, dT/dz = theta (src/dbydz.F90:4133-4171)
 * PERFECT   perfectly conducting walls (src/dbydz.F90:4227-4400), static B = x
 * KINEMATIC F = 0 with the field still evolved (src/initial.F90:314-318)
 * DYNAMO    F = 1 and a seed field at mag_init (src/initial.F90:305-309,
 *           src/magnetic.f90:29-57) */
enum { MHD_OPT_RADBC = 1, MHD_OPT_CONSTFLUX = 2, MHD_OPT_PERFECT = 4, MHD_OPT_KINEMATIC = 8,
       MHD_OPT_DYNAMO = 16 };

/* Driver scalars of src/main.F90 (dt history, AB start-up flags, counters). */
typedef struct mhd_scalars {
    double dt, dt1, dt2, ttime, bzbar;
    int nit, iter, first, second;
} mhd_scalars_t;

/* Diagnostics of src/main.F90:552-650. */
typedef struct mhd_diag {
    double ubar, vbar;        /* GlobalAverages (already subtracted) */
    double kinenergy;         /* kinenergy */
    double magenergy;         /* F * menergy */
    double nusselt;           /* nusselt(Nuss) */
    double divbmax, divbmin;  /* checkdivb */
    double dt;                /* current dt */
    int nit;
} mhd_diag_t;

typedef struct mhd_ctx mhd_ctx_t;

/* Ra, Q -> CK, F (INPUT flag 2), src/initial.F90:183-195 */
void mhd_ra_q_to_ck_f(double gamma, double sigma, double theta, double cm, double zeta,
                      double ra, double q, double* ck, double* f);

/* set_constants + init_fft + init_dbydz + alloc_arrays
 * (src/constants.f90:56-109, src/fft.F90:88-223, src/dbydz.F90:59-364,
 * src/arrays.f90:81-300).  precision: MHD_PREC_FP64 or MHD_PREC_MIXED. */
int mhd_create(const mhd_params_t* p, int nx, int ny, int nz, int precision, mhd_ctx_t** out);
/* close_arrays + close_fft (src/arrays.f90:306-361, src/fft.F90:228-250) */
void mhd_destroy(mhd_ctx_t* ctx);

/* static + ini_trans_params + mag_init (src/initial.F90:25-86,
 * src/trans_params.f90:15-69, src/magnetic.f90:5-279) */
int mhd_static(mhd_ctx_t* ctx);
/* ReadNetCDF's state hand-over + ini_trans_params + mag_init: load the 8
 * physical fields (src/netcdf.F90 ReadNetCDF) and rebuild pol/tor on device */
int mhd_set_state(mhd_ctx_t* ctx, const double* r, const double* t, const double* u,
                  const double* v, const double* w, const double* bx, const double* by,
                  const double* bz);
/* the 8 dumped fields (WriteNetCDF, src/netcdf.F90:372-580); any may be NULL */
int mhd_get_state(mhd_ctx_t* ctx, double* r, double* t, double* u, double* v, double* w,
                  double* bx, double* by, double* bz);

/* field access by id (MHD_F_*); level 1..3 selects an RHS history level
 * (udot(:,:,:,level)); ignored for other fields.  Mean fields and their
 * histories have nz entries, kappa/dkappa nz entries. */
enum {
    MHD_F_U = 0, MHD_F_V, MHD_F_W, MHD_F_T, MHD_F_R, MHD_F_BX, MHD_F_BY, MHD_F_BZ,
    MHD_F_POL, MHD_F_TOR, MHD_F_JSQ,
    MHD_F_UDOT, MHD_F_VDOT, MHD_F_WDOT, MHD_F_RDOT, MHD_F_TDOT, MHD_F_POLDOT, MHD_F_TORDOT,
    MHD_F_BXBAR, MHD_F_BYBAR, MHD_F_BXBARDOT, MHD_F_BYBARDOT, MHD_F_KAPPA, MHD_F_DKAPPA,
    MHD_F_COUNT
};
int mhd_get_field(mhd_ctx_t* ctx, int id, int level, double* dst);
int mhd_set_field(mhd_ctx_t* ctx, int id, int level, const double* src);
int mhd_get_scalars(mhd_ctx_t* ctx, mhd_scalars_t* s);
int mhd_set_scalars(mhd_ctx_t* ctx, const mhd_scalars_t* s);

/* The per-step subroutines of the reference, on device state.  The RHS
 * routines write history level 3 (udot(:,:,:,3) etc.). */
int mhd_hydro_deriv(mhd_ctx_t* ctx);   /* src/hydro_deriv.f90:5-505 */
int mhd_mag_deriv(mhd_ctx_t* ctx);     /* src/magnetic.f90:285-487 */
int mhd_bandj(mhd_ctx_t* ctx);         /* src/magnetic.f90:492-664 */
int mhd_magdealias(mhd_ctx_t* ctx);    /* src/magnetic.f90:666-717 */
int mhd_boundary(mhd_ctx_t* ctx);      /* src/dbydz.F90:3686-4197 (defaults) */
int mhd_mag_boundary(mhd_ctx_t* ctx);  /* src/dbydz.F90:4205-4791 (vertical field) */
int mhd_mag_init(mhd_ctx_t* ctx);      /* src/magnetic.f90:5-279 */

/* nsteps iterations of the time-loop body src/main.F90:340-683:
 * hydro_deriv, mag_deriv, AB3 coefficients and update (with Euler/AB2
 * start-up, density floor), dealias every 10th step, magdealias, bandj,
 * boundary, mag_boundary, diagnostics every ndu06 steps (ndu06 <= 0: never)
 * and, when dt_fixed <= 0, calcdt every 10th step.  dt_fixed > 0 holds dt
 * fixed.  `last` (may be NULL) receives the last diagnostics computed.
 * Returns MHD_E_DIVB when checkdivb exceeds 0.01, as abort_comms(12). */
int mhd_step(mhd_ctx_t* ctx, int nsteps, double dt_fixed, int ndu06, mhd_diag_t* last);

/* diagnostics (src/utils.F90, src/dbydz.F90:4796-4925) */
int mhd_global_averages(mhd_ctx_t* ctx, double out[4]); /* ubar, vbar, ke, me (unscaled); subtracts ubar/vbar */
int mhd_nusselt(mhd_ctx_t* ctx, double* nu);
int mhd_checkdivb(mhd_ctx_t* ctx, double out[2]);       /* divbmax, divbmin */
int mhd_sminmax(mhd_ctx_t* ctx, int field_id, double out[2]); /* min, max */
int mhd_calcdt(mhd_ctx_t* ctx, double* dt);

/* Operator hooks on host buffers (operators of src/fft.F90 and
 * src/dbydz.F90).  in/out are (nx,ny,nz) arrays, or nz-vectors for the 1D
 * ops; out2 is used by D2BYDX/D2BYDY (second derivative).  D1UPBYDZ reads
 * the context's w as upwind selector, like the reference's global w. */
enum {
    MHD_OP_D1BYDX = 1, MHD_OP_D1BYDY, MHD_OP_D2BYDX, MHD_OP_D2BYDY,
    MHD_OP_D1BYDZ, MHD_OP_D1UPBYDZ, MHD_OP_D2BYDZ, MHD_OP_DEALIAS,
    MHD_OP_FORWARDFFT, MHD_OP_INVERSEFFT, MHD_OP_D1BYDZ1D, MHD_OP_D2BYDZ1D
};
int mhd_op(mhd_ctx_t* ctx, int op, const double* in, double* out, double* out2, int flag);

/* Event timing of one kernel family over subsequent mhd_step calls (HIP
 * events recorded on the context's stream around every launch of it).
 * kernel: MHD_K_*; enable=1 resets and starts, enable=0 stops.  Query with
 * mhd_kernel_time(): total milliseconds and number of launches. */
enum { MHD_K_NONE = 0, MHD_K_XDERIV2 = 1, MHD_K_YDERIV2 = 2, MHD_K_ZDERIV = 3, MHD_K_HYDRO = 4,
       MHD_K_AB3 = 5 };
int mhd_kernel_timing(mhd_ctx_t* ctx, int kernel, int enable);
int mhd_kernel_time(mhd_ctx_t* ctx, double* total_ms, long* launches);

/* Dump / restart files in the reference's layout (CreateNetCDF, WriteNetCDF,
 * ReadNetCDF: src/netcdf.F90:39-1287): netCDF classic (CDF-1; CDF-2, the
 * 64-bit-offset variant, when a variable would begin past 2 GiB), dimensions
 * nx, ny, nz, one; global attributes version, history, gamma .. ymax;
 * variables time, iteration, timestep and R, T, U, V, W, Bx, By, Bz over
 * (nz, ny, nx), double (iprec 8) or float (iprec 4).  Host-only, no device
 * calls, no libnetcdf.  fields[] = {R, T, U, V, W, Bx, By, Bz}; the reader
 * accepts NULL fields to query meta only, and converts float files to double.
 * Errors: -13 cannot open/create (abort_comms(13)), -14 write failed,
 * -15 not a readable classic file with those variables. */
typedef struct mhd_dump {
    int nx, ny, nz;
    int version;
    int iteration;
    double time, timestep;
    double gamma, sigma, theta, ck, cm, f, tayl, psi, zeta, xmax, ymax;
    char history[256];
} mhd_dump_t;
int mhd_write_dump(const char* path, const mhd_dump_t* meta, const double* const fields[8], int iprec);
int mhd_read_dump(const char* path, mhd_dump_t* meta, double* const fields[8]);
/* The header mhd_write_dump writes for this shape and precision (the bytes
 * before the time variable's value): returns its length, copied to buf when
 * cap is large enough, and the format version (1 or 2) in *version; or
 * MHD_E_ARG (a field of 4 GiB or more, which neither format holds).  No
 * counterpart in the reference, whose nf_create lays the header out. */
long mhd_dump_header(const mhd_dump_t* meta, int iprec, unsigned char* buf, long cap, int* version);

/* ---- autotuner (no counterpart in the reference) ---------------------------
 * Times this box's launch shapes on this GPU and keeps the fastest in the
 * context: per transform direction and kind (first derivatives, first and
 * second derivatives, real to halfcomplex and back) the radix plan (the
 * default or one of the length's alternatives) together with the
 * persistent-grid multiple (1 or 2 x resident), on a launch of that kind
 * shaped like the step's; the planes per thread of the column-walking
 * hydro kernel, in the variant the step runs (with the fused AB3 update
 * unless a closure is attached; the magnetic column kernels keep 8 planes).
 * Needs a state; overwrites only work arrays, the buffers the next step's
 * new state goes to and the newest RHS level (all rewritten by every step
 * before they are read), so it may run between steps. */
enum { MHD_TK_DERIV1 = 0, MHD_TK_DERIV2 = 1, MHD_TK_R2HC = 2, MHD_TK_HC2R = 3, MHD_TK_N = 4 };
typedef struct mhd_tune {
    /* radix plan per transform kind (index MHD_TK_*): 0 the default, 1 and 2
     * the length's alternatives where it has them */
    int variant_x[MHD_TK_N], variant_y[MHD_TK_N];
    /* persistent-grid multiple per transform kind (1 .. 8) */
    int mult_x[MHD_TK_N], mult_y[MHD_TK_N];
    int zc;
    double ms_x, ms_y, ms_col;   /* the tuned d/dx+d2/dx2, d/dy+d2/dy2 and hydro column launches */
} mhd_tune_t;
int mhd_autotune(mhd_ctx_t* c, mhd_tune_t* out);
/* apply given choices (e.g. tuned on an earlier run); MHD_E_ARG if invalid */
int mhd_set_tune(mhd_ctx_t* c, const mhd_tune_t* t);

/* ---- ensembles of independent boxes (no counterpart in the reference) ------
 * The boxes (contexts with a state, any sizes, parameters and precisions on
 * this GPU) are brought to steady AB3 coefficients at the fixed step dt (up
 * to two start-up steps each).  The kernel launches of MHD_ENSEMBLE_BLOCK
 * steps of every box are then recorded, and the k-th launches of all boxes
 * of a class (boxes with the same horizontal lengths, precision and options
 * launch the same kernels) become one launch over a table of the boxes'
 * arguments and grids; these batched launches, one graph branch per class,
 * are captured into one hipGraph.  mhd_ensemble_step launches it
 * nsteps / block times (nsteps a multiple of the block) and advances each
 * box's driver scalars; results equal mhd_step(box, nsteps, dt, 0, NULL) per
 * box, bit for bit.  Diagnostics are taken per box between calls.  A box
 * must not be stepped on its own while in an ensemble. */
typedef struct mhd_ensemble mhd_ensemble_t;
int mhd_ensemble_create(mhd_ctx_t* const* boxes, int nbox, double dt, mhd_ensemble_t** out);
int mhd_ensemble_step(mhd_ensemble_t* e, int nsteps);
int mhd_ensemble_block(void);
long mhd_ensemble_nodes(mhd_ensemble_t* e);   /* graph nodes (batched kernels and joins) */
int mhd_ensemble_classes(mhd_ensemble_t* e);  /* classes of boxes (graph branches) */
void mhd_ensemble_destroy(mhd_ensemble_t* e);

/* ---- resampling and coarse-graining (host arrays in, host arrays out) ---- */
/* expand (interpolate/fft.F90:832-927, called per field at
 * interpolate/main.F90:299-306): a (nx,ny,nz) field becomes (2nx,2ny,nz) by
 * zero-padding the spectrum of every plane, the coarse Nyquist coefficient
 * kept as a cosine of the fine grid.  Codes: -1/-2 odd nx/ny, -20 a length of
 * nx, 2nx, ny, 2ny without a transform kernel. */
int mhd_expand(int nx, int ny, int nz, const double* in, double* out);
/* mhd_expand on device arrays, allocating nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  in
 * (nx,ny,nz) and out (2nx,2ny,nz) are device pointers of the current
 * device; in is only read.  The work is queued on `stream` (a hipStream_t;
 * NULL = the null stream) and is complete when the stream is.  Work memory
 * is a fixed device array of MHD_NOALLOC_SCRATCH_BYTES per GPU, which must
 * hold 2 nx ny nz + 6 (nx + ny) doubles: larger boxes return MHD_E_SCRATCH.
 * Every call on one GPU shares that array, so calls that overlap on one GPU
 * (other streams, other threads) race.  Results equal mhd_expand's. */
enum { MHD_NOALLOC_SCRATCH_BYTES = 64 << 20 };
int mhd_expand_noalloc(int nx, int ny, int nz, const double* in, double* out, void* stream);
/* Block box filter of factor f (the closure data path: filtered fine fields
 * as coarse targets): out(I,J,K) = mean of in over the f x f x f block at
 * (fI, fJ, fK); (nx,ny,nz) -> (nx/f,ny/f,nz/f).  No counterpart in the
 * reference.  MHD_E_ARG when f does not divide nx, ny and nz. */
int mhd_box_filter(int nx, int ny, int nz, int f, const double* in, double* out);

/* ---- learned subgrid closure for coarse runs (no counterpart in the
 * reference) ------------------------------------------------------------------
 * A stencil MLP maps the 8 coarse fields (r t u v w bx by bz) on a cell's
 * 3x3x3 neighbourhood (216 inputs; x, y periodic, z clamped) through two
 * 64-unit GELU layers to 12 subgrid fluxes: tau_xx yy zz xy xz yz, eps_x y z,
 * q_x y z.  Its targets are filters F of fine states (nx, ny, nz) onto the
 * coarse grid (nx/f, ny/f, cz = nz/f): the f x f block mean horizontally
 * (the coarse-graining of interpolate/main.F90:299-306) and, in z, centred on
 * the coarse grid's own planes Z_K = K/(cz-1): the mean over
 * [Z_K - H/2, Z_K + H/2] (H = 1/(cz-1)) of the fine profile's linear
 * interpolant, and the fine wall planes themselves at K = 0 and cz-1, so
 * filtered fields keep the wall values:
 *   tau_ij = F(r) [F(u_i u_j) - F(u_i) F(u_j)]
 *            - fmag ([F(B_i B_j) - F(B_i) F(B_j)] - delta_ij (trace) / 2)
 *   eps = F(u x B) - F(u) x F(B),   q_i = F(u_i T) - F(u_i) F(T).
 * bf16 weights and activations, FP32 accumulation (matrix cores); FP32 master
 * weights trained with Adam on the GPU.  Parameters are packed as
 * W1[216][64] b1[64] W2[64][64] b2[64] W3[64][12] b3[12]; norm[40] holds the
 * field means and deviations (8 + 8) and the target means and deviations
 * (12 + 12) that inputs and outputs are scaled by. */
typedef struct mhd_closure mhd_closure_t;
int mhd_closure_nparam(void);
int mhd_closure_create(unsigned seed, mhd_closure_t** out);
void mhd_closure_destroy(mhd_closure_t* c);
int mhd_closure_set_params(mhd_closure_t* c, const float* p, int n);
int mhd_closure_get_params(mhd_closure_t* c, float* p, int n);
/* box-filtered fields and targets of a fine state (host arrays, fields in
 * the order r t u v w bx by bz), computed on the GPU */
int mhd_closure_targets(int nx, int ny, int nz, int f, double fmag, const double* const fine[8],
                        double* const coarse[8], double* const target[12]);
/* append one coarse snapshot (all snapshots share one shape) to the data set */
int mhd_closure_add(mhd_closure_t* c, int cx, int cy, int cz, const double* const coarse[8],
                    const double* const target[12]);
/* the same for the current state of a fine run: the box-filtered fields and
 * targets (mhd_closure_targets' filter, factor f, fmag = the run's F) are
 * computed from the run's device arrays straight into the data set, with no
 * host round trip of the fields */
int mhd_closure_add_state(mhd_closure_t* c, mhd_ctx_t* fine, int f);
int mhd_closure_samples(mhd_closure_t* c, long* n);
/* the scaling fitted to the data set (until set explicitly) */
int mhd_closure_get_norm(mhd_closure_t* c, double norm[40]);
int mhd_closure_set_norm(mhd_closure_t* c, const double norm[40]);
/* `steps` Adam steps on batches of `batch` (a multiple of 32) samples drawn
 * from the data set; *loss = the last batch's mean squared error (scaled) */
int mhd_closure_train(mhd_closure_t* c, int steps, int batch, double lr, double* loss);
/* loss and gradient over the listed samples (snapshot * cells + cell), no update */
int mhd_closure_grad(mhd_closure_t* c, int ncell, const long* cells, double* loss, float* grad);
/* mean squared error (scaled) over the whole data set */
int mhd_closure_eval(mhd_closure_t* c, double* loss);
/* the 12 fluxes of coarse fields (host arrays) */
int mhd_closure_predict(mhd_closure_t* c, int nx, int ny, int nz, const double* const coarse[8],
                        double* const out[12]);
/* average times of one forward pass over the data set and of one training
 * step of `batch` samples (HIP events) */
int mhd_closure_time(mhd_closure_t* c, int batch, int reps, double* ms_fwd, double* ms_train, long* n_fwd);
/* Attach a trained closure to a coarse run (NULL detaches): every step then
 * adds -div(tau)/r to the momentum right-hand sides, -div(q) to T's and eps
 * to the EMF.  The closure must outlive the attachment.  MHD_E_ARG with the
 * eta(z) extension on. */
int mhd_closure_attach(mhd_ctx_t* ctx, mhd_closure_t* c);

/* Waits for all work of the context; returns any pending device error. */
int mhd_sync(mhd_ctx_t* ctx);
/* Bytes of device memory held by the context. */
long mhd_device_bytes(mhd_ctx_t* ctx);
const char* mhd_strerror(int code);

#ifdef __cplusplus
}
#endif
#endif
