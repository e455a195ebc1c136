/*
 * mhd_oracle.h -- CPU oracle for the per-timestep compressible-MHD core.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, serial restatement of
 * the reference solver's hot path (p3d "VarElecCond", Fortran 90).  It is the
 * checker that the HIP implementation is compared against; it is never linked
 * into, called by, or shipped as part of the product library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Array layout follows the reference exactly: every 3-D field is Fortran
 * f(nx,ny,nz) column-major, i.e. element (i,j,k) (0-based here) sits at
 * i + nx*(j + ny*k).  RHS histories hold 3 levels back to back
 * (level l at offset l*nx*ny*nz), matching udot(nx,ny,nz,3).
 *
 * Return codes mirror the reference's abort/iret codes:
 *   0 ok, -1 odd nx, -2 odd ny, -5 allocation, 1 nz < 6,
 *   7 unknown derivative flag, 12 div B > 0.01.
 */
#ifndef MHD_ORACLE_H
#define MHD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* Parameter block, in the order of the reference's INPUT file after the
 * Ra/Q -> CK/F conversion (src/initial.F90:183-195).  Extension knobs follow
 * the reference's fields. */
typedef struct orc_params {
    double gamma, sigma, theta, ck, cm, f, tayl, psi, zeta, xmax, ymax, sf;
    /* extensions (defaults reproduce the reference):                       */
    double t_amp;      /* T perturbation amplitude (reference: 0.1)          */
    double b_amp;      /* Bz perturbation amplitude (reference: 0, absent)   */
    /* magnetic diffusivity eta(z) = zeta C_k [1 + (eta_c - 1)(1 + tanh((z -
     * eta_z0) / eta_w)) / 2]; eta_c <= 1 keeps the reference's constant eta */
    double eta_c, eta_z0, eta_w;
    int opts;          /* ORC_OPT_* switches of src/control.f90:23-34 (all off) */
} orc_params;

/* the reference's compile-time switches (src/main.F90:109-116) */
enum {
    ORC_OPT_RADBC = 1,      /* radiative top T condition, src/dbydz.F90:4037-4110 */
    ORC_OPT_CONSTFLUX = 2,  /* fixed-flux bottom T condition, 4133-4171          */
    ORC_OPT_PERFECT = 4,    /* perfectly conducting walls, 4227-4400; B0 = x     */
    ORC_OPT_KINEMATIC = 8,  /* F = 0 but B evolved, src/initial.F90:314-318       */
    ORC_OPT_DYNAMO = 16     /* F = 1 and a seed field, src/magnetic.f90:29-57    */
};

/* field ids for orc_field() */
enum {
    ORC_U = 0, ORC_V, ORC_W, ORC_T, ORC_R, ORC_BX, ORC_BY, ORC_BZ,
    ORC_POL, ORC_TOR, ORC_JSQ, ORC_EX, ORC_EY, ORC_EZ,
    ORC_WK1, ORC_WK2, ORC_WK3, ORC_WK4,
    ORC_UDOT, ORC_VDOT, ORC_WDOT, ORC_TDOT, ORC_RDOT, ORC_POLDOT, ORC_TORDOT,
    ORC_BXBAR, ORC_BYBAR, ORC_KAPPA, ORC_DKAPPA, ORC_BXBARDOT, ORC_BYBARDOT,
    ORC_AKX1, ORC_AKX2, ORC_AKY1, ORC_AKY2,
    ORC_JX, ORC_JY, ORC_JZ,   /* current density, kept when eta(z) is on */
    ORC_ETAR,                 /* eta(z) / (zeta C_k), nz entries           */
    ORC_NFIELDS
};

/* scalar driver state ids for orc_scalar()/orc_set_scalar() */
enum {
    ORC_S_DT = 0, ORC_S_DT1, ORC_S_DT2, ORC_S_TTIME, ORC_S_BZBAR,
    ORC_S_NIT, ORC_S_ITER, ORC_S_FIRST, ORC_S_SECOND,
    ORC_S_GCK, ORC_S_SCK, ORC_S_ZCK, ORC_S_DZ, ORC_S_XN, ORC_S_XZN, ORC_S_NUSSTMP,
    ORC_NSCALARS
};

/* Ra/Q -> CK/F (INPUT flag 2), src/initial.F90:183-195 */
void orc_ra_q_to_ck_f(double gamma, double sigma, double theta, double cm,
                      double zeta, double ra, double q, double* ck, double* f);

orc_ctx* orc_create(const orc_params* p, int nx, int ny, int nz, int* ier);
void     orc_destroy(orc_ctx* c);
double*  orc_field(orc_ctx* c, int id);
double   orc_scalar(orc_ctx* c, int id);
void     orc_set_scalar(orc_ctx* c, int id, double v);

/* 1-D real transforms in FFTW-2 halfcomplex order, unnormalised */
void orc_r2hc(int n, const double* in, double* out);
void orc_hc2r(int n, const double* in, double* out);

/* horizontal spectral operators (src/fft.F90) */
void orc_d1bydx(orc_ctx* c, const double* fin, double* fout);
void orc_d1bydy(orc_ctx* c, const double* fin, double* fout);
void orc_d2bydx(orc_ctx* c, const double* fin, double* f1, double* f2);
void orc_d2bydy(orc_ctx* c, const double* fin, double* f1, double* f2);
void orc_dealias(orc_ctx* c, double* z);
void orc_forwardfft(orc_ctx* c, const double* fin, double* fout);
void orc_inversefft(orc_ctx* c, const double* fin, double* fout);

/* vertical finite differences (src/dbydz.F90, serial branches) */
int orc_d1bydz(orc_ctx* c, const double* fin, double* fout, int flag);
int orc_d1upbydz(orc_ctx* c, const double* fin, double* fout, int flag,
                 const double* w);
int orc_d2bydz(orc_ctx* c, const double* fin, double* fout, int flag);
int orc_d1bydz1D(orc_ctx* c, const double* fin, double* fout, int flag);
int orc_d2bydz1D(orc_ctx* c, const double* fin, double* fout, int flag);

/* physics (src/initial.F90, src/trans_params.f90, src/magnetic.f90,
 * src/hydro_deriv.f90, src/dbydz.F90) */
void orc_ran0_fill(int id, double* a, int n, double amp);
void orc_ran0_seq(int idum, int n, double* out);
void orc_ab3_coeffs(int* first, int* second, double dt, double* dt1, double* dt2, double c[3]);
void orc_static(orc_ctx* c);
void orc_ini_trans_params(orc_ctx* c);
int  orc_mag_init(orc_ctx* c);
int  orc_hydro_deriv(orc_ctx* c);
int  orc_mag_deriv(orc_ctx* c);
int  orc_bandj(orc_ctx* c);
void orc_magdealias(orc_ctx* c);
void orc_boundary(orc_ctx* c);
void orc_mag_boundary(orc_ctx* c);

/* diagnostics (src/utils.F90, src/dbydz.F90:4796-4925) */
void   orc_global_averages(orc_ctx* c, double out[4]); /* ubar,vbar,ke,me(raw) */
double orc_nusselt(orc_ctx* c);
int    orc_checkdivb(orc_ctx* c, double out[2]);        /* max, min */
void   orc_sminmax(orc_ctx* c, const double* z, double out[2]); /* min,max */
double orc_calcdt(orc_ctx* c);

/* Time loop body of src/main.F90:340-683 repeated nsteps times.
 *  dt_fixed > 0 : use that dt every step (no calcdt);
 *  dt_fixed <= 0: dt from calcdt at start (if not set) and every 10 steps.
 *  ndu06 > 0    : diagnostics (with the CoM subtraction) every ndu06 steps.
 *  diag (may be NULL) receives the last diagnostics:
 *    [ubar, vbar, kinenergy, magenergy(F-scaled), Nu, divbmax, divbmin]   */
int orc_step(orc_ctx* c, int nsteps, double dt_fixed, int ndu06, double* diag);

/* learned subgrid closure (closure_oracle.c): targets from fine fields,
 * bf16-rounded MLP forward and its gradient.  norm[40] = mu[8], sd[8] of the
 * fields, tmu[12], tsd[12] of the targets. */
int orc_closure_nparam(void);
int orc_closure_attach(orc_ctx* c, const double* params, const double norm[40]);
int orc_closure_fix_fluxes(orc_ctx* c, const double* flux);
float orc_bf16(float x);
double orc_zweight(int nz, int cz, int K, int k);
int orc_closure_targets(int nx, int ny, int nz, int f, double fmag, const double* const fine[8],
                        double* const coarse[8], double* const target[12]);
int orc_mlp_forward(int nx, int ny, int nz, const double* const coarse[8], const double* params,
                    const double norm[40], int round, double* const out[12]);
int orc_mlp_grad(int nx, int ny, int nz, const double* const* coarse, const double* const* target,
                 const double norm[40], const double* params, int ncell, const long* cells, int round,
                 double* loss, double* grad);

#ifdef __cplusplus
}
#endif
/* OpenMP threads used by the plane / point loops (results do not depend on it) */
int orc_threads(void);
void orc_set_threads(int n);

/* expand (interpolate/fft.F90:832-927): (nx,ny,nz) -> (2nx,2ny,nz); 0 or -1 / -5 */
int orc_expand(int nx, int ny, int nz, const double* fin, double* fout);
/* f-fold block box filter: (nx,ny,nz) -> (nx/f,ny/f,nz/f); 0 or -1 */
int orc_box_filter(int nx, int ny, int nz, int f, const double* fin, double* fout);

#endif
