// physics.hpp -- host interface of the vertical-stencil, right-hand-side,
// time-update, boundary and reduction kernels.
#pragma once

#include "common.hpp"

namespace mhd {

// coefficients of src/constants.f90:56-109 and src/dbydz.F90:355-362
struct Phys {
    double gck, sck, zck, scko3, sckg, fzckg, thmplus1, fo2, gm1, cor, cpsi, spsi, f, theta;
    double ho12, sc1, xn, xzn, dz, dd, g1, nusstmp, sf;
    double dtvisc0x, dtvisc0z, dtdiff0x, dtdiff0z;
    int mag;  // f != 0
};

// ---- batched z-derivatives --------------------------------------------------
constexpr int Z_MAX_JOBS = 16;
struct ZJob {
    const void* in;
    const void* in1;   // LD_PROD second factor
    int in_f32;
    int load_op;       // 0 plain, 1 scale*in*in1
    double scale;
    void* out;
    int out_f32;
    int kind;          // ZKind
    int flag;
};
struct ZLaunch {
    Geom g;
    double ho12, sc1;
    const void* w;     // upwind selector (the solver's w)
    int w_f32;
    int njobs;
    ZJob job[Z_MAX_JOBS];
};
int launch_z(const ZLaunch& L, hipStream_t s);

// ---- fused hydro right-hand side (src/hydro_deriv.f90:5-505) ----------------
// One column-walking kernel computes every z-derivative of hydro_deriv
// (d1bydz / d1upbydz / d2bydz with the reference's flags) from register
// windows, the derivatives of the three composite fluxes (lines 181-184,
// 258, 300, 391, 434) by linear splitting, the boundary d(rho)/dz
// substitution (455-470) and the division by rho (487-497).  Inputs are
// the state and the horizontal derivatives of the two spectral passes.
//
// Mixed precision: the state arrays are FP32 fluctuations about per-plane
// FP64 backgrounds (t0, r0; B about its means bxm, bym, bzm) and the kernel
// rebuilds full FP64 values on load; all arithmetic is FP64.
struct HydroColArgs {
    Geom g;
    Phys p;
    const void *u, *v, *w, *t, *r;      // state: FP64, or FP32 fluctuations (mixed)
    const double *t0, *r0;              // mixed: backgrounds of t, r (nz entries)
    const void *bx, *by, *bz, *jsq;     // working precision
    const double *bxm, *bym;            // mixed: Bx-bar(z), By-bar(z)
    double bzm;                         // mixed: Bz-bar
    const double *kappa, *dkappa;
    const double* etar;                 // eta(z)/eta_0 for the Ohmic heating, or null (1)
    // x pass: d/dx of r, t, u, v, w; d2/dx2 of t, v, w; uxxc = sck*uxx +
    // scko3*d/dx(d/dx u); xbb = d/dx(F by bx); xzx = d/dx(F bz bx);
    // dxm = d/dx(F/2 (bx^2 - by^2 - bz^2))
    const void *rx, *tx, *txx, *ux, *uxxc, *vx, *vxx, *wx, *wxx, *xbb, *xzx, *dxm;
    // y pass: the same along y with vyyc = sck*vyy + scko3*d/dy(d/dy v),
    // dym = d/dy(F/2 (by^2 - bx^2 - bz^2)), uxy = d/dy(ux), vxy = d/dy(vx)
    const void *ry, *ty, *tyy, *uy, *uyy, *vy, *vyyc, *wy, *wyy, *ybb, *yzy, *dym, *uxy, *vxy;
    void *rdot, *tdot, *udot, *vdot, *wdot;   // newest history level (working precision)
    int zc;                             // planes per thread
    // fused AB3 update of u v w r t (src/main.F90:413-443); ab3 = 0: RHS only.
    // The new state goes to separate buffers, so the column walks of other
    // chunks and the magnetic pass still read the old one.
    int ab3;
    double c0, c1, c2;
    const void* h1[5];                  // history levels 1 and 2 (u v w r t order)
    const void* h2[5];
    void* h3z[3];                       // rdot levels 1 2 3, zeroed by the density floor
    const double* x64[5];               // mixed: the FP64 state (the windows hold FP32 shadows)
    double* xnew[5];                    // new FP64 state
    float* snew[5];                     // mixed: new FP32 shadows
};
int launch_hydro_col(const HydroColArgs& a, int mixed, hipStream_t s);

// ---- magnetic part (src/magnetic.f90) --------------------------------------
struct MagArgs {
    Geom g;
    Phys p;
    int icutx, icuty;
    const double *akx1, *akx2, *aky1, *aky2;
    double *pol, *tor, *bxbar, *bybar;
    void *poldot, *tordot;        // newest history level (working precision)
    double *bxbardot, *bybardot;  // newest history level of the means (FP64)
    void *exh, *eyh, *ezh;        // spectral EMF / B (double-halfcomplex)
    void *d, *cc;                 // mag_deriv work (mag_rhs takes D_z of d itself)
    void *polz;                   // bandj: d pol / dz
    void *bxh, *byh, *bzh, *jzh, *dybzh, *dxbzh;  // bandj spectral outputs
    const void *bx, *by;          // physical B for J^2
    const double *bxm, *bym;      // mixed: means of bx, by (their d/dz enters J^2)
    void *jsq;
    void *jx, *jy, *jz;           // eta(z): J kept for the EMF, or null
    int wf32;
    int zc;                       // planes per thread of the column-walking kernels
    // mag_rhs with the AB3 update of pol, tor fused (src/main.F90:413-443) and
    // magdealias; the new values go to polnew, tornew (other threads' z-windows
    // still read the old ones); ab3 = 0: right-hand sides only
    int ab3;
    double c0, c1, c2;
    const void *ph1, *ph2, *th1, *th2;   // history levels 1 and 2 of pol, tor
    double *polnew, *tornew;
};
int launch_mag_spec1(const MagArgs& a, hipStream_t s);   // D, C from E-hat
int launch_mean_emf(const MagArgs& a, hipStream_t s);    // mean-field RHS
int launch_mag_rhs(const MagArgs& a, hipStream_t s);     // poldot, tordot
int launch_bandj_spec(const MagArgs& a, hipStream_t s);  // B-hat, J-hat
int launch_jsq(const MagArgs& a, hipStream_t s);         // J^2
int launch_mag_init_spec(const MagArgs& a, hipStream_t s, double* means_out /*device: bzbar*/);

// ---- AB3 update, density floor, history ring, magdealias -------------------
struct AB3Args {
    Geom g;
    double c0, c1, c2;
    int f0;                   // first field updated (5: pol, tor; 7: the means only)
    double* var[7];           // u v w r t pol tor
    void* hist[7][3];         // [field][level 1..3] physical slots
    int h32;                  // history arrays are float
    double* mvar[2];          // bxbar bybar
    double* mhist[2][3];
    int nfields;              // 5 (hydro only) or 7
    int icutx, icuty;
    int mag;
    // optional working-precision shadows written after the update
    void* shadow[5];          // u v w r t (float), or null
    const double* bg_t;       // per-plane background subtracted for the T shadow
    const double* bg_r;       // per-plane background subtracted for the rho shadow
};
int launch_ab3(const AB3Args& a, hipStream_t s);

// ---- boundary conditions (src/dbydz.F90:3686-4791, defaults) ---------------
struct BCArgs {
    Geom g;
    double theta;
    double *u, *v, *w, *t, *pol, *tor, *bxbar, *bybar;
    float *su, *sv, *sw, *st;  // mixed: FP32 shadows of the rows set here, or null
    const double* t0;          // mixed: background subtracted for st
    int hydro;   // boundary()
    int mag;     // mag_boundary()
    int radbc, constflux, perfect;   // switches (MHD_OPT_*)
    double dz;
};
int launch_boundary(const BCArgs& a, hipStream_t s);
int launch_magdealias(Geom g, double* pol, double* tor, int icutx, int icuty, hipStream_t s);

// ---- reductions (src/utils.F90, src/dbydz.F90:4796-4925) --------------------
enum RedOp { RED_GLOBALAVG = 0, RED_MINMAX = 1, RED_DIVB = 2, RED_CALCDT = 3, RED_NUSSELT = 4 };
struct RedArgs {
    Geom g;
    Phys p;
    int op;
    const double *u, *v, *w, *r, *t;
    const void *bx, *by, *bz;
    const void *f0, *f1, *f2;   // generic inputs (minmax: f0; divb: f0+f1+f2)
    int f_f32;
    const double* f0m;          // per-plane mean added to f0 (minmax), or null
    double f0s;                 // scalar added to f0 (minmax)
    int wf32;                   // bx/by/bz are float
    const double *bxm, *bym;    // mixed: means of bx, by (null: B holds them)
    double bzm;                 // mixed: mean of bz
    double* partial;            // nz * (RED_CHUNKS + 1) * 8
    double* result;             // 8 doubles
};
// workgroups per plane of a reduction: nz * RED_CHUNKS workgroups fill the
// device (nz alone, 160 at 256x256x160, leaves CUs idle); the finishing
// kernel combines the chunks of a plane in order, then the planes
constexpr int RED_CHUNKS = 8;
int launch_reduce(const RedArgs& a, hipStream_t s);
// subtract result[0], result[1] from u, v (GlobalAverages side effect)
int launch_com_shift(Geom g, double* u, double* v, const double* res, hipStream_t s);

// ---- mixed precision: FP32 fluctuation shadows of the FP64 state ------------
// su..sw = u..w, st = t - t0(z), sr = r - r0(z), rounded to float
int launch_shadow(Geom g, const double* u, const double* v, const double* w, const double* t, const double* r,
                  float* su, float* sv, float* sw, float* st, float* sr, const double* t0, const double* r0,
                  hipStream_t s);
// dst[k-plane] = (double)src + mean[k] + s (mean may be null)
int launch_add_mean(const float* src, double* dst, const double* mean, double sc, Geom g, hipStream_t s);

// ---- misc -------------------------------------------------------------------
int launch_copy_cvt(const void* src, int src_f32, void* dst, int dst_f32, long n, hipStream_t s);
int launch_fill(void* dst, int f32, double v, long n, hipStream_t s);

}  // namespace mhd
