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

// ---- hydro right-hand side assembly (src/hydro_deriv.f90) -------------------
struct HydroArgs {
    Geom g;
    Phys p;
    const double *u, *v, *w, *t, *r;
    const void *bx, *by, *bz, *jsq;     // working precision
    const double *kappa, *dkappa;
    // x / y / z derivative planes (working precision)
    const void *rx, *tx, *txx, *ux, *uxx, *vx, *vxx, *wx, *wxx, *xbb, *xzx;
    const void *ry, *ty, *tyy, *uy, *uyy, *vy, *vyy, *wy, *wyy, *ybb, *yzy;
    const void *tzu, *tzc, *tzz, *wzu, *wzc, *uzu, *uzz, *vzu, *vzz, *zzx, *zzy, *rzu, *rzc;
    const void *dxc1, *dyc2, *dzc3, *dzc4;
    double *rdot, *tdot, *udot, *vdot, *wdot;
    void *c1, *c2, *c3, *c4;
    int wf32;   // working arrays are float
};
int launch_hydro_a(const HydroArgs& a, hipStream_t s);
int launch_hydro_b(const HydroArgs& a, hipStream_t s);

// ---- magnetic part (src/magnetic.f90) --------------------------------------
struct MagArgs {
    Geom g;
    Phys p;
    int icutx, icuty;
    const double *akx1, *akx2, *aky1, *aky2;
    double *pol, *tor, *bxbar, *bybar;
    double *poldot, *tordot, *bxbardot, *bybardot;   // newest history level
    void *exh, *eyh, *ezh;        // spectral EMF / B (double-halfcomplex)
    void *d, *cc, *dz;            // mag_deriv work
    void *polz;                   // bandj: d pol / dz
    void *bxh, *byh, *bzh, *jzh, *dybzh, *dxbzh;  // bandj spectral outputs
    const void *bx, *by;          // physical B for J^2
    void *jsq;
    int wf32;
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
    double* var[7];           // u v w r t pol tor
    double* hist[7][3];       // [field][level 1..3] physical slots
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
    int hydro;   // boundary()
    int mag;     // mag_boundary()
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
    int wf32;                   // bx/by/bz are float
    double* partial;            // nz * 8
    double* result;             // 8 doubles
};
int launch_reduce(const RedArgs& a, hipStream_t s);
// subtract result[0], result[1] from u, v (GlobalAverages side effect)
int launch_com_shift(Geom g, double* u, double* v, const double* res, hipStream_t s);

// ---- misc -------------------------------------------------------------------
int launch_copy_cvt(const void* src, int src_f32, void* dst, int dst_f32, long n, hipStream_t s);
int launch_fill(void* dst, int f32, double v, long n, hipStream_t s);

}  // namespace mhd
