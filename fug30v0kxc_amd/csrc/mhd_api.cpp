// mhd_api.cpp -- host side of the C ABI (include/mhd.h): context, device
// memory, and the time-loop driver that sequences the gfx950 kernels in the
// order of the reference's src/main.F90:340-683.
//
// Device layout (one hipMalloc pool per context, every array 256-B aligned):
//   state (FP64):   u v w t r pol tor, bxbar bybar (nz)
//   B and J^2:      bx by bz jsq (working precision)
//   RHS history:    7 fields x 3 levels (FP64), ring-indexed so the
//                   reference's level shift (2->1, 3->2) is a relabelling
//   work:           derivative planes of the RHS assembly (working precision)
//   tables:         akx1 akx2 aky1 aky2, kappa dkappa, FFT twiddles
#include "../../include/mhd.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <array>
#include <utility>
#include <vector>

#include "closure.hpp"
#include "launch.hpp"
#include "physics.hpp"
#include "spectral.hpp"
#include "zstencil.hpp"

using namespace mhd;

namespace mhd {
LaunchRecorder*& launch_recorder()
{
    static thread_local LaunchRecorder* r = nullptr;
    return r;
}
}  // namespace mhd

int mhd_hip_error(hipError_t e, const char* what, const char* file, int line)
{
    fprintf(stderr, "mhd: HIP error %d (%s) at %s:%d: %s\n", (int)e, hipGetErrorString(e), file, line, what);
    return MHD_E_DEVICE;
}

namespace {

// work buffers of the RHS assembly
enum Wk {
    W_RX, W_TX, W_TXX, W_UX, W_UXX, W_VX, W_VXX, W_WX, W_WXX, W_XBB, W_XZX,
    W_RY, W_TY, W_TYY, W_UY, W_UYY, W_VY, W_VYY, W_WY, W_WYY, W_YBB, W_YZY,
    W_DXM, W_DYM, W_UXY, W_VXY,
    W_COUNT
};

constexpr int MAXEV = 8192;

}  // namespace

struct mhd_ctx {
    int nx, ny, nz, nxy;
    long n3;
    int prec, wf32;
    size_t wsz;  // bytes per working element
    mhd_params_t p;
    Phys ph;
    Geom g;
    int icutx, icuty;
    double dx, dy, dz, dd;
    hipStream_t s;

    char* pool;
    size_t pool_bytes;
    double *akx1, *akx2, *aky1, *aky2, *kappa, *dkappa;
    void *twx, *twy;                 // twiddles in the working precision
    void *twx64, *twy64;             // FP64 twiddles (FP64 transforms of the mixed path)
    double *u, *v, *w, *t, *r, *pol, *tor, *bxbar, *bybar;
    void *bx, *by, *bz, *jsq;
    void* hist[7][3];  // RHS history (FP32 in mixed precision)
    int h32;
    double* mhist[2][3];
    float *su, *sv, *sw, *st, *sr;   // mixed: FP32 fluctuation shadows of u v w t r
    // the other half of the state ping-pong: the fused hydro + AB3 kernel writes
    // the new u v w t r (and shadows) here, then the pointers swap
    double *u2, *v2, *w2, *t2, *r2;
    float *su2, *sv2, *sw2, *st2, *sr2;
    double *pol2, *tor2;             // and of pol, tor (the update fused into mag_rhs)
    double *t0, *r0, *bzm;           // mixed: backgrounds of t, r; Bz-bar per plane
    // learned subgrid closure: the network, its 12 fluxes, their x/y/z
    // derivatives (working precision) and a vector of -1 (eps enters the EMF
    // through the load's extra term)
    mhd_closure* closure;
    void* cl_mem;
    void *cflux[12], *cdx[4], *cdy[4], *cdz[4], *cgrid;
    double* cneg1;
    int eta_on;                      // eta(z) extension active
    double *etar, *eta1;             // eta(z)/eta_0 and zeta C_k (etar - 1), nz entries
    void *jx, *jy, *jz;              // eta(z): current density kept for the EMF
    int ring[3];
    void* wk[W_COUNT];
    double *red_partial, *red_result, *d_bzbar;
    double* stage;     // device FP64 staging (one field)
    int ncu;           // compute units (persistent transform grids)
    int spec_var[2][4]; // radix plan per direction and transform kind (SpecMode), tuned
    int spec_mult[2][4];  // persistent-grid multiple per direction and SpecMode, tuned
    int hzc;           // planes per thread of the column-walking hydro kernel
    int mzc;           // ... of the magnetic column kernels (mag_rhs, jsq)
    double* hbuf;      // pinned host staging (one field)

    mhd_scalars_t sc;
    int have_state;

    int timed_kernel;
    int nev;
    hipEvent_t ev[2 * MAXEV];
    int ev_created;
};

namespace {

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// history field index (u v w r t pol tor) for MHD_F_* ids
int hist_index(int id)
{
    switch (id) {
    case MHD_F_UDOT: return 0;
    case MHD_F_VDOT: return 1;
    case MHD_F_WDOT: return 2;
    case MHD_F_RDOT: return 3;
    case MHD_F_TDOT: return 4;
    case MHD_F_POLDOT: return 5;
    case MHD_F_TORDOT: return 6;
    }
    return -1;
}

void* level_ptr(mhd_ctx* c, int f, int level) { return c->hist[f][c->ring[level - 1]]; }
double* mlevel_ptr(mhd_ctx* c, int f, int level) { return c->mhist[f][c->ring[level - 1]]; }

void set_constants(mhd_ctx* c)
{
    // src/constants.f90:56-109, src/dbydz.F90:355-362
    const mhd_params_t& p = c->p;
    Phys& h = c->ph;
    c->dx = p.xmax / (double)c->nx;
    c->dy = p.ymax / (double)c->ny;
    c->dz = 1.0 / (double)(c->nz - 1);
    c->dd = std::fmin(c->dx, c->dy);
    const double taylt = p.tayl / std::pow(1.0 + p.theta / 2.0, 2.0 * p.cm);
    h.cpsi = std::cos(p.psi);
    h.spsi = std::sin(p.psi);
    h.cor = p.ck * p.sigma * std::sqrt(taylt);
    h.gck = p.gamma * p.ck;
    h.sck = p.sigma * p.ck;
    h.zck = p.zeta * p.ck;
    h.scko3 = h.sck / 3.0;
    h.sckg = p.sigma * p.ck * (p.gamma - 1.0);
    h.fzckg = p.f * p.zeta * p.ck * (p.gamma - 1.0);
    h.thmplus1 = p.theta * (p.cm + 1.0);
    h.fo2 = p.f / 2.0;
    h.g1 = std::sqrt(p.gamma);
    h.gm1 = p.gamma - 1.0;
    h.nusstmp = p.theta * (p.cm + 1.0) * (1.0 - (1.0 / p.gamma));
    h.f = p.f;
    h.theta = p.theta;
    h.xn = 1.0 / (double)(c->nx * c->ny);
    h.xzn = c->dx * c->dy * c->dz;
    h.dz = c->dz;
    h.dd = c->dd;
    h.sf = p.sf;
    h.dtvisc0x = 0.1 * 0.75 * c->dd * c->dd / p.sigma / p.ck;
    h.dtvisc0z = 0.2 * c->dz * c->dz / p.sigma / p.ck * (1.0 + p.theta / 7.0);
    h.dtdiff0x = 0.1 * c->dd * c->dd / p.gamma / p.ck;
    h.dtdiff0z = 0.2 * c->dz * c->dz / p.gamma / p.ck * (1.0 + p.theta / 7.0);
    h.ho12 = (double)(c->nz - 1) / 12.0;
    h.sc1 = (double)((c->nz - 1) * (c->nz - 1)) / 12.0;
    // the field evolves when F != 0 or for the kinematic problem (src/main.F90:355-375)
    h.mag = (p.f != 0.0) || (p.opts & MHD_OPT_KINEMATIC);
}

// src/fft.F90:171-221
void wavenumbers(int n, double len, std::vector<double>& a1, std::vector<double>& a2)
{
    a1.assign(n, 0.0);
    a2.assign(n, 0.0);
    const double cx = 8.0 * std::atan(1.0) / len;
    for (int i = 1; i <= n / 2 - 1; ++i) {
        a1[i] = (double)i * cx;
        a1[n - i] = (double)i * cx;
        a2[i] = -((double)i * (double)i * cx * cx);
        a2[n - i] = -((double)i * (double)i * cx * cx);
    }
    a2[n / 2] = -((double)(n / 2) * (double)(n / 2) * cx * cx);
}

// W_N^t = exp(-2 pi i t / N), rounded once from long double
void twiddles(int n, std::vector<double>& d, std::vector<float>& f)
{
    d.resize(2 * n);
    f.resize(2 * n);
    for (int t = 0; t < n; ++t) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
        d[2 * t] = (double)cosl(a);
        d[2 * t + 1] = (double)sinl(a);
        f[2 * t] = (float)cosl(a);
        f[2 * t + 1] = (float)sinl(a);
    }
}

// src/trans_params.f90:36-62
void kappa_profile(int nz, double dz, std::vector<double>& ka, std::vector<double>& dka)
{
    ka.resize(nz);
    dka.resize(nz);
    const double constA = 10.0;
    for (int k = 0; k < nz; ++k) {
        const double z = (double)k * dz;
        const double constB = constA * (z - 5.0e-1);
        ka[k] = 5.0e-1 * (std::tanh(constB) + 1);
        dka[k] = 5.0e-1 * constA * std::pow(std::cosh(constB), -2.0);
    }
}

// Park-Miller minimal standard with Schrage's factorisation and the XOR mask,
// src/ranset.f90:16-29
double ran0(int* idum)
{
    const int IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836, MASK = 123459876;
    const double AM = 1.0 / IM;
    *idum ^= MASK;
    const int k = *idum / IQ;
    *idum = IA * (*idum - k * IQ) - IR * k;
    if (*idum < 0) *idum += IM;
    const double r = AM * (*idum);
    *idum ^= MASK;
    return r;
}

int kerr(int e) { return e ? MHD_E_DEVICE : 0; }

// f64: FP64 transform in the mixed path (launched with spec64)
SpecLaunch spec_base(mhd_ctx* c, int dir, int mode, bool f64 = false)
{
    SpecLaunch L;
    memset(&L, 0, sizeof(L));
    L.g = c->g;
    L.dir = dir;
    L.mode = mode;
    L.cx = 8.0 * std::atan(1.0) / (dir == 0 ? c->p.xmax : c->p.ymax);   // src/fft.F90:184-185
    if (f64) L.tw = dir == 0 ? c->twx64 : c->twy64;
    else L.tw = dir == 0 ? c->twx : c->twy;
    L.icut = dir == 0 ? c->icutx : c->icuty;
    L.ncu = c->ncu;
    L.variant = c->spec_var[dir][mode];
    L.grid_mult = c->spec_mult[dir][mode];
    return L;
}

static_assert(MHD_TK_DERIV1 == SM_DERIV1 && MHD_TK_DERIV2 == SM_DERIV2 && MHD_TK_R2HC == SM_R2HC &&
                  MHD_TK_HC2R == SM_HC2R,
              "the tuner's transform kinds index the context's plans by SpecMode");

SpecJob sjob(const void* in, int in_f32, void* o0, void* o1, int m0, int m1, int out_f32)
{
    SpecJob j;
    memset(&j, 0, sizeof(j));
    j.in[0] = in;
    j.in_f32 = in_f32;
    j.load_op = LD_PLAIN;
    j.scale = 1.0;
    j.out[0] = o0;
    j.out[1] = o1;
    j.mul[0] = m0;
    j.mul[1] = m1;
    j.out_f32 = out_f32;
    return j;
}

ZJob zjob(const void* in, int in_f32, void* out, int out_f32, int kind, int flag)
{
    ZJob j;
    memset(&j, 0, sizeof(j));
    j.in = in;
    j.in_f32 = in_f32;
    j.load_op = 0;
    j.scale = 1.0;
    j.out = out;
    j.out_f32 = out_f32;
    j.kind = kind;
    j.flag = flag;
    return j;
}

ZLaunch zbase(mhd_ctx* c)
{
    ZLaunch L;
    memset(&L, 0, sizeof(L));
    L.g = c->g;
    L.ho12 = c->ph.ho12;
    L.sc1 = c->ph.sc1;
    L.w = c->w;
    L.w_f32 = 0;
    return L;
}

// timed launch wrapper
template <typename F>
int timed(mhd_ctx* c, int kernel, F&& launch)
{
    const bool on = (c->timed_kernel == kernel) && c->nev + 2 <= 2 * MAXEV;
    if (on && hipEventRecord(c->ev[c->nev], c->s) != hipSuccess) return MHD_E_DEVICE;
    int e = launch();
    if (on) {
        if (hipEventRecord(c->ev[c->nev + 1], c->s) != hipSuccess) return MHD_E_DEVICE;
        c->nev += 2;
    }
    return e;
}

int spec(mhd_ctx* c, const SpecLaunch& L, int kernel = MHD_K_NONE)
{
    const int cf32 = c->wf32;
    int e = timed(c, kernel, [&] { return launch_spectral(L, cf32, c->s); });
    return e ? (e == -20 ? MHD_E_SIZE : MHD_E_DEVICE) : 0;
}

// FP64 transform whatever the working precision (L from spec_base(.., true))
int spec64(mhd_ctx* c, const SpecLaunch& L)
{
    int e = launch_spectral(L, 0, c->s);
    return e ? (e == -20 ? MHD_E_SIZE : MHD_E_DEVICE) : 0;
}

// refresh the FP32 shadows after the FP64 state changed outside the AB3 update
int refresh_shadows(mhd_ctx* c)
{
    if (!c->su) return 0;
    return kerr(launch_shadow(c->g, c->u, c->v, c->w, c->t, c->r, c->su, c->sv, c->sw, c->st, c->sr, c->t0, c->r0,
                              c->s));
}

int zrun(mhd_ctx* c, const ZLaunch& L)
{
    int e = timed(c, MHD_K_ZDERIV, [&] { return launch_z(L, c->s); });
    return e == 7 ? MHD_E_FLAG : (e ? MHD_E_DEVICE : 0);
}

HydroColArgs colargs(mhd_ctx* c)
{
    HydroColArgs a;
    memset(&a, 0, sizeof(a));
    a.g = c->g;
    a.p = c->ph;
    if (c->su) {
        a.u = c->su; a.v = c->sv; a.w = c->sw; a.t = c->st; a.r = c->sr;
        a.t0 = c->t0; a.r0 = c->r0;
        a.bxm = c->bxbar; a.bym = c->bybar; a.bzm = c->sc.bzbar;
    } else {
        a.u = c->u; a.v = c->v; a.w = c->w; a.t = c->t; a.r = c->r;
    }
    a.bx = c->bx; a.by = c->by; a.bz = c->bz; a.jsq = c->jsq;
    a.kappa = c->kappa; a.dkappa = c->dkappa;
    a.etar = c->eta_on ? c->etar : nullptr;
    void** W = c->wk;
    a.rx = W[W_RX]; a.tx = W[W_TX]; a.txx = W[W_TXX]; a.ux = W[W_UX]; a.uxxc = W[W_UXX];
    a.vx = W[W_VX]; a.vxx = W[W_VXX]; a.wx = W[W_WX]; a.wxx = W[W_WXX]; a.xbb = W[W_XBB]; a.xzx = W[W_XZX];
    a.dxm = W[W_DXM];
    a.ry = W[W_RY]; a.ty = W[W_TY]; a.tyy = W[W_TYY]; a.uy = W[W_UY]; a.uyy = W[W_UYY];
    a.vy = W[W_VY]; a.vyyc = W[W_VYY]; a.wy = W[W_WY]; a.wyy = W[W_WYY]; a.ybb = W[W_YBB]; a.yzy = W[W_YZY];
    a.dym = W[W_DYM]; a.uxy = W[W_UXY]; a.vxy = W[W_VXY];
    a.rdot = level_ptr(c, 3, 3);
    a.tdot = level_ptr(c, 4, 3);
    a.udot = level_ptr(c, 0, 3);
    a.vdot = level_ptr(c, 1, 3);
    a.wdot = level_ptr(c, 2, 3);
    a.zc = c->hzc;
    return a;
}

MagArgs margs(mhd_ctx* c)
{
    MagArgs a;
    memset(&a, 0, sizeof(a));
    a.g = c->g;
    a.p = c->ph;
    a.icutx = c->icutx;
    a.icuty = c->icuty;
    a.akx1 = c->akx1; a.akx2 = c->akx2; a.aky1 = c->aky1; a.aky2 = c->aky2;
    a.pol = c->pol; a.tor = c->tor; a.bxbar = c->bxbar; a.bybar = c->bybar;
    a.poldot = level_ptr(c, 5, 3);
    a.tordot = level_ptr(c, 6, 3);
    a.bxbardot = mlevel_ptr(c, 0, 3);
    a.bybardot = mlevel_ptr(c, 1, 3);
    void** W = c->wk;
    // the RHS-assembly planes are free while the magnetic routines run
    a.exh = W[W_RX]; a.eyh = W[W_TX]; a.ezh = W[W_TXX];
    a.d = W[W_UX]; a.cc = W[W_UXX];
    a.polz = W[W_VXX];
    a.bxh = W[W_WX]; a.byh = W[W_WXX]; a.bzh = W[W_XBB]; a.jzh = W[W_XZX]; a.dybzh = W[W_RY]; a.dxbzh = W[W_TY];
    a.bx = c->bx; a.by = c->by; a.jsq = c->jsq;
    if (c->su) { a.bxm = c->bxbar; a.bym = c->bybar; }
    if (c->eta_on) { a.jx = c->jx; a.jy = c->jy; a.jz = c->jz; }
    a.wf32 = c->wf32;
    a.zc = c->mzc;
    return a;
}


#define TRY(x)                    \
    do {                          \
        int _r = (x);             \
        if (_r) return _r;        \
    } while (0)

// ---------------------------------------------------------------------------
// hydro_deriv, src/hydro_deriv.f90:5-505
// ---------------------------------------------------------------------------
// cf: AB3 coefficients to fuse the update of u v w r t into the hydro kernel
// (new state in the ping-pong buffers), or null for the right-hand sides only
int run_closure_flux(mhd_ctx* c);
int run_closure_tend(mhd_ctx* c);

// The AB3 update of u, v, w, rho, T fused into hydro_col: coefficients cf,
// history levels 1 and 2 of the ring, the new state into the ping-pong
// buffers; hz1, hz2 are the rho history levels the density floor clears.
void fuse_ab3(mhd_ctx* c, HydroColArgs& a, const double* cf, void* hz1, void* hz2)
{
    a.ab3 = 1;
    a.c0 = cf[0]; a.c1 = cf[1]; a.c2 = cf[2];
    const int hf[5] = {0, 1, 2, 3, 4};   // u v w r t
    for (int i = 0; i < 5; ++i) {
        a.h1[i] = level_ptr(c, hf[i], 1);
        a.h2[i] = level_ptr(c, hf[i], 2);
    }
    a.h3z[0] = hz1;
    a.h3z[1] = hz2;
    a.h3z[2] = level_ptr(c, 3, 3);
    const double* x64[5] = {c->u, c->v, c->w, c->r, c->t};
    double* xn[5] = {c->u2, c->v2, c->w2, c->r2, c->t2};
    float* sn[5] = {c->su2, c->sv2, c->sw2, c->sr2, c->st2};
    for (int i = 0; i < 5; ++i) {
        a.x64[i] = x64[i];
        a.xnew[i] = xn[i];
        a.snew[i] = sn[i];
    }
    if (c->su) { a.t0 = c->t0; a.r0 = c->r0; }
}

int run_hydro_deriv(mhd_ctx* c, const double* cf = nullptr)
{
    if (c->closure) {
        if (cf) return MHD_E_STATE;   // the closure's terms need the unfused update
        TRY(run_closure_flux(c));
    }
    const int wf = c->wf32;
    void** W = c->wk;
    const Phys& P = c->ph;
    const double f = c->p.f;
    // mixed precision: the transforms read the FP32 fluctuation shadows, and
    // the products the per-plane means of B (B is stored without them)
    const bool mx = c->su != nullptr;
    const void* T_ = mx ? (const void*)c->st : c->t;
    const void* U_ = mx ? (const void*)c->su : c->u;
    const void* V_ = mx ? (const void*)c->sv : c->v;
    const void* W_ = mx ? (const void*)c->sw : c->w;
    const void* R_ = mx ? (const void*)c->sr : c->r;
    const int sf = mx ? 1 : 0;
    const double* mbx = mx ? c->bxbar : nullptr;
    const double* mby = mx ? c->bybar : nullptr;
    const double* mbz = mx ? c->bzm : nullptr;
    auto means = [](SpecJob& j, const double* m0, const double* m1, const double* m2) {
        j.mean[0] = m0; j.mean[1] = m1; j.mean[2] = m2;
    };
    // x pass: d/dx, d2/dx2 of t, u, v, w (u's second output also carries the
    // scko3 d/dx d/dx u of the composite flux, lines 181, 258) and d/dx of r,
    // a single-output job of the same launch (one launch instead of two: at
    // 256^2 x 160 FP64 242 against 210 + 47 us); then d/dx of the
    // Maxwell-stress products F by bx, F bz bx, F/2 (bx^2-by^2-bz^2)
    {
        SpecLaunch L = spec_base(c, 0, SM_DERIV2);
        L.job[0] = sjob(T_, sf, W[W_TX], W[W_TXX], MUL_D1, MUL_D2, wf);
        L.job[1] = sjob(U_, sf, W[W_UX], W[W_UXX], MUL_D1, MUL_D2C, wf);
        L.job[1].mc[0] = P.sck; L.job[1].mc[1] = P.scko3;
        L.job[2] = sjob(V_, sf, W[W_VX], W[W_VXX], MUL_D1, MUL_D2, wf);
        L.job[3] = sjob(W_, sf, W[W_WX], W[W_WXX], MUL_D1, MUL_D2, wf);
        L.job[4] = sjob(R_, sf, W[W_RX], nullptr, MUL_D1, 0, wf);
        L.njobs = 5;
        TRY(spec(c, L, MHD_K_XDERIV2));
        SpecLaunch M = spec_base(c, 0, SM_DERIV1);
        M.job[0] = sjob(c->by, wf, W[W_XBB], nullptr, MUL_D1, 0, wf);
        M.job[0].load_op = LD_PROD; M.job[0].in[1] = c->bx; M.job[0].scale = f;
        means(M.job[0], mby, mbx, nullptr);
        M.job[1] = sjob(c->bz, wf, W[W_XZX], nullptr, MUL_D1, 0, wf);
        M.job[1].load_op = LD_PROD; M.job[1].in[1] = c->bx; M.job[1].scale = f;
        means(M.job[1], mbz, mbx, nullptr);
        M.job[2] = sjob(c->bx, wf, W[W_DXM], nullptr, MUL_D1, 0, wf);
        M.job[2].load_op = LD_MAXW; M.job[2].in[1] = c->by; M.job[2].in[2] = c->bz; M.job[2].scale = P.fo2;
        means(M.job[2], mbx, mby, mbz);
        M.njobs = 3;
        TRY(spec(c, M));
    }
    // y pass: the same along y (v's second output carries scko3 d/dy d/dy v),
    // with the single-output jobs d/dy of r and the mixed derivatives d/dy(ux),
    // d/dy(vx) of the composite fluxes (FP64 318 against 219 + 120 us as two
    // launches); then the Maxwell-stress products
    {
        SpecLaunch L = spec_base(c, 1, SM_DERIV2);
        L.job[0] = sjob(T_, sf, W[W_TY], W[W_TYY], MUL_D1, MUL_D2, wf);
        L.job[1] = sjob(U_, sf, W[W_UY], W[W_UYY], MUL_D1, MUL_D2, wf);
        L.job[2] = sjob(V_, sf, W[W_VY], W[W_VYY], MUL_D1, MUL_D2C, wf);
        L.job[2].mc[0] = P.sck; L.job[2].mc[1] = P.scko3;
        L.job[3] = sjob(W_, sf, W[W_WY], W[W_WYY], MUL_D1, MUL_D2, wf);
        L.job[4] = sjob(R_, sf, W[W_RY], nullptr, MUL_D1, 0, wf);
        L.job[5] = sjob(W[W_UX], wf, W[W_UXY], nullptr, MUL_D1, 0, wf);
        L.job[6] = sjob(W[W_VX], wf, W[W_VXY], nullptr, MUL_D1, 0, wf);
        L.njobs = 7;
        TRY(spec(c, L, MHD_K_YDERIV2));
        SpecLaunch M = spec_base(c, 1, SM_DERIV1);
        M.job[0] = sjob(c->by, wf, W[W_YBB], nullptr, MUL_D1, 0, wf);
        M.job[0].load_op = LD_PROD; M.job[0].in[1] = c->bx; M.job[0].scale = f;
        means(M.job[0], mby, mbx, nullptr);
        M.job[1] = sjob(c->bz, wf, W[W_YZY], nullptr, MUL_D1, 0, wf);
        M.job[1].load_op = LD_PROD; M.job[1].in[1] = c->by; M.job[1].scale = f;
        means(M.job[1], mbz, mby, nullptr);
        M.job[2] = sjob(c->by, wf, W[W_DYM], nullptr, MUL_D1, 0, wf);
        M.job[2].load_op = LD_MAXW; M.job[2].in[1] = c->bx; M.job[2].in[2] = c->bz; M.job[2].scale = P.fo2;
        means(M.job[2], mby, mbx, mbz);
        M.njobs = 3;
        TRY(spec(c, M));
    }
    HydroColArgs a = colargs(c);
    if (cf) fuse_ab3(c, a, cf, level_ptr(c, 3, 1), level_ptr(c, 3, 2));
    TRY(timed(c, MHD_K_HYDRO, [&] { return kerr(launch_hydro_col(a, c->wf32, c->s)); }));
    return c->closure ? run_closure_tend(c) : 0;
}

// the closure's fluxes of the current state (B with its means in the mixed path)
int run_closure_flux(mhd_ctx* c)
{
    ClSource src;
    memset(&src, 0, sizeof(src));
    const double* f64[5] = {c->r, c->t, c->u, c->v, c->w};
    for (int q = 0; q < 5; ++q) src.p[q] = f64[q];
    src.p[5] = c->bx;
    src.p[6] = c->by;
    src.p[7] = c->bz;
    for (int q = 5; q < 8; ++q) src.f32[q] = c->wf32;
    if (c->su) {
        src.mean[5] = c->bxbar;
        src.mean[6] = c->bybar;
        src.add[7] = c->sc.bzbar;
    }
    const int e = closure_fluxes(c->closure, src, c->nx, c->ny, c->nz, c->cflux, c->wf32, c->cgrid, c->s);
    return e ? (e == MHD_E_ALLOC ? e : MHD_E_DEVICE) : 0;
}

// -div(tau)/r into udot, vdot, wdot and -div(q) into tdot (newest level)
int run_closure_tend(mhd_ctx* c)
{
    const int wf = c->wf32;
    // x: tau_xx, tau_xy, tau_xz, q_x; y: tau_xy, tau_yy, tau_yz, q_y; z: tau_xz, tau_yz, tau_zz, q_z
    static const int ix[4] = {0, 3, 4, 9}, iy[4] = {3, 1, 5, 10}, iz[4] = {4, 5, 2, 11};
    for (int dir = 0; dir < 2; ++dir) {
        SpecLaunch L = spec_base(c, dir, SM_DERIV1);
        for (int m = 0; m < 4; ++m)
            L.job[m] = sjob(c->cflux[dir == 0 ? ix[m] : iy[m]], wf, dir == 0 ? c->cdx[m] : c->cdy[m], nullptr,
                            MUL_D1, 0, wf);
        L.njobs = 4;
        TRY(spec(c, L));
    }
    ZLaunch Z = zbase(c);
    for (int m = 0; m < 4; ++m) Z.job[m] = zjob(c->cflux[iz[m]], wf, c->cdz[m], wf, ZK_D1, 5);
    Z.njobs = 4;
    TRY(zrun(c, Z));
    void* dots[4] = {level_ptr(c, 0, 3), level_ptr(c, 1, 3), level_ptr(c, 2, 3), level_ptr(c, 4, 3)};
    return kerr(closure_tend(c->n3, c->cdx, c->cdy, c->cdz, wf, c->r, dots, c->h32, c->s));
}

// ---------------------------------------------------------------------------
// mag_deriv, src/magnetic.f90:285-487
// ---------------------------------------------------------------------------
// cf: AB3 coefficients to fuse the update of pol, tor into mag_rhs (new
// values in the ping-pong buffers), or null for the right-hand sides only
int run_mag_deriv(mhd_ctx* c, const double* cf = nullptr)
{
    const int wf = c->wf32;
    MagArgs m = margs(c);
    {
        SpecLaunch L = spec_base(c, 0, SM_R2HC);
        // E = u x B (308-316), fused into the x pass of forwardfft
        // mixed precision: FP32 velocity shadows, B without its means (added by the loads)
        const bool mx = c->su != nullptr;
        const void* U_ = mx ? (const void*)c->su : c->u;
        const void* V_ = mx ? (const void*)c->sv : c->v;
        const void* W_ = mx ? (const void*)c->sw : c->w;
        const double* mbx = mx ? c->bxbar : nullptr;
        const double* mby = mx ? c->bybar : nullptr;
        const double* mbz = mx ? c->bzm : nullptr;
        L.job[0] = sjob(V_, wf, m.exh, nullptr, 0, 0, wf);
        L.job[0].load_op = LD_CROSS; L.job[0].in[1] = c->bz; L.job[0].in[2] = W_; L.job[0].in[3] = c->by;
        L.job[0].mean[1] = mbz; L.job[0].mean[3] = mby;
        L.job[1] = sjob(W_, wf, m.eyh, nullptr, 0, 0, wf);
        L.job[1].load_op = LD_CROSS; L.job[1].in[1] = c->bx; L.job[1].in[2] = U_; L.job[1].in[3] = c->bz;
        L.job[1].mean[1] = mbx; L.job[1].mean[3] = mbz;
        L.job[2] = sjob(U_, wf, m.ezh, nullptr, 0, 0, wf);
        L.job[2].load_op = LD_CROSS; L.job[2].in[1] = c->by; L.job[2].in[2] = V_; L.job[2].in[3] = c->bx;
        L.job[2].mean[1] = mby; L.job[2].mean[3] = mbx;
        if (c->eta_on) {
            // eta(z) extension: E = u x B - zeta C_k (eta/eta_0 - 1) J
            L.job[0].jin = c->jx; L.job[0].jcoef = c->eta1;
            L.job[1].jin = c->jy; L.job[1].jcoef = c->eta1;
            L.job[2].jin = c->jz; L.job[2].jcoef = c->eta1;
        } else if (c->closure) {
            // closure: E = u x B + eps
            for (int q = 0; q < 3; ++q) {
                L.job[q].jin = c->cflux[6 + q];
                L.job[q].jcoef = c->cneg1;
            }
        }
        L.njobs = 3;
        TRY(spec(c, L));
        SpecLaunch M = spec_base(c, 1, SM_R2HC);
        M.job[0] = sjob(m.exh, wf, m.exh, nullptr, 0, 0, wf);
        M.job[1] = sjob(m.eyh, wf, m.eyh, nullptr, 0, 0, wf);
        M.job[2] = sjob(m.ezh, wf, m.ezh, nullptr, 0, 0, wf);
        M.njobs = 3;
        TRY(spec(c, M));
    }
    TRY(kerr(launch_mean_emf(m, c->s)));
    TRY(kerr(launch_mag_spec1(m, c->s)));
    // D_z of d (line 419) is taken inside mag_rhs from a window of d
    if (cf) {
        m.ab3 = 1;
        m.c0 = cf[0]; m.c1 = cf[1]; m.c2 = cf[2];
        m.ph1 = level_ptr(c, 5, 1); m.ph2 = level_ptr(c, 5, 2);
        m.th1 = level_ptr(c, 6, 1); m.th2 = level_ptr(c, 6, 2);
        m.polnew = c->pol2;
        m.tornew = c->tor2;
    }
    return kerr(launch_mag_rhs(m, c->s));
}

// ---------------------------------------------------------------------------
// bandj, src/magnetic.f90:492-664 (jsq tail shared with mag_init)
// ---------------------------------------------------------------------------
int inverse_b_and_j(mhd_ctx* c, MagArgs& m, bool with_b)
{
    const int wf = c->wf32;
    SpecLaunch L = spec_base(c, 0, SM_HC2R);
    int n = 0;
    if (with_b) {
        L.job[n++] = sjob(m.bxh, wf, m.bxh, nullptr, 0, 0, wf);
        L.job[n++] = sjob(m.byh, wf, m.byh, nullptr, 0, 0, wf);
        L.job[n++] = sjob(m.bzh, wf, m.bzh, nullptr, 0, 0, wf);
    }
    L.job[n++] = sjob(m.dybzh, wf, m.dybzh, nullptr, 0, 0, wf);
    L.job[n++] = sjob(m.dxbzh, wf, m.dxbzh, nullptr, 0, 0, wf);
    L.job[n++] = sjob(m.jzh, wf, m.jzh, nullptr, 0, 0, wf);
    L.njobs = n;
    TRY(spec(c, L));
    SpecLaunch M = spec_base(c, 1, SM_HC2R);
    n = 0;
    if (with_b) {
        // 575-593: inverse transforms, then the mean parts of B
        // (the mixed path keeps B without its means: no additions)
        const bool mx = c->su != nullptr;
        M.job[n] = sjob(m.bxh, wf, c->bx, nullptr, 0, 0, wf);
        M.job[n++].addk = mx ? nullptr : c->bxbar;
        M.job[n] = sjob(m.byh, wf, c->by, nullptr, 0, 0, wf);
        M.job[n++].addk = mx ? nullptr : c->bybar;
        M.job[n] = sjob(m.bzh, wf, c->bz, nullptr, 0, 0, wf);
        M.job[n++].adds = mx ? 0.0 : c->sc.bzbar;
    }
    M.job[n++] = sjob(m.dybzh, wf, m.dybzh, nullptr, 0, 0, wf);
    M.job[n++] = sjob(m.dxbzh, wf, m.dxbzh, nullptr, 0, 0, wf);
    M.job[n++] = sjob(m.jzh, wf, m.jzh, nullptr, 0, 0, wf);
    M.njobs = n;
    TRY(spec(c, M));
    return kerr(launch_jsq(m, c->s));
}

int run_bandj(mhd_ctx* c)
{
    const int wf = c->wf32;
    MagArgs m = margs(c);
    ZLaunch Z = zbase(c);
    Z.job[0] = zjob(c->pol, 0, m.polz, wf, ZK_D1, 5);  // 519
    Z.njobs = 1;
    TRY(zrun(c, Z));
    TRY(kerr(launch_bandj_spec(m, c->s)));
    return inverse_b_and_j(c, m, true);
}

int run_mag_boundary(mhd_ctx* c, int hydro, int mag)
{
    BCArgs b;
    memset(&b, 0, sizeof(b));
    b.g = c->g;
    b.theta = c->p.theta;
    b.u = c->u; b.v = c->v; b.w = c->w; b.t = c->t;
    b.pol = c->pol; b.tor = c->tor; b.bxbar = c->bxbar; b.bybar = c->bybar;
    b.su = c->su; b.sv = c->sv; b.sw = c->sw; b.st = c->st; b.t0 = c->t0;
    b.radbc = (c->p.opts & MHD_OPT_RADBC) != 0;
    b.constflux = (c->p.opts & MHD_OPT_CONSTFLUX) != 0;
    b.perfect = (c->p.opts & MHD_OPT_PERFECT) != 0;
    b.dz = c->dz;
    b.hydro = hydro;
    b.mag = mag;
    return kerr(launch_boundary(b, c->s));
}

// Bz-bar as a per-plane vector (the mixed path's loads take per-plane means)
int fill_bzm(mhd_ctx* c)
{
    if (!c->bzm) return 0;
    return kerr(launch_fill(c->bzm, 0, c->sc.bzbar, c->nz, c->s));
}

// mag_init, src/magnetic.f90:5-279.  The mixed path stores B without its
// means, which live in bxbar, bybar and bzbar and are kept as they are.
int run_mag_init(mhd_ctx* c)
{
    const int wf = c->wf32;
    MagArgs m = margs(c);
    const bool mx = c->su != nullptr;
    if (c->p.opts & MHD_OPT_DYNAMO) {
        // src/magnetic.f90:29-57: the seed field replaces B (1-based i, j in the cosines)
        const int nx = c->nx, ny = c->ny, nz = c->nz;
        const double seed = 0.003, tx = 8.0 * std::atan(1.0) / (double)nx, ty = 8.0 * std::atan(1.0) / (double)ny;
        const bool pc = (c->p.opts & MHD_OPT_PERFECT) != 0;
        std::vector<double> f(c->n3);
        void* dst[3] = {c->bx, c->by, c->bz};
        for (int comp = 0; comp < 3; ++comp) {
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        double v = 0.0;
                        if (pc && comp == 0) v = seed * std::cos(ty * (double)(j + 1));
                        if (pc && comp == 1) v = seed * std::cos(tx * (double)(i + 1));
                        if (!pc && comp == 2) v = seed * std::cos(tx * (double)(i + 1)) * std::cos(ty * (double)(j + 1));
                        f[((long)k * ny + j) * nx + i] = v;
                    }
            // on the context's stream, and finished before f and stage are reused
            MHD_HIP_CHECK(hipMemcpyAsync(c->stage, f.data(), c->n3 * sizeof(double), hipMemcpyHostToDevice, c->s));
            TRY(kerr(launch_copy_cvt(c->stage, 0, dst[comp], c->wf32, c->n3, c->s)));
            MHD_HIP_CHECK(hipStreamSynchronize(c->s));
        }
        if (mx) {
            // the seed has no plane means
            TRY(kerr(launch_fill(c->bxbar, 0, 0.0, nz, c->s)));
            TRY(kerr(launch_fill(c->bybar, 0, 0.0, nz, c->s)));
            c->sc.bzbar = 0.0;
        }
        MHD_HIP_CHECK(hipStreamSynchronize(c->s));
    }
    std::vector<double> keep;
    const double keep_bz = c->sc.bzbar;
    if (mx) {
        keep.resize(2 * c->nz);
        MHD_HIP_CHECK(hipMemcpyAsync(keep.data(), c->bxbar, c->nz * sizeof(double), hipMemcpyDeviceToHost, c->s));
        MHD_HIP_CHECK(hipMemcpyAsync(keep.data() + c->nz, c->bybar, c->nz * sizeof(double), hipMemcpyDeviceToHost,
                                     c->s));
        MHD_HIP_CHECK(hipStreamSynchronize(c->s));
    }
    SpecLaunch L = spec_base(c, 0, SM_R2HC);
    L.job[0] = sjob(c->bx, wf, m.exh, nullptr, 0, 0, wf);
    L.job[1] = sjob(c->by, wf, m.eyh, nullptr, 0, 0, wf);
    L.job[2] = sjob(c->bz, wf, m.ezh, nullptr, 0, 0, wf);
    L.njobs = 3;
    TRY(spec(c, L));
    SpecLaunch M = spec_base(c, 1, SM_R2HC);
    M.job[0] = sjob(m.exh, wf, m.exh, nullptr, 0, 0, wf);
    M.job[1] = sjob(m.eyh, wf, m.eyh, nullptr, 0, 0, wf);
    M.job[2] = sjob(m.ezh, wf, m.ezh, nullptr, 0, 0, wf);
    M.njobs = 3;
    TRY(spec(c, M));
    TRY(kerr(launch_mag_init_spec(m, c->s, c->d_bzbar)));
    MHD_HIP_CHECK(hipMemcpyAsync(&c->sc.bzbar, c->d_bzbar, sizeof(double), hipMemcpyDeviceToHost, c->s));
    MHD_HIP_CHECK(hipStreamSynchronize(c->s));
    if (mx) {
        MHD_HIP_CHECK(hipMemcpyAsync(c->bxbar, keep.data(), c->nz * sizeof(double), hipMemcpyHostToDevice, c->s));
        MHD_HIP_CHECK(hipMemcpyAsync(c->bybar, keep.data() + c->nz, c->nz * sizeof(double), hipMemcpyHostToDevice,
                                     c->s));
        c->sc.bzbar = keep_bz;
        TRY(fill_bzm(c));
        MHD_HIP_CHECK(hipStreamSynchronize(c->s));
    }
    TRY(inverse_b_and_j(c, m, false));
    return run_mag_boundary(c, 0, 1);
}

// 5-field dealias (src/main.F90:495-502), in place
int run_dealias5(mhd_ctx* c)
{
    // the FP64 state is filtered in FP64 in both precision modes
    double* f5[5] = {c->r, c->t, c->u, c->v, c->w};
    for (int dir = 0; dir < 2; ++dir) {
        SpecLaunch L = spec_base(c, dir, SM_DERIV1, true);
        for (int i = 0; i < 5; ++i) L.job[i] = sjob(f5[i], 0, f5[i], nullptr, MUL_MASK, 0, 0);
        L.njobs = 5;
        TRY(spec64(c, L));
    }
    return refresh_shadows(c);
}

RedArgs rargs(mhd_ctx* c, int op)
{
    RedArgs a;
    memset(&a, 0, sizeof(a));
    a.g = c->g;
    a.p = c->ph;
    a.op = op;
    a.u = c->u; a.v = c->v; a.w = c->w; a.r = c->r; a.t = c->t;
    a.bx = c->bx; a.by = c->by; a.bz = c->bz;
    a.wf32 = c->wf32;
    if (c->su) { a.bxm = c->bxbar; a.bym = c->bybar; a.bzm = c->sc.bzbar; }
    a.partial = c->red_partial;
    a.result = c->red_result;
    return a;
}

int fetch_result(mhd_ctx* c, double* out, int n)
{
    MHD_HIP_CHECK(hipMemcpyAsync(c->hbuf, c->red_result, n * sizeof(double), hipMemcpyDeviceToHost, c->s));
    MHD_HIP_CHECK(hipStreamSynchronize(c->s));
    memcpy(out, c->hbuf, n * sizeof(double));
    return 0;
}

int run_global_averages(mhd_ctx* c, double out[4])
{
    RedArgs a = rargs(c, RED_GLOBALAVG);
    TRY(kerr(launch_reduce(a, c->s)));
    TRY(kerr(launch_com_shift(c->g, c->u, c->v, c->red_result, c->s)));
    TRY(refresh_shadows(c));
    return out ? fetch_result(c, out, 4) : 0;
}

int run_nusselt(mhd_ctx* c, double* nu)
{
    RedArgs a = rargs(c, RED_NUSSELT);
    TRY(kerr(launch_reduce(a, c->s)));
    return fetch_result(c, nu, 1);
}

int run_checkdivb(mhd_ctx* c, double out[2])
{
    // src/utils.F90:262-264: d/dx bx + d/dy by + d/dz bz (flag 5)
    const int wf = c->wf32;
    void** W = c->wk;
    SpecLaunch L = spec_base(c, 0, SM_DERIV1);
    L.job[0] = sjob(c->bx, wf, W[W_UY], nullptr, MUL_D1, 0, wf);
    L.njobs = 1;
    TRY(spec(c, L));
    SpecLaunch M = spec_base(c, 1, SM_DERIV1);
    M.job[0] = sjob(c->by, wf, W[W_UYY], nullptr, MUL_D1, 0, wf);
    M.njobs = 1;
    TRY(spec(c, M));
    ZLaunch Z = zbase(c);
    Z.job[0] = zjob(c->bz, wf, W[W_VY], wf, ZK_D1, 5);
    Z.njobs = 1;
    TRY(zrun(c, Z));
    RedArgs a = rargs(c, RED_DIVB);
    a.f0 = W[W_UY]; a.f1 = W[W_UYY]; a.f2 = W[W_VY];
    a.f_f32 = wf;
    TRY(kerr(launch_reduce(a, c->s)));
    return fetch_result(c, out, 2);
}

int run_calcdt(mhd_ctx* c, double* dt)
{
    RedArgs a = rargs(c, RED_CALCDT);
    TRY(kerr(launch_reduce(a, c->s)));
    return fetch_result(c, dt, 1);
}

int check_ctx(mhd_ctx* c) { return c ? 0 : MHD_E_ARG; }

// AB3 coefficients with the Euler / AB2 start-up, src/main.F90:381-403
void ab3_coeffs(mhd_scalars_t& s, double dt, double cf[3])
{
    if (s.first) {
        s.first = 0;
        s.second = 1;
        s.dt1 = 0.0;
        s.dt2 = 0.0;
        cf[0] = dt; cf[1] = 0.0; cf[2] = 0.0;
    } else if (s.second) {
        s.first = 0;
        s.second = 0;
        s.dt2 = 0.0;
        cf[0] = dt * (1.0 + 0.5 * (dt / s.dt1));
        cf[1] = -0.5 * dt * (dt / s.dt1);
        cf[2] = 0.0;
    } else {
        const double dt1 = s.dt1, dt2 = s.dt2;
        cf[0] = dt + (((dt * dt) / (dt1 * dt2)) * ((dt / 3.0) + 0.5 * (dt1 + dt2)));
        cf[1] = ((dt * dt) / (dt1 * (dt1 - dt2))) * ((dt / 3.0) + 0.5 * dt2);
        cf[2] = -((dt * dt) / (dt2 * (dt1 - dt2))) * ((dt / 3.0) + 0.5 * dt1);
    }
}

// f0 = 5: the hydro fields were updated by the fused hydro kernel; pol, tor
// and the means here.  f0 = 7: pol and tor too (fused into mag_rhs); the means
// here.  Then the history ring turns and the state ping-pong swaps.
int run_ab3(mhd_ctx* c, const double cf[3], int f0 = 0)
{
    AB3Args a;
    memset(&a, 0, sizeof(a));
    a.g = c->g;
    a.c0 = cf[0]; a.c1 = cf[1]; a.c2 = cf[2];
    a.f0 = f0;
    double* vars[7] = {c->u, c->v, c->w, c->r, c->t, c->pol, c->tor};
    for (int f = 0; f < 7; ++f) {
        a.var[f] = vars[f];
        for (int l = 0; l < 3; ++l) a.hist[f][l] = level_ptr(c, f, l + 1);
    }
    a.mvar[0] = c->bxbar;
    a.mvar[1] = c->bybar;
    for (int f = 0; f < 2; ++f)
        for (int l = 0; l < 3; ++l) a.mhist[f][l] = mlevel_ptr(c, f, l + 1);
    a.h32 = c->h32;
    if (c->su) {
        a.shadow[0] = c->su; a.shadow[1] = c->sv; a.shadow[2] = c->sw; a.shadow[3] = c->sr; a.shadow[4] = c->st;
        a.bg_t = c->t0;
        a.bg_r = c->r0;
    }
    a.mag = c->ph.mag;
    a.nfields = a.mag ? 7 : 5;
    a.icutx = c->icutx;
    a.icuty = c->icuty;
    if (a.f0 < a.nfields || a.mag) TRY(timed(c, MHD_K_AB3, [&] { return kerr(launch_ab3(a, c->s)); }));
    if (f0 >= 5) {
        std::swap(c->u, c->u2); std::swap(c->v, c->v2); std::swap(c->w, c->w2);
        std::swap(c->t, c->t2); std::swap(c->r, c->r2);
        std::swap(c->su, c->su2); std::swap(c->sv, c->sv2); std::swap(c->sw, c->sw2);
        std::swap(c->st, c->st2); std::swap(c->sr, c->sr2);
    }
    if (f0 >= 7) {
        std::swap(c->pol, c->pol2);
        std::swap(c->tor, c->tor2);
    }
    // shift 2->1, 3->2 (src/main.F90:448-476): relabel the ring
    const int r0 = c->ring[0];
    c->ring[0] = c->ring[1];
    c->ring[1] = c->ring[2];
    c->ring[2] = r0;
    return 0;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void mhd_ra_q_to_ck_f(double gamma, double sigma, double theta, double cm, double zeta, double ra, double q,
                      double* ck, double* f)
{
    // src/initial.F90:187-193
    const double c = std::sqrt((((cm + 1.0) * theta * theta) / (sigma * ra * gamma)) * (cm + 1.0 - (cm * gamma)) *
                               std::pow(1.0 + (0.5 * theta), (2.0 * cm) - 1.0));
    *ck = c;
    *f = q * (zeta * sigma * c * c);
}

const char* mhd_strerror(int code)
{
    switch (code) {
    case MHD_OK: return "ok";
    case MHD_E_ODD_NX: return "nx must be even";
    case MHD_E_ODD_NY: return "ny must be even";
    case MHD_E_ALLOC: return "allocation failed";
    case MHD_E_NZ: return "need at least six layers";
    case MHD_E_FLAG: return "unknown derivative flag";
    case MHD_E_DIVB: return "div B is not zero";
    case MHD_E_ARG: return "bad argument";
    case MHD_E_SIZE: return "horizontal size has no transform kernel";
    case MHD_E_DEVICE: return "HIP runtime error";
    case MHD_E_STATE: return "call out of order";
    case MHD_E_SCRATCH: return "box too large for the no-allocation scratch";
    }
    return "unknown error";
}

int mhd_create(const mhd_params_t* p, int nx, int ny, int nz, int precision, mhd_ctx_t** out)
{
    if (!p || !out) return MHD_E_ARG;
    *out = nullptr;
    if (2 * (nx / 2) != nx) return MHD_E_ODD_NX;
    if (2 * (ny / 2) != ny) return MHD_E_ODD_NY;
    if (nz < 6) return MHD_E_NZ;
    if (!spectral_len_supported(nx) || !spectral_len_supported(ny)) return MHD_E_SIZE;
    if (precision != MHD_PREC_FP64 && precision != MHD_PREC_MIXED) return MHD_E_ARG;
    int ndev = 0, dev = 0, ncu = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MHD_E_DEVICE;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return MHD_E_DEVICE;

    mhd_ctx* c = new mhd_ctx();
    memset((void*)c, 0, sizeof(*c));
    c->nx = nx; c->ny = ny; c->nz = nz;
    c->nxy = nx * ny;
    c->n3 = (long)nx * ny * nz;
    c->prec = precision;
    c->wf32 = precision == MHD_PREC_MIXED;
    c->h32 = c->wf32;
    c->wsz = c->wf32 ? sizeof(float) : sizeof(double);
    c->p = *p;
    // src/initial.F90:305-318: F rescaled for the dynamo and kinematic problems
    if (p->opts & MHD_OPT_DYNAMO) c->p.f = 1.0;
    if (p->opts & MHD_OPT_KINEMATIC) c->p.f = 0.0;
    c->g.nx = nx; c->g.ny = ny; c->g.nz = nz; c->g.nxy = nx * ny; c->g.n3 = c->n3;
    c->icutx = (nx / 2) / 3;
    c->icuty = (ny / 2) / 3;
    set_constants(c);
    c->ring[0] = 0; c->ring[1] = 1; c->ring[2] = 2;
    c->sc.first = 1;
    c->ncu = ncu;
    // launch shapes until mhd_autotune / mhd_set_tune: the autotuner's choices at
    // 256x256x160 on MI355X (y transforms with two resident grids in FP64; 8
    // planes per column-walk thread)
    for (int k = 0; k < 4; ++k) c->spec_var[0][k] = c->spec_var[1][k] = 0;
    for (int k = 0; k < 4; ++k) {
        c->spec_mult[0][k] = 1;
        c->spec_mult[1][k] = c->wf32 ? 1 : 2;
    }
    {
        const int nin = c->nz > 8 ? c->nz - 8 : 1;   // interior planes
        c->hzc = 8 < nin ? 8 : nin;
        // mag_rhs and jsq: 8 planes measured fastest at 256x256x160 (8 to 76
        // tried; 24 planes cost 30-35 us more per step in both precisions)
        c->mzc = 8 < c->nz ? 8 : c->nz;
    }

    if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) { delete c; return MHD_E_DEVICE; }

    // pool layout
    const size_t f64 = al256(c->n3 * sizeof(double));
    const size_t fw = al256(c->n3 * c->wsz);
    const size_t fh = al256(c->n3 * (c->h32 ? sizeof(float) : sizeof(double)));
    const size_t fs = c->wf32 ? al256(c->n3 * sizeof(float)) : 0;   // shadows
    const size_t vz = al256(nz * sizeof(double));
    size_t bytes = 0;
    bytes += 2 * al256(nx * sizeof(double)) + 2 * al256(ny * sizeof(double)) + 2 * vz;   // tables
    bytes += 2 * al256(2 * nx * sizeof(double)) + 2 * al256(2 * ny * sizeof(double));   // twiddles
    bytes += 7 * f64 + 2 * vz;                       // state
    bytes += f64;                                    // FP64 staging for host transfers
    bytes += 4 * fw;                                  // bx by bz jsq
    bytes += 21 * fh + 6 * vz;                        // history
    bytes += 5 * fs + 3 * vz;                         // shadows, backgrounds
    bytes += 7 * f64 + 5 * fs;                        // state ping-pong (u v w t r pol tor)
    c->eta_on = p->eta_c > 1.0;
    bytes += 2 * vz + (c->eta_on ? 3 * fw : 0);       // eta(z) profile, J
    bytes += W_COUNT * fw;                            // work
    bytes += al256(nz * (RED_CHUNKS + 1) * 8 * sizeof(double)) + 2 * al256(8 * sizeof(double));  // reductions
    if (hipMalloc((void**)&c->pool, bytes) != hipSuccess) {
        (void)hipStreamDestroy(c->s);
        delete c;
        return MHD_E_ALLOC;
    }
    c->pool_bytes = bytes;
    // the fill runs on the context's non-blocking stream: it must finish before
    // the tables below are copied in (plain hipMemcpy is not ordered with it)
    if (hipMemsetAsync(c->pool, 0, bytes, c->s) != hipSuccess || hipStreamSynchronize(c->s) != hipSuccess) {
        mhd_destroy(c);
        return MHD_E_DEVICE;
    }
    char* cur = c->pool;
    auto take = [&](size_t b) { char* q = cur; cur += b; return q; };
    c->akx1 = (double*)take(al256(nx * sizeof(double)));
    c->akx2 = (double*)take(al256(nx * sizeof(double)));
    c->aky1 = (double*)take(al256(ny * sizeof(double)));
    c->aky2 = (double*)take(al256(ny * sizeof(double)));
    c->stage = (double*)take(f64);
    c->kappa = (double*)take(vz);
    c->dkappa = (double*)take(vz);
    void* twx64 = take(al256(2 * nx * sizeof(double)));
    void* twx32 = take(al256(2 * nx * sizeof(double)));
    void* twy64 = take(al256(2 * ny * sizeof(double)));
    void* twy32 = take(al256(2 * ny * sizeof(double)));
    c->twx = c->wf32 ? twx32 : twx64;
    c->twy = c->wf32 ? twy32 : twy64;
    c->twx64 = twx64;
    c->twy64 = twy64;
    c->u = (double*)take(f64); c->v = (double*)take(f64); c->w = (double*)take(f64);
    c->t = (double*)take(f64); c->r = (double*)take(f64);
    c->pol = (double*)take(f64); c->tor = (double*)take(f64);
    c->bxbar = (double*)take(vz); c->bybar = (double*)take(vz);
    c->bx = take(fw); c->by = take(fw); c->bz = take(fw); c->jsq = take(fw);
    for (int f = 0; f < 7; ++f)
        for (int l = 0; l < 3; ++l) c->hist[f][l] = take(fh);
    for (int f = 0; f < 2; ++f)
        for (int l = 0; l < 3; ++l) c->mhist[f][l] = (double*)take(vz);
    for (int i = 0; i < W_COUNT; ++i) c->wk[i] = take(fw);
    c->red_partial = (double*)take(al256(nz * (RED_CHUNKS + 1) * 8 * sizeof(double)));
    c->red_result = (double*)take(al256(8 * sizeof(double)));
    c->d_bzbar = (double*)take(al256(8 * sizeof(double)));
    c->t0 = (double*)take(vz);
    c->r0 = (double*)take(vz);
    c->bzm = (double*)take(vz);
    c->etar = (double*)take(vz);
    c->eta1 = (double*)take(vz);
    if (c->eta_on) {
        c->jx = take(fw); c->jy = take(fw); c->jz = take(fw);
    }
    if (c->wf32) {
        c->su = (float*)take(fs); c->sv = (float*)take(fs); c->sw = (float*)take(fs);
        c->st = (float*)take(fs); c->sr = (float*)take(fs);
        c->su2 = (float*)take(fs); c->sv2 = (float*)take(fs); c->sw2 = (float*)take(fs);
        c->st2 = (float*)take(fs); c->sr2 = (float*)take(fs);
    }
    c->u2 = (double*)take(f64); c->v2 = (double*)take(f64); c->w2 = (double*)take(f64);
    c->t2 = (double*)take(f64); c->r2 = (double*)take(f64);
    c->pol2 = (double*)take(f64); c->tor2 = (double*)take(f64);
    if ((size_t)(cur - c->pool) > bytes) {
        fprintf(stderr, "mhd: pool overflow\n");
        mhd_destroy(c);
        return MHD_E_ALLOC;
    }
    if (hipHostMalloc((void**)&c->hbuf, (size_t)c->n3 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        mhd_destroy(c);
        return MHD_E_ALLOC;
    }

    // tables
    std::vector<double> a1, a2, b1, b2, ka, dka, tw64;
    std::vector<float> tw32;
    wavenumbers(nx, p->xmax, a1, a2);
    wavenumbers(ny, p->ymax, b1, b2);
    kappa_profile(nz, c->dz, ka, dka);
    int e = 0;
    e |= hipMemcpy(c->akx1, a1.data(), nx * sizeof(double), hipMemcpyHostToDevice);
    e |= hipMemcpy(c->akx2, a2.data(), nx * sizeof(double), hipMemcpyHostToDevice);
    e |= hipMemcpy(c->aky1, b1.data(), ny * sizeof(double), hipMemcpyHostToDevice);
    e |= hipMemcpy(c->aky2, b2.data(), ny * sizeof(double), hipMemcpyHostToDevice);
    e |= hipMemcpy(c->kappa, ka.data(), nz * sizeof(double), hipMemcpyHostToDevice);
    e |= hipMemcpy(c->dkappa, dka.data(), nz * sizeof(double), hipMemcpyHostToDevice);
    {
        // backgrounds of the mixed path's fluctuations: the polytrope of
        // src/initial.F90:47-53, T0 = 1 + theta z, rho0 = T0^m
        std::vector<double> t0(nz), r0(nz);
        for (int k = 0; k < nz; ++k) {
            t0[k] = 1.0 + (double)k * c->dz * p->theta;
            r0[k] = std::pow(t0[k], p->cm);
        }
        e |= hipMemcpy(c->t0, t0.data(), nz * sizeof(double), hipMemcpyHostToDevice);
        e |= hipMemcpy(c->r0, r0.data(), nz * sizeof(double), hipMemcpyHostToDevice);
        // eta(z)/eta_0 = 1 + (eta_c - 1)(1 + tanh((z - eta_z0)/eta_w))/2 (extension)
        std::vector<double> er(nz, 1.0), e1(nz, 0.0);
        if (c->eta_on)
            for (int k = 0; k < nz; ++k) {
                const double z = (double)k * c->dz;
                er[k] = 1.0 + (p->eta_c - 1.0) * 0.5 * (1.0 + std::tanh((z - p->eta_z0) / p->eta_w));
                e1[k] = c->ph.zck * (er[k] - 1.0);
            }
        e |= hipMemcpy(c->etar, er.data(), nz * sizeof(double), hipMemcpyHostToDevice);
        e |= hipMemcpy(c->eta1, e1.data(), nz * sizeof(double), hipMemcpyHostToDevice);
    }
    twiddles(nx, tw64, tw32);
    e |= hipMemcpy(twx64, tw64.data(), 2 * nx * sizeof(double), hipMemcpyHostToDevice);
    e |= hipMemcpy(twx32, tw32.data(), 2 * nx * sizeof(float), hipMemcpyHostToDevice);
    twiddles(ny, tw64, tw32);
    e |= hipMemcpy(twy64, tw64.data(), 2 * ny * sizeof(double), hipMemcpyHostToDevice);
    e |= hipMemcpy(twy32, tw32.data(), 2 * ny * sizeof(float), hipMemcpyHostToDevice);
    if (e || hipDeviceSynchronize() != hipSuccess) {
        mhd_destroy(c);
        return MHD_E_DEVICE;
    }
    for (int i = 0; i < 2 * MAXEV; ++i)
        if (hipEventCreate(&c->ev[i]) != hipSuccess) {
            mhd_destroy(c);
            return MHD_E_DEVICE;
        } else
            c->ev_created = i + 1;
    *out = c;
    return MHD_OK;
}

void mhd_destroy(mhd_ctx_t* c)
{
    if (!c) return;
    if (c->s) (void)hipStreamSynchronize(c->s);
    for (int i = 0; i < c->ev_created; ++i) (void)hipEventDestroy(c->ev[i]);
    if (c->hbuf) (void)hipHostFree(c->hbuf);
    if (c->pool) (void)hipFree(c->pool);
    if (c->cl_mem) (void)hipFree(c->cl_mem);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
}

static int upload(mhd_ctx* c, void* dst, int dst_f32, const double* src, long n)
{
    if (!src) return 0;
    memcpy(c->hbuf, src, n * sizeof(double));
    if (!dst_f32) {
        MHD_HIP_CHECK(hipMemcpyAsync(dst, c->hbuf, n * sizeof(double), hipMemcpyHostToDevice, c->s));
    } else {
        MHD_HIP_CHECK(hipMemcpyAsync(c->stage, c->hbuf, n * sizeof(double), hipMemcpyHostToDevice, c->s));
        TRY(kerr(launch_copy_cvt(c->stage, 0, dst, 1, n, c->s)));
    }
    MHD_HIP_CHECK(hipStreamSynchronize(c->s));
    return 0;
}

static int download(mhd_ctx* c, double* dst, const void* src, int src_f32, long n)
{
    if (!dst) return 0;
    const void* from = src;
    if (src_f32) {
        TRY(kerr(launch_copy_cvt(src, 1, c->stage, 0, n, c->s)));
        from = c->stage;
    }
    MHD_HIP_CHECK(hipMemcpyAsync(c->hbuf, from, n * sizeof(double), hipMemcpyDeviceToHost, c->s));
    MHD_HIP_CHECK(hipStreamSynchronize(c->s));
    memcpy(dst, c->hbuf, n * sizeof(double));
    return 0;
}

// mixed precision: the per-plane mean of B component comp (0 bx, 1 by, 2 bz)
static int b_means(mhd_ctx* c, int comp, std::vector<double>& m)
{
    m.assign(c->nz, 0.0);
    if (comp == 2) {
        for (int k = 0; k < c->nz; ++k) m[k] = c->sc.bzbar;
        return 0;
    }
    MHD_HIP_CHECK(hipMemcpyAsync(m.data(), comp == 0 ? c->bxbar : c->bybar, c->nz * sizeof(double),
                                 hipMemcpyDeviceToHost, c->s));
    MHD_HIP_CHECK(hipStreamSynchronize(c->s));
    return 0;
}

// B component to the host, with its means added back in the mixed path
static int download_b(mhd_ctx* c, double* dst, int comp)
{
    if (!dst) return 0;
    const void* src = comp == 0 ? c->bx : (comp == 1 ? c->by : c->bz);
    TRY(download(c, dst, src, c->wf32, c->n3));
    if (!c->su) return 0;
    std::vector<double> m;
    TRY(b_means(c, comp, m));
    for (int k = 0; k < c->nz; ++k)
        for (long q = 0; q < c->nxy; ++q) dst[(long)k * c->nxy + q] += m[k];
    return 0;
}

int mhd_static(mhd_ctx_t* c)
{
    TRY(check_ctx(c));
    // src/initial.F90:25-86 on the host (integer RNG), then upload
    const int nz = c->nz, m = c->nxy;
    const long n3 = c->n3;
    // src/initial.F90:63-78: B = x between perfect conductors, B = z otherwise
    const bool pc = (c->p.opts & MHD_OPT_PERFECT) != 0;
    std::vector<double> t(n3), r(n3), zero(n3, 0.0), bx(n3, pc ? 1.0 : 0.0), bz(n3, pc ? 0.0 : 1.0);
    for (int k = 0; k < nz; ++k) {
        const double z = (double)k * c->dz;
        const double t0 = 1.0 + z * c->p.theta;
        const double r0 = std::pow(t0, c->p.cm);
        int iseed = 1996 + 64 * (k + 1);
        for (int q = 0; q < m; ++q) t[(long)k * m + q] = 2.0 * (ran0(&iseed) - 0.5) * c->p.t_amp;
        for (int q = 0; q < m; ++q) {
            r[(long)k * m + q] = r0;
            t[(long)k * m + q] = t[(long)k * m + q] + t0;
        }
        if (c->p.b_amp != 0.0) {
            int bseed = 1996 + 64 * (k + 1) + 32;
            for (int q = 0; q < m; ++q) bz[(long)k * m + q] += 2.0 * (ran0(&bseed) - 0.5) * c->p.b_amp;
        }
    }
    return mhd_set_state(c, r.data(), t.data(), zero.data(), zero.data(), zero.data(), bx.data(), zero.data(),
                         bz.data());
}

int mhd_set_state(mhd_ctx_t* c, const double* r, const double* t, const double* u, const double* v, const double* w,
                  const double* bx, const double* by, const double* bz)
{
    TRY(check_ctx(c));
    if (!r || !t || !u || !v || !w || !bx || !by || !bz) return MHD_E_ARG;
    const long n = c->n3;
    TRY(upload(c, c->r, 0, r, n));
    TRY(upload(c, c->t, 0, t, n));
    TRY(upload(c, c->u, 0, u, n));
    TRY(upload(c, c->v, 0, v, n));
    TRY(upload(c, c->w, 0, w, n));
    if (!c->su) {
        TRY(upload(c, c->bx, c->wf32, bx, n));
        TRY(upload(c, c->by, c->wf32, by, n));
        TRY(upload(c, c->bz, c->wf32, bz, n));
    } else {
        // mixed: B is stored as FP32 fluctuations about its means, which are
        // the plane averages of bx, by and the bottom-plane average of bz
        // (the DC terms mag_init takes, src/magnetic.f90:63-90)
        const int nz = c->nz;
        const long m = c->nxy;
        std::vector<double> mxv(nz, 0.0), myv(nz, 0.0), fl(n);
        double mz = 0.0;
        for (int k = 0; k < nz; ++k) {
            for (long q = 0; q < m; ++q) {
                mxv[k] += bx[(long)k * m + q];
                myv[k] += by[(long)k * m + q];
            }
            mxv[k] /= (double)m;
            myv[k] /= (double)m;
        }
        for (long q = 0; q < m; ++q) mz += bz[q];
        mz /= (double)m;
        for (long q = 0; q < n; ++q) fl[q] = bx[q] - mxv[q / m];
        TRY(upload(c, c->bx, 1, fl.data(), n));
        for (long q = 0; q < n; ++q) fl[q] = by[q] - myv[q / m];
        TRY(upload(c, c->by, 1, fl.data(), n));
        for (long q = 0; q < n; ++q) fl[q] = bz[q] - mz;
        TRY(upload(c, c->bz, 1, fl.data(), n));
        MHD_HIP_CHECK(hipMemcpy(c->bxbar, mxv.data(), nz * sizeof(double), hipMemcpyHostToDevice));
        MHD_HIP_CHECK(hipMemcpy(c->bybar, myv.data(), nz * sizeof(double), hipMemcpyHostToDevice));
        c->sc.bzbar = mz;
        TRY(fill_bzm(c));
    }
    if (c->ph.mag) TRY(run_mag_init(c));
    TRY(refresh_shadows(c));
    c->have_state = 1;
    return mhd_sync(c);
}

int mhd_get_state(mhd_ctx_t* c, double* r, double* t, double* u, double* v, double* w, double* bx, double* by,
                  double* bz)
{
    TRY(check_ctx(c));
    const long n = c->n3;
    TRY(download(c, r, c->r, 0, n));
    TRY(download(c, t, c->t, 0, n));
    TRY(download(c, u, c->u, 0, n));
    TRY(download(c, v, c->v, 0, n));
    TRY(download(c, w, c->w, 0, n));
    TRY(download_b(c, bx, 0));
    TRY(download_b(c, by, 1));
    TRY(download_b(c, bz, 2));
    return 0;
}

static int field_ref(mhd_ctx* c, int id, int level, void** ptr, int* f32, long* n)
{
    *f32 = 0;
    *n = c->n3;
    switch (id) {
    case MHD_F_U: *ptr = c->u; return 0;
    case MHD_F_V: *ptr = c->v; return 0;
    case MHD_F_W: *ptr = c->w; return 0;
    case MHD_F_T: *ptr = c->t; return 0;
    case MHD_F_R: *ptr = c->r; return 0;
    case MHD_F_POL: *ptr = c->pol; return 0;
    case MHD_F_TOR: *ptr = c->tor; return 0;
    case MHD_F_BX: *ptr = c->bx; *f32 = c->wf32; return 0;
    case MHD_F_BY: *ptr = c->by; *f32 = c->wf32; return 0;
    case MHD_F_BZ: *ptr = c->bz; *f32 = c->wf32; return 0;
    case MHD_F_JSQ: *ptr = c->jsq; *f32 = c->wf32; return 0;
    case MHD_F_BXBAR: *ptr = c->bxbar; *n = c->nz; return 0;
    case MHD_F_BYBAR: *ptr = c->bybar; *n = c->nz; return 0;
    case MHD_F_KAPPA: *ptr = c->kappa; *n = c->nz; return 0;
    case MHD_F_DKAPPA: *ptr = c->dkappa; *n = c->nz; return 0;
    }
    if (level < 1 || level > 3) return MHD_E_ARG;
    if (id == MHD_F_BXBARDOT || id == MHD_F_BYBARDOT) {
        *ptr = mlevel_ptr(c, id == MHD_F_BXBARDOT ? 0 : 1, level);
        *n = c->nz;
        return 0;
    }
    const int h = hist_index(id);
    if (h < 0) return MHD_E_ARG;
    *ptr = level_ptr(c, h, level);
    *f32 = c->h32;
    return 0;
}

int mhd_get_field(mhd_ctx_t* c, int id, int level, double* dst)
{
    TRY(check_ctx(c));
    if (!dst) return MHD_E_ARG;
    void* p;
    int f32;
    long n;
    TRY(field_ref(c, id, level, &p, &f32, &n));
    if (id == MHD_F_BX || id == MHD_F_BY || id == MHD_F_BZ) return download_b(c, dst, id - MHD_F_BX);
    return download(c, dst, p, f32, n);
}

int mhd_set_field(mhd_ctx_t* c, int id, int level, const double* src)
{
    TRY(check_ctx(c));
    if (!src) return MHD_E_ARG;
    void* p;
    int f32;
    long n;
    TRY(field_ref(c, id, level, &p, &f32, &n));
    // mixed: B is kept as fluctuations about its means; set it with mhd_set_state
    if (c->su && (id == MHD_F_BX || id == MHD_F_BY || id == MHD_F_BZ)) return MHD_E_ARG;
    TRY(upload(c, p, f32, src, n));
    if (id == MHD_F_U || id == MHD_F_V || id == MHD_F_W || id == MHD_F_T || id == MHD_F_R)
        TRY(refresh_shadows(c));
    return mhd_sync(c);
}

int mhd_get_scalars(mhd_ctx_t* c, mhd_scalars_t* s)
{
    TRY(check_ctx(c));
    if (!s) return MHD_E_ARG;
    *s = c->sc;
    return 0;
}

int mhd_set_scalars(mhd_ctx_t* c, const mhd_scalars_t* s)
{
    TRY(check_ctx(c));
    if (!s) return MHD_E_ARG;
    c->sc = *s;
    TRY(fill_bzm(c));
    return mhd_sync(c);
}

int mhd_hydro_deriv(mhd_ctx_t* c)
{
    TRY(check_ctx(c));
    TRY(run_hydro_deriv(c));
    return mhd_sync(c);
}

int mhd_mag_deriv(mhd_ctx_t* c)
{
    TRY(check_ctx(c));
    TRY(run_mag_deriv(c));
    return mhd_sync(c);
}

int mhd_bandj(mhd_ctx_t* c)
{
    TRY(check_ctx(c));
    TRY(run_bandj(c));
    return mhd_sync(c);
}

int mhd_magdealias(mhd_ctx_t* c)
{
    TRY(check_ctx(c));
    TRY(kerr(launch_magdealias(c->g, c->pol, c->tor, c->icutx, c->icuty, c->s)));
    return mhd_sync(c);
}

int mhd_boundary(mhd_ctx_t* c)
{
    TRY(check_ctx(c));
    TRY(run_mag_boundary(c, 1, 0));
    return mhd_sync(c);
}

int mhd_mag_boundary(mhd_ctx_t* c)
{
    TRY(check_ctx(c));
    TRY(run_mag_boundary(c, 0, 1));
    return mhd_sync(c);
}

int mhd_mag_init(mhd_ctx_t* c)
{
    TRY(check_ctx(c));
    TRY(run_mag_init(c));
    return mhd_sync(c);
}

int mhd_step(mhd_ctx_t* c, int nsteps, double dt_fixed, int ndu06, mhd_diag_t* last)
{
    TRY(check_ctx(c));
    if (!c->have_state) return MHD_E_STATE;
    const bool mag = c->ph.mag;
    if (dt_fixed > 0.0) c->sc.dt = dt_fixed;
    else if (c->sc.dt <= 0.0) TRY(run_calcdt(c, &c->sc.dt));   // src/main.F90:318
    for (int s = 0; s < nsteps; ++s) {
        mhd_scalars_t& S = c->sc;
        S.nit += 1;
        S.iter += 1;
        S.ttime += S.dt;
        double cf[3];
        const double dt = S.dt;
        ab3_coeffs(S, dt, cf);
        if (c->closure) {
            // the closure's terms join the right-hand sides, so the update stays separate
            TRY(run_hydro_deriv(c));
            if (mag) TRY(run_mag_deriv(c));
            TRY(run_ab3(c, cf));
        } else {
            TRY(run_hydro_deriv(c, cf));       // hydro RHS + AB3 of u v w r t into the ping-pong
            if (mag) TRY(run_mag_deriv(c, cf)); // reads the old u v w; + AB3 of pol, tor, magdealias
            TRY(run_ab3(c, cf, mag ? 7 : 5));  // the means (src/main.F90:508); swaps
        }
        S.dt2 = dt + S.dt1;
        S.dt1 = dt;
        if (S.nit % 10 == 0) TRY(run_dealias5(c));
        if (mag) TRY(run_bandj(c));
        TRY(run_mag_boundary(c, 1, mag ? 1 : 0));
        if (ndu06 > 0 && S.nit % ndu06 == 0) {
            double ga[4], nu, db[2];
            TRY(run_global_averages(c, ga));
            TRY(run_nusselt(c, &nu));
            TRY(run_checkdivb(c, db));
            if (last) {
                last->ubar = ga[0];
                last->vbar = ga[1];
                last->kinenergy = ga[2];
                // src/main.F90:567-571
                last->magenergy = (c->p.opts & MHD_OPT_KINEMATIC) ? ga[3] : c->p.f * ga[3];
                last->nusselt = nu;
                last->divbmax = db[0];
                last->divbmin = db[1];
                last->dt = S.dt;
                last->nit = S.nit;
            }
            if (std::fmax(std::fabs(db[0]), std::fabs(db[1])) > 0.01) return MHD_E_DIVB;
        }
        if (dt_fixed <= 0.0 && S.nit % 10 == 0) TRY(run_calcdt(c, &S.dt));
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Ensembles: independent boxes stepped by one hipGraph launch per block of
// ENS_BLOCK steps.  With a fixed dt the AB3 coefficients are constant after
// the Euler and AB2 start-up steps, the history ring has period 3 and the
// dealias cadence period 10, so the kernel sequence of 30 steps repeats
// exactly.  Each box's sequence is recorded (launch.hpp); the k-th launches
// of all boxes of a class (same kernels, i.e. the same lengths, precision
// and options) become one launch over the table of their argument blocks,
// and the graph holds those batched launches, one branch per class.
// ---------------------------------------------------------------------------
}  // extern "C"

struct mhd_ensemble {
    std::vector<mhd_ctx*> box;
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    std::vector<hipStream_t> branch;   // one stream per class of boxes
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* tables = nullptr;            // device: argument blocks, starts and grids of the batched launches
    double dt = 0.0;
    long nodes = 0;
    int classes = 0;
};

namespace {
constexpr int ENS_BLOCK = 30;

bool ab3_steady(const mhd_ctx* c, double dt)
{
    return !c->sc.first && !c->sc.second && c->sc.dt == dt && c->sc.dt1 == dt && c->sc.dt2 == dt + dt;
}

void ens_free(mhd_ensemble* e)
{
    if (e->s) (void)hipStreamSynchronize(e->s);
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    for (hipEvent_t ev : e->join)
        if (ev) (void)hipEventDestroy(ev);
    for (hipStream_t st : e->branch)
        if (st) (void)hipStreamDestroy(st);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->s) (void)hipStreamDestroy(e->s);
    if (e->tables) (void)hipFree(e->tables);
    delete e;
}

// The kernel launches of ENS_BLOCK steps of one box, recorded instead of
// issued.  The box's stream is put in capture mode meanwhile: any launch that
// bypassed the recorder would land in that throwaway graph, and is reported
// as MHD_E_STATE.  The host-side bookkeeping (driver scalars, history ring)
// is put back afterwards; the state ping-pong swaps an even number of times.
int record_block(mhd_ctx* c, double dt, LaunchRecorder& rec)
{
    const mhd_scalars_t keep = c->sc;
    const int ring[3] = {c->ring[0], c->ring[1], c->ring[2]};
    MHD_HIP_CHECK(hipStreamBeginCapture(c->s, hipStreamCaptureModeRelaxed));
    launch_recorder() = &rec;
    const int rc = mhd_step(c, ENS_BLOCK, dt, 0, nullptr);
    launch_recorder() = nullptr;
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(c->s, &g);
    size_t stray = 0;
    if (g) {
        (void)hipGraphGetNodes(g, nullptr, &stray);
        (void)hipGraphDestroy(g);
    }
    c->sc = keep;
    for (int l = 0; l < 3; ++l) c->ring[l] = ring[l];
    if (rc) return rc;
    if (ce != hipSuccess) return MHD_E_DEVICE;
    return stray ? MHD_E_STATE : 0;
}

// Boxes whose recorded sequences launch the same kernels with the same
// workgroup sizes (same lengths, precision and options) form a class; the
// k-th launches of a class's boxes become one batched launch over a table of
// the boxes' argument blocks and grids (launch.hpp).  Classes run on
// parallel branches of the graph.
int build_batched(mhd_ensemble* e)
{
    const int nbox = (int)e->box.size();
    std::vector<LaunchRecorder> rec(nbox);
    for (int b = 0; b < nbox; ++b) TRY(record_block(e->box[b], e->dt, rec[b]));
    // classes by signature
    std::vector<std::vector<int>> cls;
    auto same = [&](int a, int b) {
        const auto& x = rec[a].recs;
        const auto& y = rec[b].recs;
        if (x.size() != y.size()) return false;
        for (size_t k = 0; k < x.size(); ++k)
            if (x[k].kernel != y[k].kernel || x[k].block.x != y[k].block.x || x[k].args.size() != y[k].args.size())
                return false;
        return true;
    };
    for (int b = 0; b < nbox; ++b) {
        bool put = false;
        for (auto& c : cls)
            if (same(c[0], b)) {
                c.push_back(b);
                put = true;
                break;
            }
        if (!put) cls.push_back({b});
    }
    for (int b = 0; b < nbox; ++b)
        for (const LaunchRec& r : rec[b].recs)
            if (r.block.y != 1 || r.block.z != 1) return MHD_E_STATE;   // 1-D workgroups only
    // tables: per batched launch, the argument blocks, nbox + 1 starts and nbox grids
    struct Batch {
        size_t args_off, start_off, grid_off;
        unsigned total;
        size_t lds;
        int cls, k;
    };
    std::vector<Batch> batches;
    std::vector<unsigned char> host;
    auto put = [&](const void* p, size_t n) {
        size_t off = (host.size() + 255) & ~(size_t)255;
        host.resize(off + n);
        memcpy(host.data() + off, p, n);
        return off;
    };
    for (int ci = 0; ci < (int)cls.size(); ++ci) {
        const auto& members = cls[ci];
        const size_t nk = rec[members[0]].recs.size();
        for (size_t k = 0; k < nk; ++k) {
            const size_t stride = rec[members[0]].recs[k].args.size();
            std::vector<unsigned char> blob(stride * members.size());
            std::vector<unsigned> start(members.size() + 1, 0u);
            std::vector<uint3> grid(members.size());
            size_t lds = 0;
            for (size_t m = 0; m < members.size(); ++m) {
                const LaunchRec& r = rec[members[m]].recs[k];
                memcpy(blob.data() + m * stride, r.args.data(), stride);
                grid[m] = uint3{r.grid.x, r.grid.y, r.grid.z};
                const unsigned long long nb = (unsigned long long)r.grid.x * r.grid.y * r.grid.z;
                if (start[m] + nb > 0xffffffffULL) return MHD_E_ARG;
                start[m + 1] = start[m] + (unsigned)nb;
                lds = r.lds > lds ? r.lds : lds;
            }
            Batch bt;
            bt.args_off = put(blob.data(), blob.size());
            bt.start_off = put(start.data(), start.size() * sizeof(unsigned));
            bt.grid_off = put(grid.data(), grid.size() * sizeof(uint3));
            bt.total = start.back();
            bt.lds = lds;
            bt.cls = ci;
            bt.k = (int)k;
            batches.push_back(bt);
        }
    }
    MHD_HIP_CHECK(hipMalloc(&e->tables, host.size() ? host.size() : 256));
    MHD_HIP_CHECK(hipMemcpy(e->tables, host.data(), host.size(), hipMemcpyHostToDevice));
    e->classes = (int)cls.size();
    e->branch.assign(cls.size(), nullptr);
    e->join.assign(cls.size(), nullptr);
    for (size_t ci = 0; ci < cls.size(); ++ci) {
        MHD_HIP_CHECK(hipStreamCreateWithFlags(&e->branch[ci], hipStreamNonBlocking));
        MHD_HIP_CHECK(hipEventCreateWithFlags(&e->join[ci], hipEventDisableTiming));
    }
    // capture: one branch per class, its batched launches in recorded order
    MHD_HIP_CHECK(hipStreamBeginCapture(e->s, hipStreamCaptureModeRelaxed));
    int rc = 0;
    if (hipEventRecord(e->fork, e->s) != hipSuccess) rc = MHD_E_DEVICE;
    for (size_t ci = 0; ci < cls.size() && !rc; ++ci)
        if (hipStreamWaitEvent(e->branch[ci], e->fork, 0) != hipSuccess) rc = MHD_E_DEVICE;
    const unsigned char* dev = (const unsigned char*)e->tables;
    for (const Batch& bt : batches) {
        if (rc) break;
        const auto& members = cls[bt.cls];
        const LaunchRec& r0 = rec[members[0]].recs[bt.k];
        BatchTable t;
        t.args = dev + bt.args_off;
        t.start = (const unsigned*)(dev + bt.start_off);
        t.grid = (const uint3*)(dev + bt.grid_off);
        t.nbox = (int)members.size();
        t.stride = (int)r0.args.size();
        if (bt.total > 0 && r0.batch(t, bt.total, r0.block.x, bt.lds, e->branch[bt.cls])) rc = MHD_E_DEVICE;
    }
    for (size_t ci = 0; ci < cls.size() && !rc; ++ci) {
        if (hipEventRecord(e->join[ci], e->branch[ci]) != hipSuccess ||
            hipStreamWaitEvent(e->s, e->join[ci], 0) != hipSuccess)
            rc = MHD_E_DEVICE;
    }
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(e->s, &g);
    e->graph = g;
    if (rc) return rc;
    if (ce != hipSuccess || !g) return MHD_E_DEVICE;
    return 0;
}
}  // namespace

extern "C" {

int mhd_ensemble_create(mhd_ctx_t* const* boxes, int nbox, double dt, mhd_ensemble_t** out)
{
    if (!boxes || nbox < 1 || !(dt > 0.0) || !out) return MHD_E_ARG;
    *out = nullptr;
    for (int b = 0; b < nbox; ++b) {
        if (!boxes[b]) return MHD_E_ARG;
        if (!boxes[b]->have_state) return MHD_E_STATE;
        if (boxes[b]->closure) return MHD_E_ARG;
        for (int b2 = 0; b2 < b; ++b2)
            if (boxes[b2] == boxes[b]) return MHD_E_ARG;
    }
    // start-up steps until the AB3 coefficients are those of a constant dt
    for (int b = 0; b < nbox; ++b) {
        mhd_ctx* c = boxes[b];
        for (int n = 0; n < 4 && !ab3_steady(c, dt); ++n) TRY(mhd_step(c, 1, dt, 0, nullptr));
        if (!ab3_steady(c, dt)) return MHD_E_STATE;
        c->timed_kernel = MHD_K_NONE;
        TRY(mhd_sync(c));
    }
    mhd_ensemble* e = new mhd_ensemble();
    e->box.assign(boxes, boxes + nbox);
    e->dt = dt;
    auto fail = [&](int code) {
        ens_free(e);
        return code;
    };
    if (hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess)
        return fail(MHD_E_DEVICE);
    {
        const int rc = build_batched(e);
        if (rc) return fail(rc);
    }
    hipGraph_t g = e->graph;
    size_t nn = 0;
    if (hipGraphGetNodes(g, nullptr, &nn) == hipSuccess) e->nodes = (long)nn;
    if (hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0) != hipSuccess) return fail(MHD_E_DEVICE);
    *out = e;
    return MHD_OK;
}

int mhd_ensemble_step(mhd_ensemble_t* e, int nsteps)
{
    if (!e || nsteps < 0 || nsteps % ENS_BLOCK) return MHD_E_ARG;
    const int blocks = nsteps / ENS_BLOCK;
    for (int i = 0; i < blocks; ++i) MHD_HIP_CHECK(hipGraphLaunch(e->exec, e->s));
    MHD_HIP_CHECK(hipStreamSynchronize(e->s));
    MHD_HIP_CHECK(hipGetLastError());
    // the host-side driver scalars advance as mhd_step would advance them
    for (mhd_ctx* c : e->box) {
        for (int n = 0; n < nsteps; ++n) {
            c->sc.nit += 1;
            c->sc.iter += 1;
            c->sc.ttime += c->sc.dt;
        }
    }
    return MHD_OK;
}

int mhd_ensemble_block(void) { return ENS_BLOCK; }

// ---------------------------------------------------------------------------
// Autotuner: the launch shapes of this box timed on this GPU
// ---------------------------------------------------------------------------
}  // extern "C"
namespace {
// average milliseconds of `reps` calls of f on the context's stream
// Autotuner timing: repetitions per candidate, and the margin a later
// candidate must win by to replace an earlier one (the defaults come first),
// so run-to-run noise of about 1 % does not move the choice.
constexpr int TUNE_REPS = 5;
constexpr double TUNE_GAIN = 0.99;

template <typename F>
int time_ms(mhd_ctx* c, int reps, F&& f, double* ms)
{
    hipEvent_t a, b;
    MHD_HIP_CHECK(hipEventCreate(&a));
    MHD_HIP_CHECK(hipEventCreate(&b));
    int rc = f();   // warm-up
    if (!rc) {
        (void)hipEventRecord(a, c->s);
        for (int r = 0; r < reps && !rc; ++r) rc = f();
        (void)hipEventRecord(b, c->s);
        (void)hipEventSynchronize(b);
        float t = 0.f;
        (void)hipEventElapsedTime(&t, a, b);
        *ms = (double)t / reps;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}
}  // namespace
extern "C" {

int mhd_set_tune(mhd_ctx_t* c, const mhd_tune_t* t)
{
    TRY(check_ctx(c));
    if (!t) return MHD_E_ARG;
    for (int k = 0; k < MHD_TK_N; ++k)
        if (t->variant_x[k] < 0 || t->variant_y[k] < 0 || t->variant_x[k] >= spectral_variants(c->nx) ||
            t->variant_y[k] >= spectral_variants(c->ny))
            return MHD_E_ARG;
    for (int k = 0; k < MHD_TK_N; ++k)
        if (t->mult_x[k] < 1 || t->mult_x[k] > 8 || t->mult_y[k] < 1 || t->mult_y[k] > 8) return MHD_E_ARG;
    if (t->zc < 1) return MHD_E_ARG;
    for (int k = 0; k < MHD_TK_N; ++k) {
        c->spec_var[0][k] = t->variant_x[k];
        c->spec_var[1][k] = t->variant_y[k];
        c->spec_mult[0][k] = t->mult_x[k];
        c->spec_mult[1][k] = t->mult_y[k];
    }
    c->hzc = t->zc;
    return 0;
}

int mhd_autotune(mhd_ctx_t* c, mhd_tune_t* out)
{
    TRY(check_ctx(c));
    if (!c->have_state) return MHD_E_STATE;
    const int saved_timing = c->timed_kernel;
    c->timed_kernel = MHD_K_NONE;
    const int wf = c->wf32;
    void** W = c->wk;
    const void* in[4] = {c->su ? (const void*)c->st : c->t, c->su ? (const void*)c->su : c->u,
                         c->su ? (const void*)c->sv : c->v, c->su ? (const void*)c->sw : c->w};
    const int sf = c->su ? 1 : 0;
    double best_ms[2] = {0, 0};
    // transforms, per direction and kind: the radix plan and the
    // persistent-grid multiple, on a launch of that kind shaped like the
    // step's (outputs into work arrays); the current choice (the default at
    // first) is timed first
    const int o1[2][4] = {{W_TX, W_UX, W_VX, W_WX}, {W_TY, W_UY, W_VY, W_WY}};
    const int o2[2][4] = {{W_TXX, W_UXX, W_VXX, W_WXX}, {W_TYY, W_UYY, W_VYY, W_WYY}};
    auto kind_launch = [&](int dir, int kind) {
        SpecLaunch L = spec_base(c, dir, kind);
        for (int j = 0; j < 4; ++j) {
            void* a = W[o1[dir][j]];
            if (kind == SM_DERIV2) L.job[j] = sjob(in[j], sf, a, W[o2[dir][j]], MUL_D1, MUL_D2, wf);
            else if (kind == SM_DERIV1) L.job[j] = sjob(in[j], sf, a, nullptr, MUL_D1, 0, wf);
            else L.job[j] = sjob(in[j], sf, a, nullptr, 0, 0, wf);
        }
        L.njobs = 4;
        if (kind == SM_DERIV2) {
            // as in the step: a single-output (first-derivative) job beside the four
            L.job[4] = sjob(in[0], sf, W[dir == 0 ? W_RX : W_RY], nullptr, MUL_D1, 0, wf);
            L.njobs = 5;
        } else if (kind == SM_DERIV1) {
            // as in the step: the first derivatives of one direction's three
            // Maxwell-stress loads (products of three fields)
            L.job[0].in[0] = in[2]; L.job[0].load_op = LD_PROD; L.job[0].in[1] = in[1];
            L.job[1].in[0] = in[3]; L.job[1].load_op = LD_PROD; L.job[1].in[1] = in[1];
            L.job[2].in[0] = in[1]; L.job[2].load_op = LD_MAXW; L.job[2].in[1] = in[2]; L.job[2].in[2] = in[3];
            L.njobs = 3;
        }
        return L;
    };
    for (int dir = 0; dir < 2; ++dir) {
        const int n = dir == 0 ? c->nx : c->ny;
        const int mdef = (dir == 1 && !c->wf32) ? 2 : 1;
        for (int kind = 0; kind < MHD_TK_N; ++kind) {
            double bk = 1e30;
            int bv = 0, bm = mdef;
            for (int var = 0; var < spectral_variants(n); ++var) {
                for (int mi = 0; mi < 2; ++mi) {
                    const int mult = mi == 0 ? mdef : 3 - mdef;
                    c->spec_var[dir][kind] = var;
                    c->spec_mult[dir][kind] = mult;
                    const SpecLaunch L = kind_launch(dir, kind);
                    double ms = 0.0;
                    TRY(time_ms(c, TUNE_REPS, [&] { return spec(c, L); }, &ms));
                    if (ms < TUNE_GAIN * bk) { bk = ms; bv = var; bm = mult; }
                }
            }
            c->spec_var[dir][kind] = bv;
            c->spec_mult[dir][kind] = bm;
            if (kind == SM_DERIV2) best_ms[dir] = bk;
        }
    }
    // column-walking kernels: planes per thread
    double bestc = 1e30;
    int bz = c->hzc;
    const int nin = c->nz > 8 ? c->nz - 8 : 1;
    const int cands[] = {8, 12, 16, 20, 24, 32, 48};
    for (int zc : cands) {
        if (zc > nin && zc != cands[0]) continue;
        c->hzc = zc;
        HydroColArgs a = colargs(c);
        if (!c->closure) {
            // the variant the step runs: with the fused AB3 update (zero
            // coefficients; the new state goes to the free ping-pong buffers
            // and the floor's history clears to a free work array)
            const double cf0[3] = {0.0, 0.0, 0.0};
            fuse_ab3(c, a, cf0, W[W_RX], W[W_RX]);
        }
        double ms = 0.0;
        TRY(time_ms(c, TUNE_REPS, [&] { return kerr(launch_hydro_col(a, c->wf32, c->s)); }, &ms));
        if (ms < TUNE_GAIN * bestc) { bestc = ms; bz = zc; }
    }
    c->hzc = bz;
    c->timed_kernel = saved_timing;
    if (out) {
        for (int k = 0; k < MHD_TK_N; ++k) {
            out->variant_x[k] = c->spec_var[0][k];
            out->variant_y[k] = c->spec_var[1][k];
        }
        for (int k = 0; k < MHD_TK_N; ++k) {
            out->mult_x[k] = c->spec_mult[0][k];
            out->mult_y[k] = c->spec_mult[1][k];
        }
        out->zc = c->hzc;
        out->ms_x = best_ms[0];
        out->ms_y = best_ms[1];
        out->ms_col = bestc;
    }
    return mhd_sync(c);
}

long mhd_ensemble_nodes(mhd_ensemble_t* e) { return e ? e->nodes : 0; }
int mhd_ensemble_classes(mhd_ensemble_t* e) { return e ? e->classes : 0; }

void mhd_ensemble_destroy(mhd_ensemble_t* e)
{
    if (e) ens_free(e);
}

int mhd_closure_attach(mhd_ctx_t* c, mhd_closure_t* cl)
{
    TRY(check_ctx(c));
    if (cl && c->eta_on) return MHD_E_ARG;
    TRY(mhd_sync(c));
    if (!cl) {
        c->closure = nullptr;
        return 0;
    }
    if (!c->cl_mem) {
        const size_t fw = (size_t)c->n3 * (c->wf32 ? 4 : 8);
        const size_t fa = (fw + 255) & ~(size_t)255;
        const size_t ga = (closure_grid_bytes(c->n3) + 255) & ~(size_t)255;
        if (hipMalloc(&c->cl_mem, 24 * fa + ga + (size_t)c->nz * 8) != hipSuccess) {
            c->cl_mem = nullptr;
            return MHD_E_ALLOC;
        }
        char* p = (char*)c->cl_mem;
        for (int q = 0; q < 12; ++q, p += fa) c->cflux[q] = p;
        for (int m = 0; m < 4; ++m, p += fa) c->cdx[m] = p;
        for (int m = 0; m < 4; ++m, p += fa) c->cdy[m] = p;
        for (int m = 0; m < 4; ++m, p += fa) c->cdz[m] = p;
        c->cgrid = p;   // this run's feature grid: one closure may serve several runs
        p += ga;
        c->cneg1 = (double*)p;
        std::vector<double> neg(c->nz, -1.0);
        MHD_HIP_CHECK(hipMemcpy(c->cneg1, neg.data(), (size_t)c->nz * 8, hipMemcpyHostToDevice));
    }
    c->closure = cl;
    return 0;
}

int mhd_closure_add_state(mhd_closure_t* cl, mhd_ctx_t* fine, int f)
{
    if (!cl) return MHD_E_ARG;
    TRY(check_ctx(fine));
    if (!fine->have_state) return MHD_E_STATE;
    TRY(mhd_sync(fine));   // the state is complete before the closure's stream reads it
    ClSource src;
    memset(&src, 0, sizeof(src));
    const double* f64[5] = {fine->r, fine->t, fine->u, fine->v, fine->w};
    for (int q = 0; q < 5; ++q) src.p[q] = f64[q];
    src.p[5] = fine->bx;
    src.p[6] = fine->by;
    src.p[7] = fine->bz;
    for (int q = 5; q < 8; ++q) src.f32[q] = fine->wf32;
    if (fine->su) {
        src.mean[5] = fine->bxbar;
        src.mean[6] = fine->bybar;
        src.add[7] = fine->sc.bzbar;
    }
    return closure_add_source(cl, src, fine->nx, fine->ny, fine->nz, f, fine->p.f);
}

int mhd_global_averages(mhd_ctx_t* c, double out[4])
{
    TRY(check_ctx(c));
    return run_global_averages(c, out);
}

int mhd_nusselt(mhd_ctx_t* c, double* nu)
{
    TRY(check_ctx(c));
    return run_nusselt(c, nu);
}

int mhd_checkdivb(mhd_ctx_t* c, double out[2])
{
    TRY(check_ctx(c));
    return run_checkdivb(c, out);
}

int mhd_sminmax(mhd_ctx_t* c, int id, double out[2])
{
    TRY(check_ctx(c));
    void* p;
    int f32;
    long n;
    TRY(field_ref(c, id, 1, &p, &f32, &n));
    if (n != c->n3) return MHD_E_ARG;
    RedArgs a = rargs(c, RED_MINMAX);
    a.f0 = p;
    a.f_f32 = f32;
    if (c->su && id == MHD_F_BX) a.f0m = c->bxbar;
    if (c->su && id == MHD_F_BY) a.f0m = c->bybar;
    if (c->su && id == MHD_F_BZ) a.f0s = c->sc.bzbar;
    TRY(kerr(launch_reduce(a, c->s)));
    return fetch_result(c, out, 2);
}

int mhd_calcdt(mhd_ctx_t* c, double* dt)
{
    TRY(check_ctx(c));
    return run_calcdt(c, dt);
}

int mhd_op(mhd_ctx_t* c, int op, const double* in, double* out, double* out2, int flag)
{
    TRY(check_ctx(c));
    if (!in || !out) return MHD_E_ARG;
    const int wf = c->wf32;
    const long n3 = c->n3;
    void* A = c->wk[W_DXM];
    void* B = c->wk[W_DYM];
    void* C = c->wk[W_UXY];
    if (op == MHD_OP_D1BYDZ1D || op == MHD_OP_D2BYDZ1D) {
        // nz-vectors: run the 3-D kernel on a 1x1 column broadcast over the plane
        std::vector<double> col(n3);
        for (int k = 0; k < c->nz; ++k)
            for (int q = 0; q < c->nxy; ++q) col[(long)k * c->nxy + q] = in[k];
        TRY(upload(c, A, wf, col.data(), n3));
        ZLaunch Z = zbase(c);
        Z.job[0] = zjob(A, wf, B, wf, op == MHD_OP_D1BYDZ1D ? ZK_D1 : ZK_D2, flag);
        Z.njobs = 1;
        TRY(zrun(c, Z));
        TRY(download(c, col.data(), B, wf, n3));
        for (int k = 0; k < c->nz; ++k) out[k] = col[(long)k * c->nxy];
        return 0;
    }
    TRY(upload(c, A, wf, in, n3));
    switch (op) {
    case MHD_OP_D1BYDX:
    case MHD_OP_D1BYDY: {
        SpecLaunch L = spec_base(c, op == MHD_OP_D1BYDX ? 0 : 1, SM_DERIV1);
        L.job[0] = sjob(A, wf, B, nullptr, MUL_D1, 0, wf);
        L.njobs = 1;
        TRY(spec(c, L));
        return download(c, out, B, wf, n3);
    }
    case MHD_OP_D2BYDX:
    case MHD_OP_D2BYDY: {
        if (!out2) return MHD_E_ARG;
        SpecLaunch L = spec_base(c, op == MHD_OP_D2BYDX ? 0 : 1, SM_DERIV2);
        L.job[0] = sjob(A, wf, B, C, MUL_D1, MUL_D2, wf);
        L.njobs = 1;
        TRY(spec(c, L));
        TRY(download(c, out, B, wf, n3));
        return download(c, out2, C, wf, n3);
    }
    case MHD_OP_DEALIAS: {
        for (int dir = 0; dir < 2; ++dir) {
            SpecLaunch L = spec_base(c, dir, SM_DERIV1);
            L.job[0] = sjob(A, wf, A, nullptr, MUL_MASK, 0, wf);
            L.njobs = 1;
            TRY(spec(c, L));
        }
        return download(c, out, A, wf, n3);
    }
    case MHD_OP_FORWARDFFT:
    case MHD_OP_INVERSEFFT: {
        const int mode = op == MHD_OP_FORWARDFFT ? SM_R2HC : SM_HC2R;
        for (int dir = 0; dir < 2; ++dir) {
            SpecLaunch L = spec_base(c, dir, mode);
            L.job[0] = sjob(A, wf, A, nullptr, 0, 0, wf);
            L.njobs = 1;
            TRY(spec(c, L));
        }
        return download(c, out, A, wf, n3);
    }
    case MHD_OP_D1BYDZ:
    case MHD_OP_D1UPBYDZ:
    case MHD_OP_D2BYDZ: {
        ZLaunch Z = zbase(c);
        const int kind = op == MHD_OP_D1BYDZ ? ZK_D1 : op == MHD_OP_D2BYDZ ? ZK_D2 : ZK_D1UP;
        Z.job[0] = zjob(A, wf, B, wf, kind, flag);
        Z.njobs = 1;
        TRY(zrun(c, Z));
        return download(c, out, B, wf, n3);
    }
    }
    return MHD_E_ARG;
}

int mhd_kernel_timing(mhd_ctx_t* c, int kernel, int enable)
{
    TRY(check_ctx(c));
    MHD_HIP_CHECK(hipStreamSynchronize(c->s));
    c->nev = 0;
    c->timed_kernel = enable ? kernel : MHD_K_NONE;
    return 0;
}

int mhd_kernel_time(mhd_ctx_t* c, double* total_ms, long* launches)
{
    TRY(check_ctx(c));
    MHD_HIP_CHECK(hipStreamSynchronize(c->s));
    double tot = 0.0;
    for (int i = 0; i + 1 < c->nev; i += 2) {
        float ms = 0.f;
        MHD_HIP_CHECK(hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = c->nev / 2;
    return 0;
}

int mhd_sync(mhd_ctx_t* c)
{
    TRY(check_ctx(c));
    MHD_HIP_CHECK(hipStreamSynchronize(c->s));
    MHD_HIP_CHECK(hipGetLastError());
    return 0;
}

long mhd_device_bytes(mhd_ctx_t* c) { return c ? (long)c->pool_bytes : 0; }

}  // extern "C"
