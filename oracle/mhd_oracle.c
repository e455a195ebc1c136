/*
 * mhd_oracle.c -- CPU oracle (FP64, serial) for the per-timestep core of the
 * reference compressible magnetoconvection solver (p3d "VarElecCond").
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker / CPU baseline.  The product
 * library never links or calls this file.
 *
 * Every routine restates one reference routine; the citation above each one
 * gives the file:line it follows (paths relative to VarElecCond/).  Loop
 * bodies keep the reference's operation order so that rounding matches the
 * reference wherever the FFT does not intervene.
 *
 * The reference calls FFTW 2.1.5 (rfftw_f77, not vendored, not available
 * here) for its real transforms.  This file implements the same transform
 * (unnormalised real<->halfcomplex, FFTW-2 storage order
 * [r0, r1..r_{n/2}, i_{n/2-1}..i_1]) with its own mixed-radix FFT; it agrees
 * with an O(n^2) long-double DFT to ~1e-15 (tests/test_oracle_fft.py).
 */
#include "mhd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* OpenMP over planes / points for the CPU baseline.  Only loops whose
 * iterations write disjoint elements are parallel; every reduction keeps its
 * serial order, so results do not depend on the thread count. */
#if defined(_OPENMP)
#define ORC_PAR _Pragma("omp parallel for schedule(static)")
#else
#define ORC_PAR
#endif

/* ------------------------------------------------------------------------ */
/* complex helpers and a mixed-radix (4,2,3,5,generic) DIT FFT              */
/* ------------------------------------------------------------------------ */
typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx a, cplx b)
{
    cplx r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
    return r;
}

typedef struct fftplan {
    int n;
    int nf;
    int fac[40];
    cplx* tw;      /* tw[t] = exp(-2 pi i t / n) */
    cplx* buf_in;  /* length n */
    cplx* buf_out; /* length n */
    cplx* tmp;     /* length max radix */
} fftplan;

#define ORC_MAXPLANS 32
static fftplan* g_plans[ORC_MAXPLANS];
static int g_nplans = 0;

static fftplan* get_plan(int n)
{
    for (int i = 0; i < g_nplans; i++)
        if (g_plans[i]->n == n) return g_plans[i];
    fftplan* p = (fftplan*)calloc(1, sizeof(fftplan));
    p->n = n;
    int m = n, nf = 0;
    while (m % 4 == 0) { p->fac[nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[nf++] = 2; m /= 2; }
    while (m % 3 == 0) { p->fac[nf++] = 3; m /= 3; }
    while (m % 5 == 0) { p->fac[nf++] = 5; m /= 5; }
    for (int q = 7; m > 1; q += 2)
        while (m % q == 0) { p->fac[nf++] = q; m /= q; }
    p->nf = nf;
    p->tw = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    for (int t = 0; t < n; t++) {
        /* angles in long double, rounded once to double */
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)t
                        / (long double)n;
        p->tw[t].re = (double)cosl(a);
        p->tw[t].im = (double)sinl(a);
    }
    int maxr = 2;
    for (int i = 0; i < nf; i++) if (p->fac[i] > maxr) maxr = p->fac[i];
    p->buf_in = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    p->buf_out = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    p->tmp = (cplx*)malloc(sizeof(cplx) * (size_t)maxr);
    if (g_nplans < ORC_MAXPLANS) g_plans[g_nplans++] = p;
    return p;
}

/* twiddle W_N^e, forward (sign<0) or backward (sign>0) */
static inline cplx twid(const fftplan* p, long e, int sign)
{
    cplx w = p->tw[e % p->n];
    if (sign > 0) w.im = -w.im;
    return w;
}

/* out[k + m*q] = sum_s t[s] * W_r^{s q}, q = 0..r-1 */
static void small_dft(const fftplan* p, int r, cplx* t, cplx* out, int k, int m, int sign)
{
    const double sg = (sign < 0) ? -1.0 : 1.0;
    if (r == 2) {
        cplx a = t[0], b = t[1];
        out[k].re = a.re + b.re; out[k].im = a.im + b.im;
        out[k + m].re = a.re - b.re; out[k + m].im = a.im - b.im;
    } else if (r == 4) {
        cplx a = { t[0].re + t[2].re, t[0].im + t[2].im };
        cplx b = { t[0].re - t[2].re, t[0].im - t[2].im };
        cplx cc = { t[1].re + t[3].re, t[1].im + t[3].im };
        cplx d = { t[1].re - t[3].re, t[1].im - t[3].im };
        /* (sg*i)*d = (-sg*d.im, sg*d.re) */
        cplx id = { -sg * d.im, sg * d.re };
        out[k].re = a.re + cc.re;          out[k].im = a.im + cc.im;
        out[k + m].re = b.re + id.re;      out[k + m].im = b.im + id.im;
        out[k + 2 * m].re = a.re - cc.re;  out[k + 2 * m].im = a.im - cc.im;
        out[k + 3 * m].re = b.re - id.re;  out[k + 3 * m].im = b.im - id.im;
    } else if (r == 3) {
        const double h = 0.86602540378443864676372317075294; /* sqrt(3)/2 */
        cplx s = { t[1].re + t[2].re, t[1].im + t[2].im };
        cplx d = { t[1].re - t[2].re, t[1].im - t[2].im };
        cplx base = { t[0].re - 0.5 * s.re, t[0].im - 0.5 * s.im };
        cplx rot = { -sg * h * d.im, sg * h * d.re };
        out[k].re = t[0].re + s.re;       out[k].im = t[0].im + s.im;
        out[k + m].re = base.re + rot.re; out[k + m].im = base.im + rot.im;
        out[k + 2 * m].re = base.re - rot.re; out[k + 2 * m].im = base.im - rot.im;
    } else if (r == 5) {
        const double c1 = 0.30901699437494742410229341718282;  /* cos(2pi/5) */
        const double c2 = -0.80901699437494742410229341718282; /* cos(4pi/5) */
        const double s1 = 0.95105651629515357211643933337938;  /* sin(2pi/5) */
        const double s2 = 0.58778525229247312916870595463907;  /* sin(4pi/5) */
        cplx a1 = { t[1].re + t[4].re, t[1].im + t[4].im };
        cplx a2 = { t[2].re + t[3].re, t[2].im + t[3].im };
        cplx b1 = { t[1].re - t[4].re, t[1].im - t[4].im };
        cplx b2 = { t[2].re - t[3].re, t[2].im - t[3].im };
        cplx p1 = { t[0].re + c1 * a1.re + c2 * a2.re, t[0].im + c1 * a1.im + c2 * a2.im };
        cplx p2 = { t[0].re + c2 * a1.re + c1 * a2.re, t[0].im + c2 * a1.im + c1 * a2.im };
        cplx q1 = { s1 * b1.re + s2 * b2.re, s1 * b1.im + s2 * b2.im };
        cplx q2 = { s2 * b1.re - s1 * b2.re, s2 * b1.im - s1 * b2.im };
        cplx iq1 = { -sg * q1.im, sg * q1.re };
        cplx iq2 = { -sg * q2.im, sg * q2.re };
        out[k].re = t[0].re + a1.re + a2.re; out[k].im = t[0].im + a1.im + a2.im;
        out[k + m].re = p1.re + iq1.re;      out[k + m].im = p1.im + iq1.im;
        out[k + 4 * m].re = p1.re - iq1.re;  out[k + 4 * m].im = p1.im - iq1.im;
        out[k + 2 * m].re = p2.re + iq2.re;  out[k + 2 * m].im = p2.im + iq2.im;
        out[k + 3 * m].re = p2.re - iq2.re;  out[k + 3 * m].im = p2.im - iq2.im;
    } else {
        const long stride = p->n / r;
        for (int q = 0; q < r; q++) {
            cplx acc = { 0.0, 0.0 };
            for (int s = 0; s < r; s++) {
                cplx w = twid(p, (long)((s * q) % r) * stride, sign);
                cplx v = cmul(t[s], w);
                acc.re += v.re; acc.im += v.im;
            }
            out[k + m * q] = acc;
        }
    }
}

/* per-thread scratch of the transforms and of the row operators */
static _Thread_local cplx *tl_in, *tl_out, *tl_tmp;
static _Thread_local double *tl_row, *tl_row2;
static _Thread_local size_t tl_cap;
static void tl_ensure(size_t n)
{
    if (n <= tl_cap) return;
    tl_in = (cplx*)realloc(tl_in, sizeof(cplx) * n);
    tl_out = (cplx*)realloc(tl_out, sizeof(cplx) * n);
    tl_tmp = (cplx*)realloc(tl_tmp, sizeof(cplx) * n);
    tl_row = (double*)realloc(tl_row, sizeof(double) * n);
    tl_row2 = (double*)realloc(tl_row2, sizeof(double) * n);
    tl_cap = n;
}

static double* orc_row(size_t n, int which)
{
    tl_ensure(n);
    return which ? tl_row2 : tl_row;
}

static void fft_rec(const fftplan* p, int n, const cplx* in, long is, cplx* out,
                    int level, int sign)
{
    if (n == 1) { out[0] = in[0]; return; }
    const int r = p->fac[level];
    const int m = n / r;
    for (int s = 0; s < r; s++)
        fft_rec(p, m, in + (long)s * is, is * r, out + (long)s * m, level + 1, sign);
    const long twstride = p->n / n;
    cplx* t = tl_tmp;
    for (int k = 0; k < m; k++) {
        for (int s = 0; s < r; s++) {
            cplx y = out[(long)s * m + k];
            t[s] = (s == 0 || k == 0) ? y : cmul(y, twid(p, (long)s * k * twstride, sign));
        }
        small_dft(p, r, t, out, k, m, sign);
    }
}

/* unnormalised complex DFT, sign -1 forward, +1 backward */
static void cfft(fftplan* p, const cplx* in, cplx* out, int sign)
{
    fft_rec(p, p->n, in, 1, out, 0, sign);
}

/* FFTW-2 rfftw REAL_TO_COMPLEX: halfcomplex [r0,r1..r_{n/2},i_{(n+1)/2-1}..i_1] */
static void r2hc_strided(fftplan* p, const double* in, long istride, double* out, long ostride)
{
    const int n = p->n;
    tl_ensure((size_t)n);
    cplx *bi = tl_in, *bo = tl_out;
    for (int t = 0; t < n; t++) { bi[t].re = in[t * istride]; bi[t].im = 0.0; }
    cfft(p, bi, bo, -1);
    out[0] = bo[0].re;
    for (int m = 1; m <= n / 2; m++) out[m * ostride] = bo[m].re;
    for (int m = 1; m < (n + 1) / 2; m++) out[(n - m) * ostride] = bo[m].im;
}

/* FFTW-2 rfftw COMPLEX_TO_REAL (unnormalised) */
static void hc2r_strided(fftplan* p, const double* in, long istride, double* out, long ostride)
{
    const int n = p->n;
    tl_ensure((size_t)n);
    cplx *bi = tl_in, *bo = tl_out;
    bi[0].re = in[0]; bi[0].im = 0.0;
    for (int m = 1; m < (n + 1) / 2; m++) {
        bi[m].re = in[m * istride];
        bi[m].im = in[(n - m) * istride];
        bi[n - m].re = bi[m].re;
        bi[n - m].im = -bi[m].im;
    }
    if (n % 2 == 0) { bi[n / 2].re = in[(n / 2) * istride]; bi[n / 2].im = 0.0; }
    cfft(p, bi, bo, +1);
    for (int t = 0; t < n; t++) out[t * ostride] = bo[t].re;
}

void orc_r2hc(int n, const double* in, double* out) { r2hc_strided(get_plan(n), in, 1, out, 1); }
void orc_hc2r(int n, const double* in, double* out) { hc2r_strided(get_plan(n), in, 1, out, 1); }

/* ------------------------------------------------------------------------ */
/* context                                                                  */
/* ------------------------------------------------------------------------ */
struct orc_ctx {
    int nx, ny, nz;
    size_t nxy, n3;
    orc_params p;
    /* src/constants.f90:56-109 */
    double dx, dy, dz, dd, xn, xzn, taylt, cpsi, spsi, cor;
    double gck, sck, zck, scko3, sckg, fzckg, thmplus1, fo2, g1, gm1, nusstmp;
    double dtvisc0x, dtvisc0z, dtdiff0x, dtdiff0z;
    /* src/dbydz.F90:355-362 */
    double ho12, sc1;
    /* src/fft.F90:171-221 */
    int icutoffx, icutoffy;
    double *akx1, *akx2, *aky1, *aky2;
    fftplan *px, *py;
    double *rowbuf, *rowbuf2;
    /* src/arrays.f90:14-74 */
    double* fld[ORC_NFIELDS];
    double bzbar;
    /* driver state, src/main.F90:84-116 */
    double dt, dt1, dt2, ttime;
    int nit, iter, first, second;
    /* learned subgrid closure (closure_oracle.c; no counterpart in the
     * reference): parameters, scaling and the 12 fluxes of the step */
    double* cl_params;
    double cl_norm[40];
    double* cl_flux;
    int cl_fixed;      /* fluxes given by orc_closure_fix_fluxes, not evaluated */
};

static const double third = 1.0 / 3.0; /* src/constants.f90:15 */
static const double elf = 1.0 / 11.0;  /* src/constants.f90:16 */

#define IDX(c, i, j, k) ((size_t)(i) + (size_t)(c)->nx * ((size_t)(j) + (size_t)(c)->ny * (size_t)(k)))

void orc_ra_q_to_ck_f(double gamma, double sigma, double theta, double cm, double zeta,
                      double ra, double q, double* ck, double* f)
{
    /* src/initial.F90:183-195 (flag 2) */
    double rayl = ra, chand = q;
    double c = sqrt((((cm + 1.0) * theta * theta) / (sigma * rayl * gamma))
                    * (cm + 1.0 - (cm * gamma))
                    * pow(1.0 + (0.5 * theta), (2.0 * cm) - 1.0));
    *ck = c;
    *f = chand * (zeta * sigma * c * c);
}

static void set_constants(orc_ctx* c)
{
    /* src/constants.f90:56-109 */
    const orc_params* p = &c->p;
    c->dx = p->xmax / (double)c->nx;
    c->dy = p->ymax / (double)c->ny;
    c->dz = 1.0 / (double)(c->nz - 1);
    c->dd = fmin(c->dx, c->dy);
    c->xn = 1.0 / (double)(c->nx * c->ny);
    c->xzn = c->dx * c->dy * c->dz;
    c->taylt = p->tayl / pow(1.0 + p->theta / 2.0, 2.0 * p->cm);
    c->cpsi = cos(p->psi);
    c->spsi = sin(p->psi);
    c->cor = p->ck * p->sigma * sqrt(c->taylt);
    c->gck = p->gamma * p->ck;
    c->sck = p->sigma * p->ck;
    c->zck = p->zeta * p->ck;
    c->scko3 = c->sck / 3.0;
    c->sckg = p->sigma * p->ck * (p->gamma - 1.0);
    c->fzckg = p->f * p->zeta * p->ck * (p->gamma - 1.0);
    c->thmplus1 = p->theta * (p->cm + 1.0);
    c->fo2 = p->f / 2.0;
    c->g1 = sqrt(p->gamma);
    c->gm1 = p->gamma - 1.0;
    c->nusstmp = p->theta * (p->cm + 1.0) * (1.0 - (1.0 / p->gamma));
    c->dtvisc0x = 0.1 * 0.75 * c->dd * c->dd / p->sigma / p->ck;
    c->dtvisc0z = 0.2 * c->dz * c->dz / p->sigma / p->ck * (1.0 + p->theta / 7.0);
    c->dtdiff0x = 0.1 * c->dd * c->dd / p->gamma / p->ck;
    c->dtdiff0z = 0.2 * c->dz * c->dz / p->gamma / p->ck * (1.0 + p->theta / 7.0);
    /* src/dbydz.F90:355-362 */
    c->ho12 = (double)(c->nz - 1) / 12.0;
    c->sc1 = (double)((c->nz - 1) * (c->nz - 1)) / 12.0;
}

static void init_fft(orc_ctx* c)
{
    /* src/fft.F90:171-221 */
    const int nx = c->nx, ny = c->ny;
    c->icutoffx = (nx / 2) / 3;
    c->icutoffy = (ny / 2) / 3;
    const double cx = 8.0 * atan(1.0) / c->p.xmax;
    const double cy = 8.0 * atan(1.0) / c->p.ymax;
    double *akx1 = c->akx1, *akx2 = c->akx2, *aky1 = c->aky1, *aky2 = c->aky2;
    for (int i = 1; i <= nx / 2 - 1; i++) {
        akx1[i] = (double)i * cx;
        akx1[nx - i] = (double)i * cx;
        akx2[i] = -((double)i * (double)i * cx * cx);
        akx2[nx - i] = -((double)i * (double)i * cx * cx);
    }
    akx2[nx / 2] = -((double)(nx / 2) * (double)(nx / 2) * cx * cx);
    for (int i = 1; i <= ny / 2 - 1; i++) {
        aky1[i] = (double)i * cy;
        aky1[ny - i] = (double)i * cy;
        aky2[i] = -((double)i * (double)i * cy * cy);
        aky2[ny - i] = -((double)i * (double)i * cy * cy);
    }
    aky2[ny / 2] = -((double)(ny / 2) * (double)(ny / 2) * cy * cy);
    c->px = get_plan(nx);
    c->py = get_plan(ny);
}

orc_ctx* orc_create(const orc_params* p, int nx, int ny, int nz, int* ier)
{
    *ier = 0;
    /* src/fft.F90:112-119 */
    if (2 * (nx / 2) != nx) { *ier = -1; return NULL; }
    if (2 * (ny / 2) != ny) { *ier = -2; return NULL; }
    /* src/main.F90:216-219 */
    if (nz < 6) { *ier = 1; return NULL; }
    orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    if (!c) { *ier = -5; return NULL; }
    c->nx = nx; c->ny = ny; c->nz = nz;
    c->nxy = (size_t)nx * ny;
    c->n3 = c->nxy * nz;
    c->p = *p;
    /* src/initial.F90:305-318: F rescaled for the dynamo and kinematic problems */
    if (c->p.opts & ORC_OPT_DYNAMO) c->p.f = 1.0;
    if (c->p.opts & ORC_OPT_KINEMATIC) c->p.f = 0.0;
    for (int id = 0; id < ORC_NFIELDS; id++) {
        size_t len;
        if (id >= ORC_UDOT && id <= ORC_TORDOT) len = 3 * c->n3;
        else if (id == ORC_BXBARDOT || id == ORC_BYBARDOT) len = 3 * (size_t)nz;
        else if (id >= ORC_BXBAR && id <= ORC_DKAPPA) len = (size_t)nz;
        else if (id == ORC_AKX1 || id == ORC_AKX2) len = (size_t)nx;
        else if (id == ORC_AKY1 || id == ORC_AKY2) len = (size_t)ny;
        else if (id == ORC_ETAR) len = (size_t)nz;
        else len = c->n3;
        c->fld[id] = (double*)calloc(len, sizeof(double));
        if (!c->fld[id]) { *ier = -5; orc_destroy(c); return NULL; }
    }
    c->akx1 = c->fld[ORC_AKX1]; c->akx2 = c->fld[ORC_AKX2];
    c->aky1 = c->fld[ORC_AKY1]; c->aky2 = c->fld[ORC_AKY2];
    int nmax = nx > ny ? nx : ny;
    c->rowbuf = (double*)calloc((size_t)nmax, sizeof(double));
    c->rowbuf2 = (double*)calloc((size_t)nmax, sizeof(double));
    set_constants(c);
    init_fft(c);
    c->first = 1; c->second = 0;
    for (int k = 0; k < nz; k++) c->fld[ORC_ETAR][k] = 1.0;
    if (c->p.eta_c > 1.0) orc_ini_trans_params(c);
    return c;
}

void orc_destroy(orc_ctx* c)
{
    if (!c) return;
    for (int id = 0; id < ORC_NFIELDS; id++) free(c->fld[id]);
    free(c->rowbuf); free(c->rowbuf2);
    free(c->cl_params); free(c->cl_flux);
    free(c);
}

/* Couple a closure network to the step (params NULL detaches): every
 * hydro_deriv then evaluates the network on the state (bf16 rounding as on
 * the GPU) and subtracts div(tau)/r from the momentum right-hand sides and
 * div(q) from T's (x, y spectral, z with the flag-5 stencils); mag_deriv adds
 * eps to u x B. */
int orc_closure_attach(orc_ctx* c, const double* params, const double norm[40])
{
    free(c->cl_params);
    free(c->cl_flux);
    c->cl_params = NULL;
    c->cl_flux = NULL;
    c->cl_fixed = 0;
    if (!params) return 0;
    const int np = orc_closure_nparam();
    c->cl_params = (double*)malloc(sizeof(double) * (size_t)np);
    c->cl_flux = (double*)calloc(12 * c->n3, sizeof(double));
    if (!c->cl_params || !c->cl_flux) return -5;
    memcpy(c->cl_params, params, sizeof(double) * (size_t)np);
    memcpy(c->cl_norm, norm, sizeof(c->cl_norm));
    return 0;
}

/* The same coupling with the 12 fluxes held fixed (checks the coupling on its
 * own, with the fluxes a GPU network produced); NULL returns to no closure. */
int orc_closure_fix_fluxes(orc_ctx* c, const double* flux)
{
    orc_closure_attach(c, NULL, NULL);
    if (!flux) return 0;
    c->cl_flux = (double*)malloc(sizeof(double) * 12 * c->n3);
    if (!c->cl_flux) return -5;
    memcpy(c->cl_flux, flux, sizeof(double) * 12 * c->n3);
    c->cl_fixed = 1;
    return 0;
}

static int closure_tend(orc_ctx* c)
{
    const size_t n3 = c->n3;
    double* f[8] = {c->fld[ORC_R], c->fld[ORC_T], c->fld[ORC_U], c->fld[ORC_V],
                    c->fld[ORC_W], c->fld[ORC_BX], c->fld[ORC_BY], c->fld[ORC_BZ]};
    double* out[12];
    for (int o = 0; o < 12; o++) out[o] = c->cl_flux + (size_t)o * n3;
    if (!c->cl_fixed)
        orc_mlp_forward(c->nx, c->ny, c->nz, (const double* const*)f, c->cl_params, c->cl_norm, 1, out);
    double* acc = (double*)malloc(sizeof(double) * n3);
    double* d = (double*)malloc(sizeof(double) * n3);
    if (!acc || !d) { free(acc); free(d); return -5; }
    /* u: tau_xx, tau_xy, tau_xz; v: tau_xy, tau_yy, tau_yz; w: tau_xz, tau_yz,
     * tau_zz; T: q_x, q_y, q_z (flux index along x, y, z) */
    static const int ix[4] = {0, 3, 4, 9}, iy[4] = {3, 1, 5, 10}, iz[4] = {4, 5, 2, 11};
    double* dots[4] = {c->fld[ORC_UDOT] + 2 * n3, c->fld[ORC_VDOT] + 2 * n3, c->fld[ORC_WDOT] + 2 * n3,
                       c->fld[ORC_TDOT] + 2 * n3};
    const double* r = c->fld[ORC_R];
    int e = 0;
    for (int m = 0; m < 4; m++) {
        orc_d1bydx(c, out[ix[m]], acc);
        orc_d1bydy(c, out[iy[m]], d);
        for (size_t q = 0; q < n3; q++) acc[q] = acc[q] + d[q];
        e |= orc_d1bydz(c, out[iz[m]], d, 5);
        for (size_t q = 0; q < n3; q++) {
            const double div = acc[q] + d[q];
            dots[m][q] = m < 3 ? dots[m][q] - div / r[q] : dots[m][q] - div;
        }
    }
    free(acc);
    free(d);
    return e ? 7 : 0;
}

double* orc_field(orc_ctx* c, int id) { return (id >= 0 && id < ORC_NFIELDS) ? c->fld[id] : NULL; }

double orc_scalar(orc_ctx* c, int id)
{
    switch (id) {
    case ORC_S_DT: return c->dt;
    case ORC_S_DT1: return c->dt1;
    case ORC_S_DT2: return c->dt2;
    case ORC_S_TTIME: return c->ttime;
    case ORC_S_BZBAR: return c->bzbar;
    case ORC_S_NIT: return c->nit;
    case ORC_S_ITER: return c->iter;
    case ORC_S_FIRST: return c->first;
    case ORC_S_SECOND: return c->second;
    case ORC_S_GCK: return c->gck;
    case ORC_S_SCK: return c->sck;
    case ORC_S_ZCK: return c->zck;
    case ORC_S_DZ: return c->dz;
    case ORC_S_XN: return c->xn;
    case ORC_S_XZN: return c->xzn;
    case ORC_S_NUSSTMP: return c->nusstmp;
    }
    return NAN;
}

void orc_set_scalar(orc_ctx* c, int id, double v)
{
    switch (id) {
    case ORC_S_DT: c->dt = v; break;
    case ORC_S_DT1: c->dt1 = v; break;
    case ORC_S_DT2: c->dt2 = v; break;
    case ORC_S_TTIME: c->ttime = v; break;
    case ORC_S_BZBAR: c->bzbar = v; break;
    case ORC_S_NIT: c->nit = (int)v; break;
    case ORC_S_ITER: c->iter = (int)v; break;
    case ORC_S_FIRST: c->first = (int)v; break;
    case ORC_S_SECOND: c->second = (int)v; break;
    }
}

/* ------------------------------------------------------------------------ */
/* horizontal spectral operators, src/fft.F90                               */
/* ------------------------------------------------------------------------ */

/* src/fft.F90:255-324 */
void orc_d1bydx(orc_ctx* c, const double* fin, double* fout)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    ORC_PAR
    for (int k = 0; k < nz; k++) {
        double* t = orc_row((size_t)(nx > ny ? nx : ny), 0);
        double* s = orc_row((size_t)(nx > ny ? nx : ny), 1);
        (void)t;
        (void)s;
        for (int j = 0; j < ny; j++) {
            double* o = fout + IDX(c, 0, j, k);
            r2hc_strided(c->px, fin + IDX(c, 0, j, k), 1, o, 1);
            o[0] = 0.0;
            o[nx / 2] = 0.0;
            for (int i = 1; i <= nx / 2 - 1; i++) {
                double fac = o[i];
                o[i] = -o[nx - i] * c->akx1[i];
                o[nx - i] = fac * c->akx1[nx - i];
            }
            hc2r_strided(c->px, o, 1, t, 1);
            for (int i = 0; i < nx; i++) o[i] = t[i] / (double)nx;
        }
    }
}

/* src/fft.F90:329-398 */
void orc_d1bydy(orc_ctx* c, const double* fin, double* fout)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    ORC_PAR
    for (int k = 0; k < nz; k++) {
        double* t = orc_row((size_t)(nx > ny ? nx : ny), 0);
        double* s = orc_row((size_t)(nx > ny ? nx : ny), 1);
        (void)t;
        (void)s;
        for (int i = 0; i < nx; i++) {
            double* o = fout + IDX(c, i, 0, k);
            r2hc_strided(c->py, fin + IDX(c, i, 0, k), nx, o, nx);
            o[0] = 0.0;
            o[(size_t)(ny / 2) * nx] = 0.0;
            for (int j = 1; j <= ny / 2 - 1; j++) {
                double fac = o[(size_t)j * nx];
                o[(size_t)j * nx] = -o[(size_t)(ny - j) * nx] * c->aky1[j];
                o[(size_t)(ny - j) * nx] = fac * c->aky1[ny - j];
            }
            hc2r_strided(c->py, o, nx, t, 1);
            for (int j = 0; j < ny; j++) o[(size_t)j * nx] = t[j] / (double)ny;
        }
    }
}

/* src/fft.F90:403-489 */
void orc_d2bydx(orc_ctx* c, const double* fin, double* f1, double* f2)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    ORC_PAR
    for (int k = 0; k < nz; k++) {
        double* t = orc_row((size_t)(nx > ny ? nx : ny), 0);
        double* s = orc_row((size_t)(nx > ny ? nx : ny), 1);
        (void)t;
        (void)s;
        for (int j = 0; j < ny; j++) {
            double* a = f1 + IDX(c, 0, j, k);
            double* b = f2 + IDX(c, 0, j, k);
            r2hc_strided(c->px, fin + IDX(c, 0, j, k), 1, b, 1);
            a[0] = 0.0;
            a[nx / 2] = 0.0;
            b[0] = 0.0;
            for (int i = 1; i <= nx / 2 - 1; i++) {
                a[i] = -b[nx - i] * c->akx1[i];
                a[nx - i] = b[i] * c->akx1[nx - i];
                b[i] = b[i] * c->akx2[i];
                b[nx - i] = b[nx - i] * c->akx2[nx - i];
            }
            b[nx / 2] = b[nx / 2] * c->akx2[nx / 2];
            hc2r_strided(c->px, a, 1, t, 1);
            for (int i = 0; i < nx; i++) a[i] = t[i] / (double)nx;
            hc2r_strided(c->px, b, 1, t, 1);
            for (int i = 0; i < nx; i++) b[i] = t[i] / (double)nx;
        }
    }
}

/* src/fft.F90:495-581 */
void orc_d2bydy(orc_ctx* c, const double* fin, double* f1, double* f2)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    ORC_PAR
    for (int k = 0; k < nz; k++) {
        double* t = orc_row((size_t)(nx > ny ? nx : ny), 0);
        double* s = orc_row((size_t)(nx > ny ? nx : ny), 1);
        (void)t;
        (void)s;
        for (int i = 0; i < nx; i++) {
            double* a = f1 + IDX(c, i, 0, k);
            double* b = f2 + IDX(c, i, 0, k);
            r2hc_strided(c->py, fin + IDX(c, i, 0, k), nx, s, 1);
            for (int j = 1; j <= ny / 2 - 1; j++) {
                a[(size_t)j * nx] = -s[ny - j] * c->aky1[j];
                a[(size_t)(ny - j) * nx] = s[j] * c->aky1[ny - j];
                b[(size_t)j * nx] = s[j] * c->aky2[j];
                b[(size_t)(ny - j) * nx] = s[ny - j] * c->aky2[ny - j];
            }
            a[0] = 0.0;
            b[0] = 0.0;
            a[(size_t)(ny / 2) * nx] = 0.0;
            b[(size_t)(ny / 2) * nx] = s[ny / 2] * c->aky2[ny / 2];
            hc2r_strided(c->py, a, nx, t, 1);
            for (int j = 0; j < ny; j++) a[(size_t)j * nx] = t[j] / (double)ny;
            hc2r_strided(c->py, b, nx, t, 1);
            for (int j = 0; j < ny; j++) b[(size_t)j * nx] = t[j] / (double)ny;
        }
    }
}

/* src/fft.F90:586-682 */
void orc_dealias(orc_ctx* c, double* z)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    ORC_PAR
    for (int k = 0; k < nz; k++) {
        double* t = orc_row((size_t)(nx > ny ? nx : ny), 0);
        double* s = orc_row((size_t)(nx > ny ? nx : ny), 1);
        (void)t;
        (void)s;
        for (int j = 0; j < ny; j++) {
            double* o = z + IDX(c, 0, j, k);
            r2hc_strided(c->px, o, 1, t, 1);
            memcpy(o, t, sizeof(double) * (size_t)nx);
            for (int i = 1; i <= c->icutoffx; i++) { o[nx / 2 + i] = 0.0; o[nx / 2 - i] = 0.0; }
            o[nx / 2] = 0.0;
            hc2r_strided(c->px, o, 1, t, 1);
            for (int i = 0; i < nx; i++) o[i] = t[i] / (double)nx;
        }
    }
    ORC_PAR
    for (int k = 0; k < nz; k++) {
        double* t = orc_row((size_t)(nx > ny ? nx : ny), 0);
        double* s = orc_row((size_t)(nx > ny ? nx : ny), 1);
        (void)t;
        (void)s;
        for (int i = 0; i < nx; i++) {
            double* o = z + IDX(c, i, 0, k);
            r2hc_strided(c->py, o, nx, t, 1);
            for (int j = 0; j < ny; j++) o[(size_t)j * nx] = t[j];
            o[(size_t)(ny / 2) * nx] = 0.0;
            for (int j = 1; j <= c->icutoffy; j++) {
                o[(size_t)(ny / 2 + j) * nx] = 0.0;
                o[(size_t)(ny / 2 - j) * nx] = 0.0;
            }
            hc2r_strided(c->py, o, nx, t, 1);
            for (int j = 0; j < ny; j++) o[(size_t)j * nx] = t[j] / (double)ny;
        }
    }
}

/* src/fft.F90:687-735 */
void orc_forwardfft(orc_ctx* c, const double* fin, double* fout)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    ORC_PAR
    for (int k = 0; k < nz; k++) {
        double* t = orc_row((size_t)(nx > ny ? nx : ny), 0);
        double* s = orc_row((size_t)(nx > ny ? nx : ny), 1);
        (void)t;
        (void)s;
        for (int j = 0; j < ny; j++)
            r2hc_strided(c->px, fin + IDX(c, 0, j, k), 1, fout + IDX(c, 0, j, k), 1);
    }
    ORC_PAR
    for (int k = 0; k < nz; k++) {
        double* t = orc_row((size_t)(nx > ny ? nx : ny), 0);
        double* s = orc_row((size_t)(nx > ny ? nx : ny), 1);
        (void)t;
        (void)s;
        for (int i = 0; i < nx; i++) {
            double* o = fout + IDX(c, i, 0, k);
            r2hc_strided(c->py, o, nx, t, 1);
            for (int j = 0; j < ny; j++) o[(size_t)j * nx] = t[j];
        }
    }
}

/* src/fft.F90:740-799 */
void orc_inversefft(orc_ctx* c, const double* fin, double* fout)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    if (fout != fin) memcpy(fout, fin, sizeof(double) * c->n3);
    ORC_PAR
    for (int k = 0; k < nz; k++) {
        double* t = orc_row((size_t)(nx > ny ? nx : ny), 0);
        double* s = orc_row((size_t)(nx > ny ? nx : ny), 1);
        (void)t;
        (void)s;
        for (int j = 0; j < ny; j++) {
            double* o = fout + IDX(c, 0, j, k);
            hc2r_strided(c->px, o, 1, t, 1);
            for (int i = 0; i < nx; i++) o[i] = t[i] / (double)nx;
        }
    }
    ORC_PAR
    for (int k = 0; k < nz; k++) {
        double* t = orc_row((size_t)(nx > ny ? nx : ny), 0);
        double* s = orc_row((size_t)(nx > ny ? nx : ny), 1);
        (void)t;
        (void)s;
        for (int i = 0; i < nx; i++) {
            double* o = fout + IDX(c, i, 0, k);
            hc2r_strided(c->py, o, nx, t, 1);
            for (int j = 0; j < ny; j++) o[(size_t)j * nx] = t[j] / (double)ny;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* vertical finite differences, src/dbydz.F90 (serial branches)             */
/* ------------------------------------------------------------------------ */
/* Each row routine below is one reference stencil subroutine applied to one
 * point; the 3-D drivers apply them plane by plane. */

/* topdz, src/dbydz.F90:927-998 */
static int topdz(double ho12, double* o, double f1, double f2, double f3, double f4, int fl)
{
    if (fl == 0 || fl == 2) *o = 0.0;
    else if (fl == 1) *o = ho12 * (14.0 * f1 + 36.0 * f2 - 18.0 * f3 + 4.0 * f4);
    else if (fl == 5) *o = ho12 * (-22.0 * f1 + 36.0 * f2 - 18.0 * f3 + 4.0 * f4);
    else return 7;
    return 0;
}
/* topdz_low, src/dbydz.F90:1053-1126 (same row for every flag) */
static int topdz_low(double ho12, double* o, double f1, double f2, double f3, double f4, int fl)
{
    if (fl != 0 && fl != 1 && fl != 2 && fl != 5) return 7;
    *o = ho12 * (-4.0 * f1 - 6.0 * f2 + 12.0 * f3 - 2.0 * f4);
    return 0;
}
/* lastdz_up, src/dbydz.F90:776-861 (same row for every flag) */
static int lastdz_up(double ho12, double* o, double f0, double fm1, double fm2, double fm3, int fl)
{
    if (fl != 0 && fl != 1 && fl != 2 && fl != 5) return 7;
    *o = ho12 * (4.0 * f0 + 6.0 * fm1 - 12.0 * fm2 + 2.0 * fm3);
    return 0;
}
/* lastdz, src/dbydz.F90:626-709 */
static int lastdz(double ho12, double* o, double f0, double fm1, double fm2, double fm3, int fl)
{
    if (fl == 0 || fl == 2) *o = 0.0;
    else if (fl == 1) *o = ho12 * (-14.0 * f0 - 36.0 * fm1 + 18.0 * fm2 - 4.0 * fm3);
    else if (fl == 5) *o = ho12 * (22.0 * f0 - 36.0 * fm1 + 18.0 * fm2 - 4.0 * fm3);
    else return 7;
    return 0;
}
/* interior, src/dbydz.F90:371-400 */
static inline double d1int(double ho12, double fm2, double fm1, double fp1, double fp2)
{
    return ho12 * (fm2 - 8.0 * fm1 + 8.0 * fp1 - fp2);
}
/* d2interior, src/dbydz.F90:428-459 */
static inline double d2int(double sc1, double fm2, double fm1, double f0, double fp1, double fp2)
{
    return sc1 * (-fm2 + 16.0 * fm1 - 30.0 * f0 + 16.0 * fp1 - fp2);
}
/* d2topdz / d2lastdz, src/dbydz.F90:1368-1418 / 1180-1228 */
static int d2edge(double sc1, double* o, double f0, double f1, double f2, double f3, double f4,
                  double f5, int fl)
{
    if (fl == 0) *o = 0.0;
    else if (fl == 3)
        *o = sc1 * (45.0 * f0 - 154.0 * f1 + 214.0 * f2 - 156.0 * f3 + 61.0 * f4 - 10.0 * f5);
    else return 7;
    return 0;
}
/* d2topdz_low / d2lastdz_up, src/dbydz.F90:1462-1517 / 1271-1324 */
static int d2edge_low(double sc1, double* o, double f0, double f1, double f2, double f3,
                      double f4, double f5, int fl)
{
    if (fl != 0 && fl != 3) return 7;
    *o = sc1 * (10.0 * f0 - 15.0 * f1 - 4.0 * f2 + 14.0 * f3 - 6.0 * f4 + f5);
    return 0;
}
/* dupinterior, src/dbydz.F90:485-527 */
static inline double dupint(double ho12, double wv, double fm3, double fm2, double fm1, double f0,
                            double fp1, double fp2, double fp3)
{
    if (wv >= 0.0) return ho12 * (-fm3 + 6.0 * fm2 - 18.0 * fm1 + 10.0 * f0 + 3.0 * fp1);
    return ho12 * (fp3 - 6.0 * fp2 + 18.0 * fp1 - 10.0 * f0 - 3.0 * fm1);
}
/* dup_up, src/dbydz.F90:532-574 */
static inline double dupup(double ho12, double wv, double fm2, double fm1, double f0, double fp1,
                           double fp2, double fp3)
{
    if (wv >= 0.0) return ho12 * (2.0 * fm2 - 12.0 * fm1 + 6.0 * f0 + 4.0 * fp1);
    return ho12 * (fp3 - 6.0 * fp2 + 18.0 * fp1 - 10.0 * f0 - 3.0 * fm1);
}
/* dup_low, src/dbydz.F90:579-621 */
static inline double duplow(double ho12, double wv, double fm3, double fm2, double fm1, double f0,
                            double fp1, double fp2)
{
    if (wv >= 0.0) return ho12 * (-fm3 + 6.0 * fm2 - 18.0 * fm1 + 10.0 * f0 + 3.0 * fp1);
    return ho12 * (-2.0 * fp2 + 12.0 * fp1 - 6.0 * f0 - 4.0 * fm1);
}

/* d1bydz serial branch, src/dbydz.F90:1685-1732 */
int orc_d1bydz(orc_ctx* c, const double* fin, double* fout, int fl)
{
    const size_t m = c->nxy;
    const int nz = c->nz;
    const double h = c->ho12;
    int e = 0;
    const double* F = fin;
    for (size_t q = 0; q < m; q++) {
        e |= topdz(h, &fout[q], F[q], F[m + q], F[2 * m + q], F[3 * m + q], fl);
        e |= topdz_low(h, &fout[m + q], F[q], F[m + q], F[2 * m + q], F[3 * m + q], fl);
    }
    if (e) return 7;
    ORC_PAR
    for (int k = 2; k <= nz - 3; k++)
        for (size_t q = 0; q < m; q++)
            fout[k * m + q] = d1int(h, F[(k - 2) * m + q], F[(k - 1) * m + q],
                                    F[(k + 1) * m + q], F[(k + 2) * m + q]);
    const size_t L = (size_t)(nz - 1) * m;
    for (size_t q = 0; q < m; q++) {
        lastdz_up(h, &fout[L - m + q], F[L + q], F[L - m + q], F[L - 2 * m + q], F[L - 3 * m + q], fl);
        lastdz(h, &fout[L + q], F[L + q], F[L - m + q], F[L - 2 * m + q], F[L - 3 * m + q], fl);
    }
    return 0;
}

/* d1upbydz serial branch, src/dbydz.F90:2353-2423; w is the global w(i,j,k) */
int orc_d1upbydz(orc_ctx* c, const double* fin, double* fout, int fl, const double* w)
{
    const size_t m = c->nxy;
    const int nz = c->nz;
    const double h = c->ho12;
    const double* F = fin;
    int e = 0;
    if (!w) w = c->fld[ORC_W];
    for (size_t q = 0; q < m; q++) {
        e |= topdz(h, &fout[q], F[q], F[m + q], F[2 * m + q], F[3 * m + q], fl);
        e |= topdz_low(h, &fout[m + q], F[q], F[m + q], F[2 * m + q], F[3 * m + q], fl);
        fout[2 * m + q] = dupup(h, w[2 * m + q], F[q], F[m + q], F[2 * m + q], F[3 * m + q],
                                F[4 * m + q], F[5 * m + q]);
    }
    if (e) return 7;
    ORC_PAR
    for (int k = 3; k <= nz - 4; k++)
        for (size_t q = 0; q < m; q++)
            fout[k * m + q] = dupint(h, w[k * m + q], F[(k - 3) * m + q], F[(k - 2) * m + q],
                                     F[(k - 1) * m + q], F[k * m + q], F[(k + 1) * m + q],
                                     F[(k + 2) * m + q], F[(k + 3) * m + q]);
    const size_t L = (size_t)(nz - 1) * m;
    for (size_t q = 0; q < m; q++) {
        fout[L - 2 * m + q] = duplow(h, w[L - 2 * m + q], F[L - 5 * m + q], F[L - 4 * m + q],
                                     F[L - 3 * m + q], F[L - 2 * m + q], F[L - m + q], F[L + q]);
        lastdz_up(h, &fout[L - m + q], F[L + q], F[L - m + q], F[L - 2 * m + q], F[L - 3 * m + q], fl);
        lastdz(h, &fout[L + q], F[L + q], F[L - m + q], F[L - 2 * m + q], F[L - 3 * m + q], fl);
    }
    return 0;
}

/* d2bydz serial branch, src/dbydz.F90:3010-3082 */
int orc_d2bydz(orc_ctx* c, const double* fin, double* fout, int fl)
{
    const size_t m = c->nxy;
    const int nz = c->nz;
    const double s = c->sc1;
    const double* F = fin;
    int e = 0;
    for (size_t q = 0; q < m; q++) {
        e |= d2edge(s, &fout[q], F[q], F[m + q], F[2 * m + q], F[3 * m + q], F[4 * m + q],
                    F[5 * m + q], fl);
        e |= d2edge_low(s, &fout[m + q], F[q], F[m + q], F[2 * m + q], F[3 * m + q],
                        F[4 * m + q], F[5 * m + q], fl);
    }
    if (e) return 7;
    ORC_PAR
    for (int k = 2; k <= nz - 3; k++)
        for (size_t q = 0; q < m; q++)
            fout[k * m + q] = d2int(s, F[(k - 2) * m + q], F[(k - 1) * m + q], F[k * m + q],
                                    F[(k + 1) * m + q], F[(k + 2) * m + q]);
    const size_t L = (size_t)(nz - 1) * m;
    for (size_t q = 0; q < m; q++) {
        d2edge_low(s, &fout[L - m + q], F[L + q], F[L - m + q], F[L - 2 * m + q],
                   F[L - 3 * m + q], F[L - 4 * m + q], F[L - 5 * m + q], fl);
        d2edge(s, &fout[L + q], F[L + q], F[L - m + q], F[L - 2 * m + q], F[L - 3 * m + q],
               F[L - 4 * m + q], F[L - 5 * m + q], fl);
    }
    return 0;
}

/* d1bydz1D serial branch, src/dbydz.F90:2025-2066 */
int orc_d1bydz1D(orc_ctx* c, const double* F, double* o, int fl)
{
    const int n = c->nz;
    const double h = c->ho12;
    int e = topdz(h, &o[0], F[0], F[1], F[2], F[3], fl);
    e |= topdz_low(h, &o[1], F[0], F[1], F[2], F[3], fl);
    if (e) return 7;
    for (int k = 2; k <= n - 3; k++) o[k] = d1int(h, F[k - 2], F[k - 1], F[k + 1], F[k + 2]);
    lastdz_up(h, &o[n - 2], F[n - 1], F[n - 2], F[n - 3], F[n - 4], fl);
    lastdz(h, &o[n - 1], F[n - 1], F[n - 2], F[n - 3], F[n - 4], fl);
    return 0;
}

/* d2bydz1D serial branch, src/dbydz.F90:3405-3476 */
int orc_d2bydz1D(orc_ctx* c, const double* F, double* o, int fl)
{
    const int n = c->nz;
    const double s = c->sc1;
    int e = d2edge(s, &o[0], F[0], F[1], F[2], F[3], F[4], F[5], fl);
    e |= d2edge_low(s, &o[1], F[0], F[1], F[2], F[3], F[4], F[5], fl);
    if (e) return 7;
    for (int k = 2; k <= n - 3; k++) o[k] = d2int(s, F[k - 2], F[k - 1], F[k], F[k + 1], F[k + 2]);
    d2edge_low(s, &o[n - 2], F[n - 1], F[n - 2], F[n - 3], F[n - 4], F[n - 5], F[n - 6], fl);
    d2edge(s, &o[n - 1], F[n - 1], F[n - 2], F[n - 3], F[n - 4], F[n - 5], F[n - 6], fl);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* initial state, src/ranset.f90 and src/initial.F90:25-86                  */
/* ------------------------------------------------------------------------ */
/* Park-Miller "minimal standard" generator with Schrage's factorisation and
 * the XOR mask, src/ranset.f90:16-29 */
static double ran0(int* idum)
{
    const int IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836, MASK = 123459876;
    const double AM = 1.0 / IM;
    *idum ^= MASK;
    int k = *idum / IQ;
    *idum = IA * (*idum - k * IQ) - IR * k;
    if (*idum < 0) *idum += IM;
    double r = AM * (*idum);
    *idum ^= MASK;
    return r;
}

/* raw ran0 stream from an explicit idum (test hook) */
void orc_ran0_seq(int idum, int n, double* out)
{
    for (int i = 0; i < n; i++) out[i] = ran0(&idum);
}

/* ranset, src/ranset.f90:1-12 (id is the 1-based global layer index) */
void orc_ran0_fill(int id, double* a, int n, double amp)
{
    int iseed = 1996 + 64 * id;
    for (int i = 0; i < n; i++) a[i] = 2.0 * (ran0(&iseed) - 0.5) * amp;
}

/* static, src/initial.F90:25-86 (amplitude made a parameter; B noise is an
 * extension seeded with 1996+64*id+32, which no T layer seed can equal) */
void orc_static(orc_ctx* c)
{
    const int nz = c->nz;
    const size_t m = c->nxy;
    double *u = c->fld[ORC_U], *v = c->fld[ORC_V], *w = c->fld[ORC_W], *t = c->fld[ORC_T],
           *r = c->fld[ORC_R], *bx = c->fld[ORC_BX], *by = c->fld[ORC_BY], *bz = c->fld[ORC_BZ];
    for (int k = 0; k < nz; k++) {
        double z = (double)k * c->dz;
        double t0 = 1.0 + z * c->p.theta;
        double r0 = pow(t0, c->p.cm);
        orc_ran0_fill(k + 1, t + k * m, (int)m, c->p.t_amp);
        for (size_t q = 0; q < m; q++) {
            u[k * m + q] = 0.0;
            v[k * m + q] = 0.0;
            w[k * m + q] = 0.0;
            r[k * m + q] = r0;
            t[k * m + q] = t[k * m + q] + t0;
            /* src/initial.F90:63-78: horizontal field between perfect conductors */
            const int pc = (c->p.opts & ORC_OPT_PERFECT) != 0;
            bx[k * m + q] = pc ? 1.0 : 0.0;
            by[k * m + q] = 0.0;
            bz[k * m + q] = pc ? 0.0 : 1.0;
        }
        if (c->p.b_amp != 0.0) {
            int iseed = 1996 + 64 * (k + 1) + 32;
            for (size_t q = 0; q < m; q++)
                bz[k * m + q] += 2.0 * (ran0(&iseed) - 0.5) * c->p.b_amp;
        }
    }
}

/* ini_trans_params, src/trans_params.f90:15-69 */
void orc_ini_trans_params(orc_ctx* c)
{
    const double constA = 10.0;
    double *ka = c->fld[ORC_KAPPA], *dka = c->fld[ORC_DKAPPA];
    for (int k = 0; k < c->nz; k++) {
        double z = (double)k * c->dz;
        double constB = constA * (z - 5.0e-1);
        ka[k] = 5.0e-1 * (tanh(constB) + 1);
        dka[k] = 5.0e-1 * constA * pow(cosh(constB), -2.0);
    }
    /* eta(z) / eta_0 (1 everywhere for the reference's constant eta) */
    double* er = c->fld[ORC_ETAR];
    for (int k = 0; k < c->nz; k++) {
        const double z = (double)k * c->dz;
        er[k] = (c->p.eta_c > 1.0)
                    ? 1.0 + (c->p.eta_c - 1.0) * 0.5 * (1.0 + tanh((z - c->p.eta_z0) / c->p.eta_w))
                    : 1.0;
    }
}

/* ------------------------------------------------------------------------ */
/* spectral helpers on the double-halfcomplex layout                        */
/* ------------------------------------------------------------------------ */
/* out = d/dx in spectral space (x-halfcomplex index swap), e.g.
 * src/magnetic.f90:141-150 */
static void spec_ddx(orc_ctx* c, const double* in, double* out)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            const double* a = in + IDX(c, 0, j, k);
            double* o = out + IDX(c, 0, j, k);
            o[0] = 0.0;
            o[nx / 2] = 0.0;
            for (int i = 1; i <= nx / 2 - 1; i++) {
                o[i] = -a[nx - i] * c->akx1[i];
                o[nx - i] = a[i] * c->akx1[nx - i];
            }
        }
}

/* out = d/dy in spectral space (y-halfcomplex index swap), e.g.
 * src/magnetic.f90:155-164 */
static void spec_ddy(orc_ctx* c, const double* in, double* out)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    for (int k = 0; k < nz; k++)
        for (int j = 1; j <= ny / 2 - 1; j++)
            for (int i = 0; i < nx; i++) {
                out[IDX(c, i, 0, k)] = 0.0;
                out[IDX(c, i, ny / 2, k)] = 0.0;
                out[IDX(c, i, j, k)] = -in[IDX(c, i, ny - j, k)] * c->aky1[j];
                out[IDX(c, i, ny - j, k)] = in[IDX(c, i, j, k)] * c->aky1[ny - j];
            }
}

/* 2/3 truncation of one double-halfcomplex field, src/magnetic.f90:70-103 */
static void spec_truncate(orc_ctx* c, double* e)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            for (int i = 1; i <= c->icutoffx; i++) {
                e[IDX(c, nx / 2 + i, j, k)] = 0.0;
                e[IDX(c, nx / 2 - i, j, k)] = 0.0;
            }
            e[IDX(c, nx / 2, j, k)] = 0.0;
        }
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) {
            e[IDX(c, i, ny / 2, k)] = 0.0;
            for (int j = 1; j <= c->icutoffy; j++) {
                e[IDX(c, i, ny / 2 + j, k)] = 0.0;
                e[IDX(c, i, ny / 2 - j, k)] = 0.0;
            }
        }
}

/* out = in / (akx2+aky2) with the DC set to 0, src/magnetic.f90:429-446 */
static void spec_invlap(orc_ctx* c, double* a)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    for (int k = 0; k < nz; k++) {
        a[IDX(c, 0, 0, k)] = 0.0;
        for (int i = 1; i < nx; i++) a[IDX(c, i, 0, k)] = a[IDX(c, i, 0, k)] / c->akx2[i];
        for (int j = 1; j < ny; j++) a[IDX(c, 0, j, k)] = a[IDX(c, 0, j, k)] / c->aky2[j];
        for (int j = 1; j < ny; j++)
            for (int i = 1; i < nx; i++)
                a[IDX(c, i, j, k)] = a[IDX(c, i, j, k)] / (c->akx2[i] + c->aky2[j]);
    }
}

/* ------------------------------------------------------------------------ */
/* magnetic part, src/magnetic.f90                                          */
/* ------------------------------------------------------------------------ */

/* J^2 from (spectral) ez and physical bx,by; shared tail of mag_init and
 * bandj, src/magnetic.f90:199-261 and 584-645 */
/* eta(z) extension: on when eta_c > 1 */
static int eta_on(const orc_ctx* c) { return c->p.eta_c > 1.0; }

static int jsq_tail(orc_ctx* c)
{
    const size_t n3 = c->n3;
    double *ex = c->fld[ORC_EX], *ey = c->fld[ORC_EY], *ez = c->fld[ORC_EZ];
    double *wk1 = c->fld[ORC_WK1], *wk2 = c->fld[ORC_WK2], *wk3 = c->fld[ORC_WK3];
    double* jsq = c->fld[ORC_JSQ];
    /* dBz/dy and dBz/dx in Fourier space */
    spec_ddy(c, ez, wk1);
    spec_ddx(c, ez, wk2);
    orc_inversefft(c, wk1, ex);
    orc_inversefft(c, wk2, ey);
    orc_inversefft(c, wk3, ez);
    int e = orc_d1bydz(c, c->fld[ORC_BY], wk1, 5);
    e |= orc_d1bydz(c, c->fld[ORC_BX], wk2, 5);
    if (e) return 7;
    ORC_PAR
    for (size_t q = 0; q < n3; q++)
        jsq[q] = ez[q] * ez[q] + ((ex[q] - wk1[q]) * (ex[q] - wk1[q]))
                 + ((wk2[q] - ey[q]) * (wk2[q] - ey[q]));
    if (eta_on(c)) {
        /* J = (dBz/dy - dBy/dz, dBx/dz - dBz/dx, Jz) for the eta(z) EMF */
        double *jx = c->fld[ORC_JX], *jy = c->fld[ORC_JY], *jz = c->fld[ORC_JZ];
        ORC_PAR
        for (size_t q = 0; q < n3; q++) {
            jx[q] = ex[q] - wk1[q];
            jy[q] = wk2[q] - ey[q];
            jz[q] = ez[q];
        }
    }
    return 0;
}

static void zero_work(orc_ctx* c)
{
    for (int id = ORC_EX; id <= ORC_WK4; id++) memset(c->fld[id], 0, sizeof(double) * c->n3);
}

/* mag_init, src/magnetic.f90:5-279 (dynamo seed off) */
int orc_mag_init(orc_ctx* c)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    double *ex = c->fld[ORC_EX], *ey = c->fld[ORC_EY], *ez = c->fld[ORC_EZ];
    double *wk3 = c->fld[ORC_WK3], *wk4 = c->fld[ORC_WK4];
    double *pol = c->fld[ORC_POL], *tor = c->fld[ORC_TOR];
    if (c->p.opts & ORC_OPT_DYNAMO) {
        /* src/magnetic.f90:29-57: seed field (1-based i, j in the cosines) */
        const double seed = 0.003, tx = 8.0 * atan(1.0) / (double)nx, ty = 8.0 * atan(1.0) / (double)ny;
        const int pc = (c->p.opts & ORC_OPT_PERFECT) != 0;
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    const size_t q = IDX(c, i, j, k);
                    if (pc) {
                        c->fld[ORC_BX][q] = seed * cos(ty * (double)(j + 1));
                        c->fld[ORC_BY][q] = seed * cos(tx * (double)(i + 1));
                        c->fld[ORC_BZ][q] = 0.0;
                    } else {
                        c->fld[ORC_BX][q] = 0.0;
                        c->fld[ORC_BY][q] = 0.0;
                        c->fld[ORC_BZ][q] = seed * cos(tx * (double)(i + 1)) * cos(ty * (double)(j + 1));
                    }
                }
    }
    orc_forwardfft(c, c->fld[ORC_BX], ex);
    orc_forwardfft(c, c->fld[ORC_BY], ey);
    orc_forwardfft(c, c->fld[ORC_BZ], ez);
    spec_truncate(c, ex);
    spec_truncate(c, ey);
    spec_truncate(c, ez);
    for (int k = 0; k < nz; k++) {
        c->fld[ORC_BXBAR][k] = ex[IDX(c, 0, 0, k)];
        c->fld[ORC_BYBAR][k] = ey[IDX(c, 0, 0, k)];
    }
    for (int k = 0; k < nz; k++) {
        c->fld[ORC_BXBAR][k] *= c->xn;
        c->fld[ORC_BYBAR][k] *= c->xn;
    }
    c->bzbar = ez[IDX(c, 0, 0, 0)] * c->xn;
    /* pol from FFT(Bz) = FFT(-del_H^2 pol) */
    for (int k = 0; k < nz; k++) {
        pol[IDX(c, 0, 0, k)] = 0.0;
        for (int i = 1; i < nx; i++) pol[IDX(c, i, 0, k)] = -1.0 * ez[IDX(c, i, 0, k)] / c->akx2[i];
        for (int j = 1; j < ny; j++) pol[IDX(c, 0, j, k)] = -1.0 * ez[IDX(c, 0, j, k)] / c->aky2[j];
        for (int j = 1; j < ny; j++)
            for (int i = 1; i < nx; i++)
                pol[IDX(c, i, j, k)] = -1.0 * ez[IDX(c, i, j, k)] / (c->akx2[i] + c->aky2[j]);
    }
    /* FFT(dBy/dx) in wk3, FFT(dBx/dy) in wk4, wk3 = FFT(jz) */
    spec_ddx(c, ey, wk3);
    spec_ddy(c, ex, wk4);
    ORC_PAR
    for (size_t q = 0; q < c->n3; q++) wk3[q] = wk3[q] - wk4[q];
    for (int k = 0; k < nz; k++) {
        tor[IDX(c, 0, 0, k)] = 0.0;
        for (int i = 1; i < nx; i++) tor[IDX(c, i, 0, k)] = -1.0 * wk3[IDX(c, i, 0, k)] / c->akx2[i];
        for (int j = 1; j < ny; j++) tor[IDX(c, 0, j, k)] = -1.0 * wk3[IDX(c, 0, j, k)] / c->aky2[j];
        for (int j = 1; j < ny; j++)
            for (int i = 1; i < nx; i++)
                tor[IDX(c, i, j, k)] = -1.0 * wk3[IDX(c, i, j, k)] / (c->akx2[i] + c->aky2[j]);
    }
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                wk3[IDX(c, i, j, k)] = -1.0 * tor[IDX(c, i, j, k)] * (c->akx2[i] + c->aky2[j]);
    int e = jsq_tail(c);
    if (e) return e;
    orc_mag_boundary(c);
    zero_work(c);
    return 0;
}

/* mag_deriv, src/magnetic.f90:285-487 */
int orc_mag_deriv(orc_ctx* c)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t n3 = c->n3;
    double *u = c->fld[ORC_U], *v = c->fld[ORC_V], *w = c->fld[ORC_W];
    double *bx = c->fld[ORC_BX], *by = c->fld[ORC_BY], *bz = c->fld[ORC_BZ];
    double *ex = c->fld[ORC_EX], *ey = c->fld[ORC_EY], *ez = c->fld[ORC_EZ];
    double *wk1 = c->fld[ORC_WK1], *wk2 = c->fld[ORC_WK2], *wk3 = c->fld[ORC_WK3],
           *wk4 = c->fld[ORC_WK4];
    double *pol = c->fld[ORC_POL], *tor = c->fld[ORC_TOR];
    double* poldot3 = c->fld[ORC_POLDOT] + 2 * n3;
    double* tordot3 = c->fld[ORC_TORDOT] + 2 * n3;
    double* bxbardot3 = c->fld[ORC_BXBARDOT] + 2 * (size_t)nz;
    double* bybardot3 = c->fld[ORC_BYBARDOT] + 2 * (size_t)nz;
    double* w1d = (double*)malloc(sizeof(double) * (size_t)nz);
    double* w2d = (double*)malloc(sizeof(double) * (size_t)nz);
    int e = 0;

    ORC_PAR
    for (size_t q = 0; q < n3; q++) {
        wk1[q] = v[q] * bz[q] - w[q] * by[q];
        wk2[q] = w[q] * bx[q] - u[q] * bz[q];
        wk3[q] = u[q] * by[q] - v[q] * bx[q];
    }
    if (eta_on(c)) {
        /* eta(z) extension: E = u x B - (eta(z) - zeta C_k) J, so that curl E
         * adds -curl((eta - eta_0) J) to the constant-eta induction terms */
        const double* er = c->fld[ORC_ETAR];
        const double *jx = c->fld[ORC_JX], *jy = c->fld[ORC_JY], *jz = c->fld[ORC_JZ];
        const size_t m = c->nxy;
        ORC_PAR
        for (int k = 0; k < nz; k++) {
            const double e1 = c->zck * (er[k] - 1.0);
            for (size_t p = 0; p < m; p++) {
                const size_t q = (size_t)k * m + p;
                wk1[q] = wk1[q] - e1 * jx[q];
                wk2[q] = wk2[q] - e1 * jy[q];
                wk3[q] = wk3[q] - e1 * jz[q];
            }
        }
    }
    if (c->cl_flux) {
        /* closure: E = u x B + eps */
        const double* fl = c->cl_flux;
        for (size_t q = 0; q < n3; q++) {
            wk1[q] = wk1[q] + fl[6 * n3 + q];
            wk2[q] = wk2[q] + fl[7 * n3 + q];
            wk3[q] = wk3[q] + fl[8 * n3 + q];
        }
    }
    orc_forwardfft(c, wk1, ex);
    orc_forwardfft(c, wk2, ey);
    orc_forwardfft(c, wk3, ez);
    for (int k = 0; k < nz; k++) {
        w1d[k] = c->xn * ex[IDX(c, 0, 0, k)];
        w2d[k] = c->xn * ey[IDX(c, 0, 0, k)];
    }
    e |= orc_d1bydz1D(c, w2d, bxbardot3, 5);
    e |= orc_d1bydz1D(c, w1d, bybardot3, 5);
    e |= orc_d2bydz1D(c, c->fld[ORC_BXBAR], w1d, 3);
    e |= orc_d2bydz1D(c, c->fld[ORC_BYBAR], w2d, 3);
    for (int k = 0; k < nz; k++) {
        bxbardot3[k] = c->zck * w1d[k] - bxbardot3[k];
        bybardot3[k] = c->zck * w2d[k] + bybardot3[k];
    }
    /* x and y derivatives of ex, ey in Fourier space */
    spec_ddx(c, ex, wk1);
    spec_ddx(c, ey, wk2);
    spec_ddy(c, ex, wk3);
    spec_ddy(c, ey, wk4);
    ORC_PAR
    for (size_t q = 0; q < n3; q++) {
        wk2[q] = wk3[q] - wk2[q];
        wk1[q] = wk1[q] + wk4[q];
    }
    e |= orc_d1bydz(c, wk1, wk3, 5);
    /* inverse del_H^2 of wk2 and wk3 */
    spec_invlap(c, wk2);
    spec_invlap(c, wk3);
    e |= orc_d2bydz(c, pol, wk1, 3);
    e |= orc_d2bydz(c, tor, wk4, 3);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t q = IDX(c, i, j, k);
                poldot3[q] = c->zck * pol[q] * (c->akx2[i] + c->aky2[j]) + c->zck * wk1[q] + wk2[q];
                tordot3[q] = c->zck * tor[q] * (c->akx2[i] + c->aky2[j]) + c->zck * wk4[q] - wk3[q]
                             + ez[q];
            }
    free(w1d);
    free(w2d);
    return e ? 7 : 0;
}

/* bandj, src/magnetic.f90:492-664 */
int orc_bandj(orc_ctx* c)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t n3 = c->n3;
    double *ex = c->fld[ORC_EX], *ey = c->fld[ORC_EY], *ez = c->fld[ORC_EZ];
    double *wk1 = c->fld[ORC_WK1], *wk2 = c->fld[ORC_WK2], *wk3 = c->fld[ORC_WK3];
    double *pol = c->fld[ORC_POL], *tor = c->fld[ORC_TOR];
    double *bx = c->fld[ORC_BX], *by = c->fld[ORC_BY], *bz = c->fld[ORC_BZ];
    if (orc_d1bydz(c, pol, wk1, 5)) return 7;
    spec_ddx(c, wk1, ex);
    spec_ddx(c, tor, wk3);
    spec_ddy(c, wk1, ey);
    spec_ddy(c, tor, wk2);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t q = IDX(c, i, j, k);
                ex[q] = ex[q] + wk2[q];
                ey[q] = ey[q] - wk3[q];
                ez[q] = -1.0 * pol[q] * (c->akx2[i] + c->aky2[j]);
            }
    orc_inversefft(c, ex, bx);
    orc_inversefft(c, ey, by);
    orc_inversefft(c, ez, bz);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t q = IDX(c, i, j, k);
                bx[q] = bx[q] + c->fld[ORC_BXBAR][k];
                by[q] = by[q] + c->fld[ORC_BYBAR][k];
                bz[q] = bz[q] + c->bzbar;
                wk3[q] = -1.0 * tor[q] * (c->akx2[i] + c->aky2[j]);
            }
    (void)n3;
    int e = jsq_tail(c);
    zero_work(c);
    return e;
}

/* magdealias, src/magnetic.f90:666-717 */
void orc_magdealias(orc_ctx* c)
{
    spec_truncate(c, c->fld[ORC_POL]);
    spec_truncate(c, c->fld[ORC_TOR]);
}

/* ------------------------------------------------------------------------ */
/* hydro_deriv, src/hydro_deriv.f90:5-505                                   */
/* ------------------------------------------------------------------------ */
int orc_hydro_deriv(orc_ctx* c)
{
    const size_t n3 = c->n3, m = c->nxy;
    const int nz = c->nz;
    double *u = c->fld[ORC_U], *v = c->fld[ORC_V], *w = c->fld[ORC_W], *t = c->fld[ORC_T],
           *r = c->fld[ORC_R], *bx = c->fld[ORC_BX], *by = c->fld[ORC_BY], *bz = c->fld[ORC_BZ],
           *jsq = c->fld[ORC_JSQ];
    double *ex = c->fld[ORC_EX], *ey = c->fld[ORC_EY], *ez = c->fld[ORC_EZ];
    double *wk1 = c->fld[ORC_WK1], *wk2 = c->fld[ORC_WK2], *wk3 = c->fld[ORC_WK3],
           *wk4 = c->fld[ORC_WK4];
    double* udot = c->fld[ORC_UDOT] + 2 * n3;
    double* vdot = c->fld[ORC_VDOT] + 2 * n3;
    double* wdot = c->fld[ORC_WDOT] + 2 * n3;
    double* tdot = c->fld[ORC_TDOT] + 2 * n3;
    double* rdot = c->fld[ORC_RDOT] + 2 * n3;
    const double f = c->p.f;
    int e = 0;

    /* 33-53 */
    orc_d1bydx(c, r, wk1);
    orc_d1bydy(c, r, wk2);
    ORC_PAR
    for (size_t q = 0; q < n3; q++) {
        rdot[q] = -u[q] * wk1[q] - v[q] * wk2[q];
        ex[q] = -t[q] * wk1[q];
        ey[q] = -t[q] * wk2[q];
    }
    /* 64-110 */
    orc_d2bydx(c, t, udot, wk1);
    orc_d2bydy(c, t, vdot, wk2);
    e |= orc_d1upbydz(c, t, wdot, 5, w);
    e |= orc_d1bydz(c, t, wk4, 5);
    e |= orc_d2bydz(c, t, wk3, 0);
    orc_ini_trans_params(c);
    for (int k = 0; k < nz; k++) {
        const double kap = c->fld[ORC_KAPPA][k], dkap = c->fld[ORC_DKAPPA][k];
        for (size_t p = 0; p < m; p++) {
            size_t q = (size_t)k * m + p;
            double u0 = u[q], v0 = v[q], w0 = w[q];
            tdot[q] = -r[q] * (u0 * udot[q] + v0 * vdot[q] + w0 * wdot[q])
                      + c->gck * kap * (wk1[q] + wk2[q] + wk3[q])
                      + c->gck * dkap * wdot[q]
                      + c->fzckg * c->fld[ORC_ETAR][k] * jsq[q];
            udot[q] = ex[q] - r[q] * udot[q];
            vdot[q] = ey[q] - r[q] * vdot[q];
            wdot[q] = -r[q] * wk4[q];
            ex[q] = 0.0;
            ey[q] = 0.0;
        }
    }
    /* 122-168 */
    orc_d1bydx(c, u, wk1);
    orc_d1bydy(c, v, wk2);
    e |= orc_d1upbydz(c, w, wk3, 1, w);
    e |= orc_d1bydz(c, w, wk4, 1);
    ORC_PAR
    for (size_t q = 0; q < n3; q++) {
        double z1 = wk1[q], z2 = wk2[q], z3 = wk3[q], z4 = wk4[q];
        double tt = t[q], rr = r[q];
        double divu = z1 + z2 + z4;
        rdot[q] = rdot[q] - rr * divu;
        tdot[q] = tdot[q] - c->gm1 * tt * divu * rr
                  + 2 * c->sckg * (z1 * z1 + z2 * z2 + z4 * z4 - third * divu * divu);
        double temp1 = c->fo2 * bx[q] * bx[q];
        double temp2 = c->fo2 * by[q] * by[q];
        double temp3 = c->fo2 * bz[q] * bz[q];
        divu = c->scko3 * divu;
        double u0 = u[q], v0 = v[q], w0 = w[q];
        udot[q] = udot[q] - rr * u0 * z1;
        vdot[q] = vdot[q] - rr * v0 * z2;
        wdot[q] = wdot[q] - rr * w0 * z3;
        wk1[q] = divu + temp1 - temp2 - temp3;
        wk2[q] = divu + temp2 - temp1 - temp3;
        wk3[q] = c->scko3 * z1 + c->scko3 * z2 + temp3 - temp1 - temp2;
        wk4[q] = 4.0 * c->scko3 * z4;
    }
    /* 181-219 */
    orc_d1bydx(c, wk1, ex);
    orc_d1bydy(c, wk2, ey);
    e |= orc_d1bydz(c, wk3, ez, 2);
    e |= orc_d1bydz(c, wk4, wk3, 5);
    orc_d1bydx(c, v, wk2);
    orc_d1bydy(c, u, wk1);
    ORC_PAR
    for (size_t q = 0; q < n3; q++) {
        udot[q] = udot[q] + ex[q];
        vdot[q] = vdot[q] + ey[q];
        wdot[q] = wdot[q] + ez[q] + wk3[q];
        ex[q] = 0.0;
        ey[q] = 0.0;
        ez[q] = 0.0;
        double rr = r[q], w1 = wk1[q], w2 = wk2[q];
        udot[q] = udot[q] - rr * v[q] * w1;
        vdot[q] = vdot[q] - rr * u[q] * w2;
        wk3[q] = f * by[q] * bx[q];
        tdot[q] = tdot[q] + c->sckg * (w1 + w2) * (w1 + w2);
    }
    /* 230-261 */
    orc_d1bydx(c, wk3, wk2);
    orc_d1bydy(c, wk3, wk1);
    orc_d1bydx(c, w, wk3);
    e |= orc_d1upbydz(c, u, wk4, 2, w);
    ORC_PAR
    for (size_t q = 0; q < n3; q++) {
        udot[q] = udot[q] + wk1[q];
        vdot[q] = vdot[q] + wk2[q];
        double rr = r[q], w3 = wk3[q], w4 = wk4[q];
        udot[q] = udot[q] - rr * w[q] * w4;
        wdot[q] = wdot[q] - rr * u[q] * w3;
        double temp1 = f * bz[q] * bx[q];
        wk1[q] = temp1;
        wk2[q] = temp1;
        tdot[q] = tdot[q] + c->sckg * (w3 + w4) * (w3 + w4);
    }
    /* 272-309 */
    e |= orc_d1bydz(c, wk1, wk3, 5);
    orc_d1bydx(c, wk2, wk4);
    orc_d1bydy(c, w, wk1);
    e |= orc_d1upbydz(c, v, wk2, 2, w);
    ORC_PAR
    for (size_t q = 0; q < n3; q++) {
        double rr = r[q];
        udot[q] = udot[q] + wk3[q];
        wdot[q] = wdot[q] + wk4[q] + c->thmplus1 * rr;
        double w1 = wk1[q], w2 = wk2[q];
        vdot[q] = vdot[q] - rr * w[q] * w2;
        wdot[q] = wdot[q] - rr * v[q] * w1;
        double temp1 = f * bz[q] * by[q];
        wk3[q] = temp1;
        wk4[q] = temp1;
        tdot[q] = tdot[q] + c->sckg * (w1 + w2) * (w1 + w2);
        udot[q] = udot[q] - rr * c->cor * (w[q] * c->cpsi + v[q] * c->spsi);
        vdot[q] = vdot[q] - rr * c->cor * (-u[q] * c->spsi);
        wdot[q] = wdot[q] - rr * c->cor * (-u[q] * c->cpsi);
    }
    /* 320-338 */
    orc_d1bydy(c, wk3, wk1);
    e |= orc_d1bydz(c, wk4, wk2, 5);
    ORC_PAR
    for (size_t q = 0; q < n3; q++) {
        vdot[q] = vdot[q] + wk2[q];
        wdot[q] = wdot[q] + wk1[q];
    }
    /* 349-367 */
    e |= orc_d2bydz(c, u, wk1, 3);
    e |= orc_d2bydz(c, v, wk2, 3);
    ORC_PAR
    for (size_t q = 0; q < n3; q++) {
        udot[q] = udot[q] + c->sck * wk1[q];
        vdot[q] = vdot[q] + c->sck * wk2[q];
    }
    /* 378-398 */
    orc_d2bydx(c, u, ex, wk1);
    orc_d2bydx(c, v, ey, wk2);
    orc_d2bydx(c, w, ez, wk3);
    ORC_PAR
    for (size_t q = 0; q < n3; q++) {
        udot[q] = udot[q] + c->sck * wk1[q];
        vdot[q] = vdot[q] + c->sck * wk2[q];
        wdot[q] = wdot[q] + c->sck * wk3[q];
    }
    /* 409-432 */
    orc_d2bydy(c, u, ex, wk1);
    orc_d2bydy(c, v, ey, wk2);
    orc_d2bydy(c, w, ez, wk3);
    ORC_PAR
    for (size_t q = 0; q < n3; q++) {
        udot[q] = udot[q] + c->sck * wk1[q];
        vdot[q] = vdot[q] + c->sck * wk2[q];
        wdot[q] = wdot[q] + c->sck * wk3[q];
        ex[q] = 0.0;
        ey[q] = 0.0;
        ez[q] = 0.0;
    }
    /* 443-482: drho/dz, boundary rows from wdot/T */
    e |= orc_d1upbydz(c, r, wk1, 0, w);
    e |= orc_d1bydz(c, r, wk2, 0);
    {
        const size_t L = (size_t)(nz - 1) * m;
        for (size_t p = 0; p < m; p++) {
            wk1[p] = wdot[p] / t[p];
            wk2[p] = wdot[p] / t[p];
            wk1[L + p] = wdot[L + p] / t[L + p];
            wk2[L + p] = wdot[L + p] / t[L + p];
        }
    }
    ORC_PAR
    for (size_t q = 0; q < n3; q++) {
        rdot[q] = rdot[q] - w[q] * wk1[q];
        wdot[q] = wdot[q] - t[q] * wk2[q];
    }
    /* 487-497 */
    ORC_PAR
    for (size_t q = 0; q < n3; q++) {
        double rr = 1.0 / (r[q]);
        tdot[q] = tdot[q] * rr;
        udot[q] = udot[q] * rr;
        vdot[q] = vdot[q] * rr;
        wdot[q] = wdot[q] * rr;
    }
    if (c->cl_flux) {
        const int ec = closure_tend(c);
        if (ec) return ec;
    }
    return e ? 7 : 0;
}

/* ------------------------------------------------------------------------ */
/* boundary conditions (defaults), src/dbydz.F90:3686-4197 and 4205-4791     */
/* ------------------------------------------------------------------------ */
void orc_boundary(orc_ctx* c)
{
    const size_t m = c->nxy;
    const size_t L = (size_t)(c->nz - 1) * m;
    double *u = c->fld[ORC_U], *v = c->fld[ORC_V], *w = c->fld[ORC_W], *t = c->fld[ORC_T];
    for (size_t p = 0; p < m; p++) w[p] = 0.0;
    for (size_t p = 0; p < m; p++) {
        u[p] = elf * (18.0 * u[m + p] - 9.0 * u[2 * m + p] + 2.0 * u[3 * m + p]);
        v[p] = elf * (18.0 * v[m + p] - 9.0 * v[2 * m + p] + 2.0 * v[3 * m + p]);
    }
    if (c->p.opts & ORC_OPT_RADBC) {
        /* 4037-4057: heat flux matched to a blackbody term, dT/dz = theta T^4
         * by the one-sided third-order stencil; ten Newton steps */
        const double th = c->p.theta, dz = c->dz;
        for (size_t p = 0; p < m; p++) {
            double t1 = t[p], ft1 = 0.0, dft1 = 1.0;
            for (int it = 0; it < 10; it++) {
                ft1 = 6.0 * th * dz * t1 * t1 * t1 * t1 + 11.0 * t1 - 18.0 * t[m + p] + 9.0 * t[2 * m + p]
                      - 2.0 * t[3 * m + p];
                dft1 = 24.0 * th * dz * t1 * t1 * t1 + 11.0;
                t1 = t1 - (ft1 / dft1);
            }
            t[p] = t1;
        }
    } else {
        for (size_t p = 0; p < m; p++) t[p] = 1.0;      /* 4119 */
    }
    for (size_t p = 0; p < m; p++) w[L + p] = 0.0;
    for (size_t p = 0; p < m; p++) {
        u[L + p] = elf * (18.0 * u[L - m + p] - 9.0 * u[L - 2 * m + p] + 2.0 * u[L - 3 * m + p]);
        v[L + p] = elf * (18.0 * v[L - m + p] - 9.0 * v[L - 2 * m + p] + 2.0 * v[L - 3 * m + p]);
    }
    if (c->p.opts & ORC_OPT_CONSTFLUX) {
        /* 4133-4143: dT/dz = theta at the bottom */
        for (size_t p = 0; p < m; p++)
            t[L + p] = elf * (18.0 * t[L - m + p] - 9.0 * t[L - 2 * m + p] + 2.0 * t[L - 3 * m + p]
                              + 6.0 * c->p.theta * c->dz);
    } else {
        for (size_t p = 0; p < m; p++) t[L + p] = 1.0 + c->p.theta; /* 4183 */
    }
}

void orc_mag_boundary(orc_ctx* c)
{
    const size_t m = c->nxy;
    const int nz = c->nz;
    const size_t L = (size_t)(nz - 1) * m;
    double *pol = c->fld[ORC_POL], *tor = c->fld[ORC_TOR];
    if (c->p.opts & ORC_OPT_PERFECT) {
        /* 4314-4400: pol = 0 (bz = 0), zero d/dz of bx, by (tor and the means) */
        double *bxb = c->fld[ORC_BXBAR], *byb = c->fld[ORC_BYBAR];
        for (size_t p = 0; p < m; p++) pol[p] = 0.0;
        bxb[0] = elf * (18.0 * bxb[1] - 9.0 * bxb[2] + 2.0 * bxb[3]);
        byb[0] = elf * (18.0 * byb[1] - 9.0 * byb[2] + 2.0 * byb[3]);
        for (size_t p = 0; p < m; p++)
            tor[p] = elf * (18.0 * tor[m + p] - 9.0 * tor[2 * m + p] + 2.0 * tor[3 * m + p]);
        for (size_t p = 0; p < m; p++) pol[L + p] = 0.0;
        bxb[nz - 1] = elf * (18.0 * bxb[nz - 2] - 9.0 * bxb[nz - 3] + 2.0 * bxb[nz - 4]);
        byb[nz - 1] = elf * (18.0 * byb[nz - 2] - 9.0 * byb[nz - 3] + 2.0 * byb[nz - 4]);
        for (size_t p = 0; p < m; p++)
            tor[L + p] = elf * (18.0 * tor[L - m + p] - 9.0 * tor[L - 2 * m + p] + 2.0 * tor[L - 3 * m + p]);
        return;
    }
    c->fld[ORC_BXBAR][0] = 0.0;
    c->fld[ORC_BYBAR][0] = 0.0;
    for (size_t p = 0; p < m; p++) tor[p] = 0.0;
    for (size_t p = 0; p < m; p++)
        pol[p] = elf * (18.0 * pol[m + p] - 9.0 * pol[2 * m + p] + 2.0 * pol[3 * m + p]);
    c->fld[ORC_BXBAR][nz - 1] = 0.0;
    c->fld[ORC_BYBAR][nz - 1] = 0.0;
    for (size_t p = 0; p < m; p++) tor[L + p] = 0.0;
    for (size_t p = 0; p < m; p++)
        pol[L + p] = elf * (18.0 * pol[L - m + p] - 9.0 * pol[L - 2 * m + p] + 2.0 * pol[L - 3 * m + p]);
}

/* ------------------------------------------------------------------------ */
/* diagnostics, src/utils.F90 and src/dbydz.F90:4796-4925                    */
/* ------------------------------------------------------------------------ */
/* GlobalAverages, src/utils.F90:22-139 (subtracts ubar, vbar from u, v) */
void orc_global_averages(orc_ctx* c, double out[4])
{
    const int nz = c->nz;
    const size_t m = c->nxy;
    double *u = c->fld[ORC_U], *v = c->fld[ORC_V], *w = c->fld[ORC_W], *r = c->fld[ORC_R];
    double *bx = c->fld[ORC_BX], *by = c->fld[ORC_BY], *bz = c->fld[ORC_BZ];
    double urbar = 0.0, vrbar = 0.0, rbar = 0.0, menergy = 0.0, kenergy = 0.0;
    for (int k = 0; k < nz; k++) {
        double urav = 0.0, vrav = 0.0, rav = 0.0, keav = 0.0, bsqav = 0.0;
        for (size_t p = 0; p < m; p++) {
            size_t q = (size_t)k * m + p;
            double uu = u[q], vv = v[q], rr = r[q], ww = w[q];
            rav = rav + rr;
            urav = urav + rr * uu;
            vrav = vrav + rr * vv;
            keav = keav + 0.5 * rr * (uu * uu + vv * vv + ww * ww);
            bsqav = bsqav + bx[q] * bx[q] + by[q] * by[q] + bz[q] * bz[q];
        }
        if (k == 0 || k == nz - 1) {
            urbar = urbar + 0.5 * urav;
            vrbar = vrbar + 0.5 * vrav;
            rbar = rbar + 0.5 * rav;
            menergy = menergy + 0.25 * bsqav;
            kenergy = kenergy + 0.5 * keav;
        } else {
            urbar = urbar + urav;
            vrbar = vrbar + vrav;
            rbar = rbar + rav;
            menergy = menergy + 0.5 * bsqav;
            kenergy = kenergy + keav;
        }
    }
    double ubar = urbar / rbar, vbar = vrbar / rbar;
    menergy = menergy * c->xzn;
    kenergy = kenergy * c->xzn;
    ORC_PAR
    for (size_t q = 0; q < c->n3; q++) {
        u[q] = u[q] - ubar;
        v[q] = v[q] - vbar;
    }
    out[0] = ubar; out[1] = vbar; out[2] = kenergy; out[3] = menergy;
}

/* nusselt, src/dbydz.F90:4796-4925 (serial branch) */
double orc_nusselt(orc_ctx* c)
{
    const size_t m = c->nxy;
    const size_t L = (size_t)(c->nz - 1) * m;
    const double* t = c->fld[ORC_T];
    double tzbar = 0.0;
    for (size_t p = 0; p < m; p++)
        tzbar = tzbar + ((1.0 / (6.0 * c->dz))
                         * (11.0 * t[L + p] - 18.0 * t[L - m + p] + 9.0 * t[L - 2 * m + p]
                            - 2.0 * t[L - 3 * m + p]));
    tzbar = tzbar * c->xn;
    return (tzbar - c->nusstmp) / (c->p.theta - c->nusstmp);
}

/* checkdivb, src/utils.F90:249-318 */
int orc_checkdivb(orc_ctx* c, double out[2])
{
    const int nz = c->nz;
    const size_t m = c->nxy;
    double *wk1 = c->fld[ORC_WK1], *wk2 = c->fld[ORC_WK2], *wk3 = c->fld[ORC_WK3];
    orc_d1bydx(c, c->fld[ORC_BX], wk1);
    orc_d1bydy(c, c->fld[ORC_BY], wk2);
    int e = orc_d1bydz(c, c->fld[ORC_BZ], wk3, 5);
    double divbmax = -1.0e30, divbmin = 1.0;
    for (int k = 0; k < nz; k++) {
        double tlmax = -1.0e30, tlmin = 1.0e30;
        for (size_t p = 0; p < m; p++) {
            size_t q = (size_t)k * m + p;
            double t1 = wk1[q] + wk2[q] + wk3[q];
            tlmax = fmax(tlmax, t1);
            tlmin = fmin(tlmin, t1);
        }
        divbmax = fmax(divbmax, tlmax);
        divbmin = fmin(divbmin, tlmin);
    }
    out[0] = divbmax;
    out[1] = divbmin;
    return e ? 7 : 0;
}

/* sminmax, src/utils.F90:323-371 */
void orc_sminmax(orc_ctx* c, const double* z, double out[2])
{
    double zmin = z[0], zmax = z[0];
    for (size_t q = 0; q < c->n3; q++) {
        zmin = fmin(zmin, z[q]);
        zmax = fmax(zmax, z[q]);
    }
    out[0] = zmin;
    out[1] = zmax;
}

/* calcdt, src/utils.F90:144-244 */
double orc_calcdt(orc_ctx* c)
{
    const double f = c->p.f;
    double *u = c->fld[ORC_U], *v = c->fld[ORC_V], *w = c->fld[ORC_W], *r = c->fld[ORC_R],
           *t = c->fld[ORC_T], *bx = c->fld[ORC_BX], *by = c->fld[ORC_BY], *bz = c->fld[ORC_BZ];
    double denmn = r[0], velmxx = -1.0e-50, velmxz = -1.0e-50;
    for (size_t q = 0; q < c->n3; q++) {
        double u0 = u[q], v0 = v[q], w0 = w[q], r0 = r[q], t0 = t[q];
        double bz0 = bz[q], bx0 = bx[q], by0 = by[q];
        double bsq0 = bx0 * bx0 + by0 * by0 + bz0 * bz0;
        double xx = sqrt(u0 * u0 + v0 * v0) + c->g1 * sqrt(t0) + sqrt((bsq0 * f) / r0);
        double xz = sqrt(w0 * w0) + c->g1 * sqrt(t0) + sqrt((bsq0 * f) / r0);
        denmn = fmin(denmn, r0);
        velmxx = fmax(velmxx, xx);
        velmxz = fmax(velmxz, xz);
    }
    double dtcflx = 0.16 * c->dd / velmxx;
    double dtcflz = 0.3 * c->dz / velmxz;
    double dtviscx = c->dtvisc0x * denmn;
    double dtviscz = c->dtvisc0z * denmn;
    double dtdiffx = c->dtdiff0x * denmn;
    double dtdiffz = c->dtdiff0z * denmn;
    double mn = fmin(fmin(fmin(dtcflx, dtcflz), fmin(dtviscx, dtviscz)), fmin(dtdiffx, dtdiffz));
    return c->p.sf * mn;
}

/* ------------------------------------------------------------------------ */
/* time loop body, src/main.F90:340-683                                     */
/* ------------------------------------------------------------------------ */
/* AB3 coefficients with the Euler / AB2 start-up, src/main.F90:381-403 */
void orc_ab3_coeffs(int* first, int* second, double dt, double* dt1p, double* dt2p, double c[3])
{
    if (*first) {
        *first = 0;
        *second = 1;
        *dt1p = 0.0;
        *dt2p = 0.0;
        c[0] = dt; c[1] = 0.0; c[2] = 0.0;
    } else if (*second) {
        *first = 0;
        *second = 0;
        *dt2p = 0.0;
        c[0] = dt * (1.0 + 0.5 * (dt / *dt1p));
        c[1] = -0.5 * dt * (dt / *dt1p);
        c[2] = 0.0;
    } else {
        const double dt1 = *dt1p, dt2 = *dt2p;
        c[0] = dt + (((dt * dt) / (dt1 * dt2)) * ((dt / 3.0) + 0.5 * (dt1 + dt2)));
        c[1] = ((dt * dt) / (dt1 * (dt1 - dt2))) * ((dt / 3.0) + 0.5 * dt2);
        c[2] = -((dt * dt) / (dt2 * (dt1 - dt2))) * ((dt / 3.0) + 0.5 * dt1);
    }
}

int orc_step(orc_ctx* c, int nsteps, double dt_fixed, int ndu06, double* diag)
{
    const size_t n3 = c->n3;
    const int nz = c->nz;
    /* src/main.F90:355-375: the field evolves when F != 0 or for the kinematic problem */
    const int mag = (c->p.f != 0.0) || (c->p.opts & ORC_OPT_KINEMATIC);
    double* dots[7] = { c->fld[ORC_UDOT], c->fld[ORC_VDOT], c->fld[ORC_WDOT], c->fld[ORC_RDOT],
                        c->fld[ORC_TDOT], c->fld[ORC_POLDOT], c->fld[ORC_TORDOT] };
    double* vars[7] = { c->fld[ORC_U], c->fld[ORC_V], c->fld[ORC_W], c->fld[ORC_R],
                        c->fld[ORC_T], c->fld[ORC_POL], c->fld[ORC_TOR] };
    if (dt_fixed > 0.0) c->dt = dt_fixed;
    else if (c->dt <= 0.0) c->dt = orc_calcdt(c); /* src/main.F90:318 */

    for (int s = 0; s < nsteps; s++) {
        int e;
        c->nit += 1;
        c->iter += 1;
        c->ttime += c->dt;
        if ((e = orc_hydro_deriv(c))) return e;
        if (mag) {
            if ((e = orc_mag_deriv(c))) return e;
        } else {
            memset(c->fld[ORC_POLDOT] + 2 * n3, 0, sizeof(double) * n3);
            memset(c->fld[ORC_TORDOT] + 2 * n3, 0, sizeof(double) * n3);
            memset(c->fld[ORC_BXBARDOT] + 2 * (size_t)nz, 0, sizeof(double) * (size_t)nz);
            memset(c->fld[ORC_BYBARDOT] + 2 * (size_t)nz, 0, sizeof(double) * (size_t)nz);
        }
        /* 381-403 */
        double dt = c->dt, cf[3];
        orc_ab3_coeffs(&c->first, &c->second, dt, &c->dt1, &c->dt2, cf);
        const double c0 = cf[0], c1 = cf[1], c2 = cf[2];
        /* 413-477 */
        ORC_PAR
        for (size_t q = 0; q < n3; q++) {
            for (int fi = 0; fi < 7; fi++) {
                double* d = dots[fi];
                vars[fi][q] = vars[fi][q] + c0 * d[2 * n3 + q] + c1 * d[n3 + q] + c2 * d[q];
            }
            /* density floor, 438-443 */
            double* rr = c->fld[ORC_R];
            if (rr[q] < 0.05) {
                rr[q] = 0.05;
                double* rd = c->fld[ORC_RDOT];
                rd[q] = 0.0; rd[n3 + q] = 0.0; rd[2 * n3 + q] = 0.0;
            }
            for (int fi = 0; fi < 7; fi++) {
                double* d = dots[fi];
                d[q] = d[n3 + q];
                d[n3 + q] = d[2 * n3 + q];
            }
        }
        for (int k = 0; k < nz; k++) {
            double *bxb = c->fld[ORC_BXBAR], *byb = c->fld[ORC_BYBAR];
            double *bxd = c->fld[ORC_BXBARDOT], *byd = c->fld[ORC_BYBARDOT];
            bxb[k] = bxb[k] + c0 * bxd[2 * nz + k] + c1 * bxd[nz + k] + c2 * bxd[k];
            byb[k] = byb[k] + c0 * byd[2 * nz + k] + c1 * byd[nz + k] + c2 * byd[k];
            bxd[k] = bxd[nz + k]; bxd[nz + k] = bxd[2 * nz + k];
            byd[k] = byd[nz + k]; byd[nz + k] = byd[2 * nz + k];
        }
        /* 488-489 */
        c->dt2 = dt + c->dt1;
        c->dt1 = dt;
        /* 495-510 */
        if (c->nit % 10 == 0) {
            orc_dealias(c, c->fld[ORC_R]);
            orc_dealias(c, c->fld[ORC_T]);
            orc_dealias(c, c->fld[ORC_U]);
            orc_dealias(c, c->fld[ORC_V]);
            orc_dealias(c, c->fld[ORC_W]);
        }
        if (mag) orc_magdealias(c);
        /* 515-521 */
        if (mag && (e = orc_bandj(c))) return e;
        /* 531-538 */
        orc_boundary(c);
        if (mag) orc_mag_boundary(c);
        /* 552-650 */
        if (ndu06 > 0 && c->nit % ndu06 == 0) {
            double ga[4], db[2];
            orc_global_averages(c, ga);
            double nu = orc_nusselt(c);
            if ((e = orc_checkdivb(c, db))) return e;
            for (int id = ORC_WK1; id <= ORC_WK3; id++) memset(c->fld[id], 0, sizeof(double) * n3);
            if (diag) {
                diag[0] = ga[0]; diag[1] = ga[1]; diag[2] = ga[2];
                /* src/main.F90:567-571 */
                diag[3] = (c->p.opts & ORC_OPT_KINEMATIC) ? ga[3] : c->p.f * ga[3];
                diag[4] = nu; diag[5] = db[0]; diag[6] = db[1];
            }
            double divbtest = fmax(fabs(db[0]), fabs(db[1]));
            if (divbtest > 0.01) return 12;
        }
        /* 672-683 */
        if (dt_fixed <= 0.0 && c->nit % 10 == 0) c->dt = orc_calcdt(c);
    }
    return 0;
}

/* threads of the OpenMP loops (1 when built without OpenMP) */
#if defined(_OPENMP)
#include <omp.h>
int orc_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
#else
int orc_threads(void) { return 1; }
void orc_set_threads(int n) { (void)n; }
#endif

/* ------------------------------------------------------------------------ */
/* spectral resampling and box filter                                       */
/* ------------------------------------------------------------------------ */

/* expand, interpolate/fft.F90:832-927: each plane of fin (nx, ny) becomes
 * fout (2nx, 2ny) by zero-padding its spectrum.  Along x: forward r2hc of
 * length nx; the 2nx-long halfcomplex line takes r_0, the r_i and i_i of
 * i = 1..nx/2-1 (at i and 2nx-i) and the coarse Nyquist r_{nx/2} at index
 * nx/2 as an ordinary cosine coefficient (lines 879-887); inverse hc2r of
 * length 2nx divided by nx (889-896).  Then the same along y (898-925). */
int orc_expand(int nx, int ny, int nz, const double* fin, double* fout)
{
    if (nx % 2 || ny % 2 || nx < 2 || ny < 2 || nz < 1) return -1;
    const int nx2 = 2 * nx, ny2 = 2 * ny;
    fftplan *px = get_plan(nx), *px2 = get_plan(nx2), *py = get_plan(ny), *py2 = get_plan(ny2);
    double* mid = (double*)calloc((size_t)nx2 * ny * nz, sizeof(double));
    if (!mid) return -5;
    ORC_PAR
    for (int k = 0; k < nz; k++) {
        double* h = orc_row((size_t)(nx2 > ny2 ? nx2 : ny2), 0);
        double* q = orc_row((size_t)(nx2 > ny2 ? nx2 : ny2), 1);
        for (int j = 0; j < ny; j++) {
            r2hc_strided(px, fin + ((size_t)k * ny + j) * nx, 1, h, 1);
            for (int i = 0; i < nx2; i++) q[i] = 0.0;
            q[0] = h[0];
            for (int i = 1; i <= nx / 2 - 1; i++) {
                q[i] = h[i];
                q[nx2 - i] = h[nx - i];
            }
            q[nx / 2] = h[nx / 2];
            double* o = mid + ((size_t)k * ny + j) * nx2;
            hc2r_strided(px2, q, 1, o, 1);
            for (int i = 0; i < nx2; i++) o[i] = o[i] / (double)nx;
        }
        for (int i = 0; i < nx2; i++) {
            const double* col = mid + (size_t)k * ny * nx2 + i;
            r2hc_strided(py, col, nx2, h, 1);
            for (int j = 0; j < ny2; j++) q[j] = 0.0;
            q[0] = h[0];
            for (int j = 1; j <= ny / 2 - 1; j++) {
                q[j] = h[j];
                q[ny2 - j] = h[ny - j];
            }
            q[ny / 2] = h[ny / 2];
            double* o = fout + (size_t)k * ny2 * nx2 + i;
            hc2r_strided(py2, q, 1, o, nx2);
            for (int j = 0; j < ny2; j++) o[(size_t)j * nx2] = o[(size_t)j * nx2] / (double)ny;
        }
    }
    free(mid);
    return 0;
}

/* Box filter of factor f for coarse-grained data: coarse point (I, J, K) is
 * the mean of the fine block i in [fI, fI+f), j in [fJ, fJ+f), k in [fK,
 * fK+f) of a (nx, ny, nz) field, giving (nx/f, ny/f, nz/f).  Summed over k,
 * then j, then i, and divided by f^3 once. */
int orc_box_filter(int nx, int ny, int nz, int f, const double* fin, double* fout)
{
    if (f < 1 || nx % f || ny % f || nz % f) return -1;
    const int cx = nx / f, cy = ny / f, cz = nz / f;
    const double inv = 1.0 / ((double)f * (double)f * (double)f);
    ORC_PAR
    for (int K = 0; K < cz; K++)
        for (int J = 0; J < cy; J++)
            for (int I = 0; I < cx; I++) {
                double s = 0.0;
                for (int kk = 0; kk < f; kk++)
                    for (int jj = 0; jj < f; jj++)
                        for (int ii = 0; ii < f; ii++)
                            s += fin[((size_t)(f * K + kk) * ny + (f * J + jj)) * nx + (f * I + ii)];
                fout[((size_t)K * cy + J) * cx + I] = s * inv;
            }
    return 0;
}
