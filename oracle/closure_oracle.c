/*
 * closure_oracle.c -- CPU oracle of the learned subgrid closure.
 *
 * TEST INFRASTRUCTURE ONLY, like mhd_oracle.c: it is the checker the HIP
 * closure kernels are compared against and is never linked into the product.
 * The closure has no counterpart in the reference; what it builds on is the
 * reference's box filter of coarse-grained data (interpolate/main.F90:299-306,
 * restated in orc_box_filter) and its field order (src/netcdf.F90: R, T, U, V,
 * W, Bx, By, Bz).
 *
 * Definitions (F = the filter of orc_closure_targets: the f x f block mean
 * horizontally, a window centred on each coarse plane in z; fields r t u v w
 * bx by bz):
 *   R_ij = F(u_i u_j) - F(u_i) F(u_j),  M_ij = F(B_i B_j) - F(B_i) F(B_j)
 *   tau_ij = F(r) R_ij - fmag (M_ij - delta_ij (M_xx + M_yy + M_zz) / 2)
 *            order xx yy zz xy xz yz                         (targets 0..5)
 *   eps = F(u x B) - F(u) x F(B)                             (targets 6..8)
 *   q_i = F(u_i T) - F(u_i) F(T)                             (targets 9..11)
 *
 * Network: 216 inputs (the 8 normalised fields on the 3x3x3 neighbourhood,
 * feature = nb * 8 + field, nb = (dz+1) 9 + (dy+1) 3 + (dx+1), x and y
 * periodic, z clamped to the walls), two hidden layers of 64 tanh-form GELU
 * units, 12 outputs (normalised targets).  Parameters (packed):
 *   W1[216][64] b1[64] W2[64][64] b2[64] W3[64][12] b3[12]  (18828 values)
 * With round != 0 the products use bf16 (round to nearest even) operands at
 * the points where the GPU kernels round: the features, the weights, the
 * hidden activations, dL/dy and the pre-activation gradients; sums and the
 * biases stay wide.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mhd_oracle.h"

#define NF 8
#define NIN 216
#define NH 64
#define NOUT 12
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + NIN * NH)
#define OFF_W2 (OFF_B1 + NH)
#define OFF_B2 (OFF_W2 + NH * NH)
#define OFF_W3 (OFF_B2 + NH)
#define OFF_B3 (OFF_W3 + NH * NOUT)
#define NPARAM (OFF_B3 + NOUT)

int orc_closure_nparam(void) { return NPARAM; }

/* float -> bf16 -> float, round to nearest even (NaN not expected here) */
float orc_bf16(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    u &= 0xffff0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

static double rnd(double x, int round) { return round ? (double)orc_bf16((float)x) : x; }

/* tanh-form GELU, z (1 + tanh(k (z + a z^3))) / 2 with k = sqrt(2/pi), a = 0.044715 */
#define GK 0.7978845608028654
#define GKA 0.035677408136300125
static double gelu(double z) { return 0.5 * z * (1.0 + tanh(z * (GK + GKA * z * z))); }
static double dgelu(double z)
{
    const double t = tanh(z * (GK + GKA * z * z));
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * (GK + 3.0 * GKA * z * z);
}

/* The filter of the closure data: the f x f block mean horizontally; in z,
 * centred on the coarse planes Z_K = K / (cz - 1): the fine wall planes at
 * K = 0 and cz - 1, elsewhere the mean over [Z_K - H/2, Z_K + H/2]
 * (H = 1 / (cz - 1)) of the linear interpolant of the fine profile (fine
 * planes at z_k = k / (nz - 1)).  orc_zweight returns the weight of fine
 * plane k in coarse plane K, integrating the interpolant's hat function of
 * plane k over the window, normalised so a plane's weights sum to 1. */
static double zwin_raw(int nz, int cz, int K, int k)
{
    const double Z = (double)K / (double)(cz - 1), half = 0.5 / (double)(cz - 1);
    const double a = Z - half, b = Z + half;
    double w = 0.0;
    /* the hat of plane k spans [z_{k-1}, z_{k+1}]: its left half rises on
     * interval k-1, its right half falls on interval k */
    for (int side = 0; side < 2; side++) {
        const int i = side == 0 ? k - 1 : k;          /* interval [z_i, z_{i+1}] */
        if (i < 0 || i > nz - 2) continue;
        const double zi = (double)i / (double)(nz - 1), zi1 = (double)(i + 1) / (double)(nz - 1);
        const double p = a > zi ? a : zi, q = b < zi1 ? b : zi1;
        if (!(q > p)) continue;
        const double d = zi1 - zi, tp = (p - zi) / d, tq = (q - zi) / d;
        /* integral over [p, q] of t (rising side) or 1 - t (falling side) */
        const double rise = d * 0.5 * (tq * tq - tp * tp);
        w += side == 0 ? rise : d * (tq - tp) - rise;
    }
    return w;
}

double orc_zweight(int nz, int cz, int K, int k)
{
    if (K == 0) return k == 0 ? 1.0 : 0.0;
    if (K == cz - 1) return k == nz - 1 ? 1.0 : 0.0;
    double sum = 0.0;
    for (int j = 0; j < nz; j++) sum += zwin_raw(nz, cz, K, j);
    return zwin_raw(nz, cz, K, k) / sum;
}

int orc_closure_targets(int nx, int ny, int nz, int f, double fmag, const double* const fine[8],
                        double* const coarse[8], double* const target[12])
{
    if (f < 1 || nx % f || ny % f || nz % f || nz / f < 2) return -1;
    const int cx = nx / f, cy = ny / f, cz = nz / f;
    const double inv = 1.0 / ((double)f * f);
    double* wz = (double*)malloc(sizeof(double) * (size_t)nz);
    if (!wz) return -5;
    for (int K = 0; K < cz; K++) {
        for (int k = 0; k < nz; k++) wz[k] = orc_zweight(nz, cz, K, k);
        for (int J = 0; J < cy; J++)
            for (int I = 0; I < cx; I++) {
                double s[8] = {0}, uu[6] = {0}, bb[6] = {0}, uxb[3] = {0}, ut[3] = {0};
                for (int k = 0; k < nz; k++) {
                    const double w = wz[k];
                    if (w == 0.0) continue;
                    for (int jj = 0; jj < f; jj++)
                        for (int ii = 0; ii < f; ii++) {
                            const size_t q = ((size_t)k * ny + (f * J + jj)) * nx + (f * I + ii);
                            double v[8];
                            for (int n = 0; n < 8; n++) {
                                v[n] = fine[n][q];
                                s[n] += w * v[n];
                            }
                            const double u = v[2], vv = v[3], ww = v[4], bx = v[5], by = v[6], bz = v[7], t = v[1];
                            uu[0] += w * (u * u); uu[1] += w * (vv * vv); uu[2] += w * (ww * ww);
                            uu[3] += w * (u * vv); uu[4] += w * (u * ww); uu[5] += w * (vv * ww);
                            bb[0] += w * (bx * bx); bb[1] += w * (by * by); bb[2] += w * (bz * bz);
                            bb[3] += w * (bx * by); bb[4] += w * (bx * bz); bb[5] += w * (by * bz);
                            uxb[0] += w * (vv * bz - ww * by);
                            uxb[1] += w * (ww * bx - u * bz);
                            uxb[2] += w * (u * by - vv * bx);
                            ut[0] += w * (u * t); ut[1] += w * (vv * t); ut[2] += w * (ww * t);
                        }
                }
                double m[8];
                for (int n = 0; n < 8; n++) m[n] = s[n] * inv;
                const double r = m[0], T = m[1], U[3] = {m[2], m[3], m[4]}, B[3] = {m[5], m[6], m[7]};
                static const int pi[6] = {0, 1, 2, 0, 0, 1}, pj[6] = {0, 1, 2, 1, 2, 2};
                double R[6], M[6];
                for (int c = 0; c < 6; c++) {
                    R[c] = uu[c] * inv - U[pi[c]] * U[pj[c]];
                    M[c] = bb[c] * inv - B[pi[c]] * B[pj[c]];
                }
                const double trace = M[0] + M[1] + M[2];
                const size_t o = ((size_t)K * cy + J) * cx + I;
                for (int n = 0; n < 8; n++) coarse[n][o] = m[n];
                for (int c = 0; c < 6; c++) target[c][o] = r * R[c] - fmag * (M[c] - (c < 3 ? 0.5 * trace : 0.0));
                target[6][o] = uxb[0] * inv - (U[1] * B[2] - U[2] * B[1]);
                target[7][o] = uxb[1] * inv - (U[2] * B[0] - U[0] * B[2]);
                target[8][o] = uxb[2] * inv - (U[0] * B[1] - U[1] * B[0]);
                for (int c = 0; c < 3; c++) target[9 + c][o] = ut[c] * inv - U[c] * T;
            }
    }
    free(wz);
    return 0;
}

/* normalised, optionally bf16-rounded features of one coarse cell */
static void features(int nx, int ny, int nz, const double* const fld[8], const double norm[40], size_t q,
                     int round, double* x)
{
    const int i = (int)(q % (size_t)nx), j = (int)((q / (size_t)nx) % (size_t)ny), k = (int)(q / ((size_t)nx * ny));
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int nb = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
                int kk = k + dz;
                kk = kk < 0 ? 0 : (kk > nz - 1 ? nz - 1 : kk);
                const int jj = (j + dy + ny) % ny, ii = (i + dx + nx) % nx;
                const size_t p = ((size_t)kk * ny + jj) * nx + ii;
                for (int n = 0; n < NF; n++) {
                    /* (x - mu) / sd in double, then to float, then bf16 */
                    const float v = (float)((fld[n][p] - norm[n]) / norm[8 + n]);
                    x[nb * NF + n] = round ? (double)orc_bf16(v) : (double)v;
                }
            }
}

typedef struct {
    double z1[NH], h1[NH], z2[NH], h2[NH], y[NOUT];
} act_t;

static void forward(const double* x, const double* p, int round, act_t* a)
{
    for (int h = 0; h < NH; h++) {
        double s = 0.0;
        for (int n = 0; n < NIN; n++) s += x[n] * rnd(p[OFF_W1 + n * NH + h], round);
        a->z1[h] = s + p[OFF_B1 + h];
        a->h1[h] = rnd(gelu(a->z1[h]), round);
    }
    for (int h = 0; h < NH; h++) {
        double s = 0.0;
        for (int n = 0; n < NH; n++) s += a->h1[n] * rnd(p[OFF_W2 + n * NH + h], round);
        a->z2[h] = s + p[OFF_B2 + h];
        a->h2[h] = rnd(gelu(a->z2[h]), round);
    }
    for (int o = 0; o < NOUT; o++) {
        double s = 0.0;
        for (int n = 0; n < NH; n++) s += a->h2[n] * rnd(p[OFF_W3 + n * NOUT + o], round);
        a->y[o] = s + p[OFF_B3 + o];
    }
}

int orc_mlp_forward(int nx, int ny, int nz, const double* const coarse[8], const double* params,
                    const double norm[40], int round, double* const out[12])
{
    const size_t n3 = (size_t)nx * ny * nz;
    double x[NIN];
    act_t a;
    for (size_t q = 0; q < n3; q++) {
        features(nx, ny, nz, coarse, norm, q, round, x);
        forward(x, params, round, &a);
        for (int o = 0; o < NOUT; o++) out[o][q] = a.y[o] * norm[28 + o] + norm[16 + o];
    }
    return 0;
}

/* Loss L = sum over the listed cells and the 12 outputs of (y - t)^2 / (12 n)
 * with t the normalised target, and its gradient with respect to the packed
 * parameters.  cells[c] = snapshot * (nx ny nz) + cell; coarse holds 8
 * pointers per snapshot, target 12 per snapshot. */
int orc_mlp_grad(int nx, int ny, int nz, const double* const* coarse, const double* const* target,
                 const double norm[40], const double* params, int ncell, const long* cells, int round,
                 double* loss, double* grad)
{
    const size_t n3 = (size_t)nx * ny * nz;
    memset(grad, 0, sizeof(double) * NPARAM);
    double L = 0.0;
    const double sc = 2.0 / (12.0 * (double)ncell);
    double x[NIN];
    act_t a;
    for (int c = 0; c < ncell; c++) {
        const size_t snap = (size_t)cells[c] / n3, q = (size_t)cells[c] % n3;
        features(nx, ny, nz, coarse + 8 * snap, norm, q, round, x);
        forward(x, params, round, &a);
        double dy[NOUT], dyr[NOUT];
        for (int o = 0; o < NOUT; o++) {
            const double t = (double)(float)((target[12 * snap + o][q] - norm[16 + o]) / norm[28 + o]);
            const double e = a.y[o] - t;
            L += e * e;
            dy[o] = sc * e;
            dyr[o] = rnd(dy[o], round);
            grad[OFF_B3 + o] += dy[o];
        }
        double dz2[NH], dz2r[NH];
        for (int h = 0; h < NH; h++) {
            double s = 0.0;
            for (int o = 0; o < NOUT; o++) {
                s += dyr[o] * rnd(params[OFF_W3 + h * NOUT + o], round);
                grad[OFF_W3 + h * NOUT + o] += a.h2[h] * dyr[o];
            }
            dz2[h] = s * dgelu(a.z2[h]);
            dz2r[h] = rnd(dz2[h], round);
            grad[OFF_B2 + h] += dz2[h];
        }
        double dz1[NH], dz1r[NH];
        for (int h = 0; h < NH; h++) {
            double s = 0.0;
            for (int n = 0; n < NH; n++) {
                s += dz2r[n] * rnd(params[OFF_W2 + h * NH + n], round);
                grad[OFF_W2 + h * NH + n] += a.h1[h] * dz2r[n];
            }
            dz1[h] = s * dgelu(a.z1[h]);
            dz1r[h] = rnd(dz1[h], round);
            grad[OFF_B1 + h] += dz1[h];
        }
        for (int n = 0; n < NIN; n++)
            for (int h = 0; h < NH; h++) grad[OFF_W1 + n * NH + h] += x[n] * dz1r[h];
    }
    *loss = L / (12.0 * (double)ncell);
    return 0;
}
