// closure.hip -- the learned subgrid closure of coarse runs, on gfx950.
//
// Not part of the reference: a stencil MLP predicts, from the 8 coarse fields
// (r t u v w bx by bz) on the 3x3x3 neighbourhood of a cell, the 12 subgrid
// fluxes that a box filter of a fine run leaves behind (momentum flux tau_ij,
// EMF eps, heat flux q).  Targets come from the reference's block box filter
// of coarse-grained data (interpolate/main.F90:299-306) applied to fine
// states of this solver.
//
// Network: 216 -> 64 -> 64 -> 12, tanh-form GELU, bf16 weights and
// activations, FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16)..+7],
// B[8(l/16)..+7][l%16] and gets D[4(l/16)..+3][l%16]).
// Every product is arranged so both operands are K-contiguous rows, and each
// fragment is one 16-byte load:
//   * the features of a cell are its 8 normalised fields, interleaved as bf16
//     (16 bytes per cell), so a layer-1 A fragment (8 consecutive features =
//     one neighbour's 8 fields) is a single load from the neighbour cell;
//   * weights are kept transposed ([out][in]) for the forward pass and in
//     natural order ([in][out]) for the backward products;
//   * activations are staged in LDS as [cell][unit] (A of the next layer) and
//     [unit][cell] (A of the weight gradients, which reduce over cells).
// Training runs the forward and backward pass of 32 cells per wave in one
// kernel, writes a per-wave gradient, reduces those in a fixed order and
// applies Adam to FP32 master weights, then repacks the bf16 copies.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/mhd.h"
#include "closure.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

namespace mhd {
namespace {

constexpr int NIN = 216, NINP = 224, NH = 64, NOUTP = 16;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NIN * NH;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH * CL_NOUT;
constexpr int NPARAM = OFF_B3 + CL_NOUT;

// LDS leading dimensions in bf16 elements (rows padded by 16 bytes)
constexpr int LD1 = NINP + 8;   // layer-1 weights [64][224]
constexpr int LDH = NH + 8;     // [cell][unit] activations, [out][64] weights
constexpr int LDC = 32 + 8;     // [unit][cell] activations of one wave's 32 cells

constexpr int FWD_NT = 256;     // 4 waves x 32 cells
constexpr int CELLS_PER_WAVE = 32;

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 frag(const __bf16* p) { return *(const bf16x8*)p; }
// A normalised FP64 feature to bf16 in two explicit roundings, each to
// nearest even: FP64 -> FP32, then FP32 -> bf16 on the bits.  The second
// rounding is written out so that a value whose FP32 rounding lands exactly
// on a bf16 tie is resolved the same way everywhere (a direct FP64 -> bf16
// rounding would differ there).
__device__ __forceinline__ __bf16 feature_bf16(double x)
{
    const float f = (float)x;
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __builtin_bit_cast(__bf16, (unsigned short)(u >> 16));
}
// tanh-form GELU: g(z) = z (1 + tanh(k (z + a z^3))) / 2, k = sqrt(2/pi), a = 0.044715.
// With s = 1 / (1 + e^{-2k(z + a z^3)}) (so (1 + tanh) / 2 = s):
//   g(z) = z s,   g'(z) = s + 2 z s (1 - s) k (1 + 3 a z^2)
// e^x is taken as the hardware 2^(x log2 e): GE0, GE1 fold -2 log2(e) into k, k a.
constexpr float GK = 0.7978845608028654f, GKA = 0.035677408136300125f;
constexpr float GE0 = -2.0f * 1.4426950408889634f * 0.7978845608028654f;
constexpr float GE1 = -2.0f * 1.4426950408889634f * 0.035677408136300125f;
__device__ __forceinline__ float gelu_s(float z)
{
    // z -> -inf: 2^(+inf) = inf, s = 0;  z -> +inf: 2^(-inf) = 0, s = 1
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * fmaf(z * z, GE1, GE0)));
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_s(z); }
__device__ __forceinline__ float dgelu(float z)
{
    const float s = gelu_s(z);
    return s + 2.0f * z * s * (1.0f - s) * fmaf(3.0f * GKA, z * z, GK);
}
// orders the wave's own LDS accesses (an area no other wave touches): the
// wave's LDS operations are performed in order, so only the compiler needs fencing
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long x)
{
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}

// ---------------------------------------------------------------------------
// subgrid targets of a fine state: one thread per coarse cell.  The filter F
// is the f x f block mean horizontally and, in z, the coarse-vertex window
// of closure_zweights: coarse plane K (z = K / (cz - 1)) is the mean of the
// fine profile's linear interpolant over [z - H/2, z + H/2], H the coarse
// spacing, and the wall planes are the fine wall planes, so the filtered
// fields keep the walls' values (T = 1 and 1 + theta, w = 0).
// ---------------------------------------------------------------------------
struct TargetArgs {
    ClSource src;            // the fine fields
    void* coarse[CL_NF];     // outputs (float when out_f32, else double)
    void* target[CL_NOUT];
    int out_f32;
    int nx, ny, f, cx, cy;
    long nc;
    double fmag;
    const int* zk0;          // per coarse plane: its first fine plane
    const int* znk;          // and how many fine planes its window covers
    const double* zw;        // their weights, [cz][zmax]
    int zmax;
};

// One coarse cell's filtered fields and fluxes.  Each of the f lanes of a
// cell (LANES = true: lane ii of the cell's f consecutive lanes, f a power of
// two up to 64) sums fine column f I + ii over the cell's f rows and z window,
// so a wave's loads cover 64 consecutive fine columns; the lanes' sums are then
// added by xor butterflies within the cell's lane group.  LANES = false: one
// thread per cell walks all f x f columns (any f).
template <bool LANES>
__global__ void __launch_bounds__(256) targets_kernel(TargetArgs a)
{
    const long g = (long)blockIdx.x * 256 + threadIdx.x;
    const int f = a.f;
    const long q = LANES ? g / f : g;
    const int lane = LANES ? (int)(g - q * f) : 0;
    const bool live = q < a.nc;
    if (!LANES && !live) return;
    const long nxy = (long)a.nx * a.ny;
    double s[CL_NF] = {0, 0, 0, 0, 0, 0, 0, 0};
    double uu[6] = {0, 0, 0, 0, 0, 0}, bb[6] = {0, 0, 0, 0, 0, 0}, uxb[3] = {0, 0, 0}, ut[3] = {0, 0, 0};
    int K = 0, J = 0, I = 0;
    if (live) {
        const long pc = (long)a.cx * a.cy;
        K = (int)(q / pc);
        const long rem = q - (long)K * pc;
        J = (int)(rem / a.cx);
        I = (int)(rem - (long)J * a.cx);
        const int nk = a.znk[K];
        const int i0 = LANES ? lane : 0, i1 = LANES ? lane + 1 : f;
        for (int kk = 0; kk < nk; ++kk) {
            const int k = a.zk0[K] + kk;
            const double wz = a.zw[K * a.zmax + kk];
            for (int jj = 0; jj < f; ++jj)
                for (int ii = i0; ii < i1; ++ii) {
                    const long p = (long)k * nxy + (long)(f * J + jj) * a.nx + (long)(f * I + ii);
                    double v[CL_NF];
#pragma unroll
                    for (int n = 0; n < CL_NF; ++n) {
                        double x = a.src.f32[n] ? (double)((const float*)a.src.p[n])[p] : ((const double*)a.src.p[n])[p];
                        if (a.src.mean[n]) x += a.src.mean[n][k];
                        v[n] = x + a.src.add[n];
                        s[n] += wz * v[n];
                    }
                    const double t = v[1], u = v[2], w2 = v[3], w = v[4], bx = v[5], by = v[6], bz = v[7];
                    uu[0] += wz * (u * u); uu[1] += wz * (w2 * w2); uu[2] += wz * (w * w);
                    uu[3] += wz * (u * w2); uu[4] += wz * (u * w); uu[5] += wz * (w2 * w);
                    bb[0] += wz * (bx * bx); bb[1] += wz * (by * by); bb[2] += wz * (bz * bz);
                    bb[3] += wz * (bx * by); bb[4] += wz * (bx * bz); bb[5] += wz * (by * bz);
                    uxb[0] += wz * (w2 * bz - w * by);
                    uxb[1] += wz * (w * bx - u * bz);
                    uxb[2] += wz * (u * by - w2 * bx);
                    ut[0] += wz * (u * t); ut[1] += wz * (w2 * t); ut[2] += wz * (w * t);
                }
        }
    }
    if constexpr (LANES) {
        // every lane takes part in the butterflies; groups of f lanes are aligned
        for (int o = 1; o < f; o <<= 1) {
#pragma unroll
            for (int n = 0; n < CL_NF; ++n) s[n] += __shfl_xor(s[n], o);
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                uu[c] += __shfl_xor(uu[c], o);
                bb[c] += __shfl_xor(bb[c], o);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                uxb[c] += __shfl_xor(uxb[c], o);
                ut[c] += __shfl_xor(ut[c], o);
            }
        }
        if (!live || lane != 0) return;
    }
    const double inv = 1.0 / ((double)f * f);
    double m[CL_NF];
#pragma unroll
    for (int n = 0; n < CL_NF; ++n) m[n] = s[n] * inv;
    const double r = m[0], T = m[1], U[3] = {m[2], m[3], m[4]}, B[3] = {m[5], m[6], m[7]};
    const int pi[6] = {0, 1, 2, 0, 0, 1}, pj[6] = {0, 1, 2, 1, 2, 2};
    double R[6], M[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        R[c] = uu[c] * inv - U[pi[c]] * U[pj[c]];
        M[c] = bb[c] * inv - B[pi[c]] * B[pj[c]];
    }
    const double trace = M[0] + M[1] + M[2];
    double o[CL_NOUT];
#pragma unroll
    for (int c = 0; c < 6; ++c) o[c] = r * R[c] - a.fmag * (M[c] - (c < 3 ? 0.5 * trace : 0.0));
    o[6] = uxb[0] * inv - (U[1] * B[2] - U[2] * B[1]);
    o[7] = uxb[1] * inv - (U[2] * B[0] - U[0] * B[2]);
    o[8] = uxb[2] * inv - (U[0] * B[1] - U[1] * B[0]);
#pragma unroll
    for (int c = 0; c < 3; ++c) o[9 + c] = ut[c] * inv - U[c] * T;
    if (a.out_f32) {
#pragma unroll
        for (int n = 0; n < CL_NF; ++n) ((float*)a.coarse[n])[q] = (float)m[n];
#pragma unroll
        for (int c = 0; c < CL_NOUT; ++c) ((float*)a.target[c])[q] = (float)o[c];
    } else {
#pragma unroll
        for (int n = 0; n < CL_NF; ++n) ((double*)a.coarse[n])[q] = m[n];
#pragma unroll
        for (int c = 0; c < CL_NOUT; ++c) ((double*)a.target[c])[q] = o[c];
    }
}

// ---------------------------------------------------------------------------
// features: normalised fields of each cell, interleaved bf16
// ---------------------------------------------------------------------------
struct FeatArgs {
    ClSource src;
    int nxy;
    long n;
    double mu[CL_NF], sd[CL_NF];
    bf16x8* grid;
};

__global__ void __launch_bounds__(256) feat_kernel(FeatArgs a)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q >= a.n) return;
    const int k = (int)(q / a.nxy);
    bf16x8 g;
#pragma unroll
    for (int n = 0; n < CL_NF; ++n) {
        double v = a.src.f32[n] ? (double)((const float*)a.src.p[n])[q] : ((const double*)a.src.p[n])[q];
        if (a.src.mean[n]) v += a.src.mean[n][k];
        v += a.src.add[n];
        g[n] = feature_bf16((v - a.mu[n]) / a.sd[n]);
    }
    a.grid[q] = g;
}

// dataset snapshots: float fields and targets -> features and normalised targets
struct PrepArgs {
    const float* raw;      // [snap][20][n3]: 8 fields, then 12 targets
    long n3, n;            // cells per snapshot, all cells
    double mu[CL_NF], sd[CL_NF], tmu[CL_NOUT], tsd[CL_NOUT];
    bf16x8* grid;
    float* tn;             // [snap][12][n3]
};

__global__ void __launch_bounds__(256) prep_kernel(PrepArgs a)
{
    const long g = (long)blockIdx.x * 256 + threadIdx.x;
    if (g >= a.n) return;
    const long snap = g / a.n3, q = g - snap * a.n3;
    const float* r = a.raw + snap * 20 * a.n3 + q;
    bf16x8 v;
#pragma unroll
    for (int n = 0; n < CL_NF; ++n) v[n] = feature_bf16(((double)r[n * a.n3] - a.mu[n]) / a.sd[n]);
    a.grid[g] = v;
#pragma unroll
    for (int o = 0; o < CL_NOUT; ++o)
        a.tn[(snap * CL_NOUT + o) * a.n3 + q] =
            (float)(((double)r[(CL_NF + o) * a.n3] - a.tmu[o]) / a.tsd[o]);
}

// ---------------------------------------------------------------------------
// forward pass: 128 cells per workgroup, 32 per wave (two 16-row blocks)
// ---------------------------------------------------------------------------
struct FwdArgs {
    const bf16x8* grid;
    int nx, ny, nz;
    long n3;               // cells per snapshot
    long n;                // all cells (snapshots back to back)
    const __bf16 *w1t, *w2t, *w3t;
    const float *b1, *b2, *b3;
    double tmu[CL_NOUT], tsd[CL_NOUT];
    void* out[CL_NOUT];    // de-normalised outputs (or null when only the loss is wanted)
    int out_f32;
    const float* tn;       // normalised targets [snap][12][n3] for the loss, or null
    double* lossp;         // per-workgroup sum of squared errors
    int tiled;             // nx and the cell count multiples of 32: neighbourhoods staged in LDS
};

// the 27 neighbour offsets of a cell, from its (i, j, k) once: x and y wrap, z clamps
struct Nbhd {
    long base;                  // snapshot start
    int xm, x0, xp, ym, y0, yp; // x indices; y offsets (j * nx)
    long zm, z0, zp;            // z offsets (k * nx * ny)
    __device__ __forceinline__ Nbhd(long c, long n3, int nx, int ny, int nz)
    {
        const long snap = c / n3;
        base = snap * n3;
        const int cc = (int)(c - base);
        const int nxy = nx * ny;
        const int k = cc / nxy;
        const int rem = cc - k * nxy;
        const int j = rem / nx, i = rem - j * nx;
        xm = i == 0 ? nx - 1 : i - 1;
        x0 = i;
        xp = i == nx - 1 ? 0 : i + 1;
        ym = (j == 0 ? ny - 1 : j - 1) * nx;
        y0 = j * nx;
        yp = (j == ny - 1 ? 0 : j + 1) * nx;
        zm = (long)(k == 0 ? 0 : k - 1) * nxy;
        z0 = (long)k * nxy;
        zp = (long)(k == nz - 1 ? nz - 1 : k + 1) * nxy;
    }
    // neighbour nb = (dz+1) 9 + (dy+1) 3 + (dx+1), nb < 27
    __device__ __forceinline__ long at(int nb) const
    {
        const int dz = nb / 9, r9 = nb - 9 * dz, dy = r9 / 3, dx = r9 - 3 * dy;
        return base + (dz == 0 ? zm : (dz == 1 ? z0 : zp)) + (dy == 0 ? ym : (dy == 1 ? y0 : yp)) +
               (dx == 0 ? xm : (dx == 1 ? x0 : xp));
    }
};

// persistent: each workgroup stages the weights once and walks 128-cell tiles
__global__ void __launch_bounds__(FWD_NT, 2) mlp_fwd_kernel(FwdArgs a)
{
    __shared__ __bf16 sW1[NH * LD1];
    __shared__ __bf16 sW2[NH * LDH];
    __shared__ __bf16 sW3[NOUTP * LDH];
    // per wave: activations [32][LDH], or first the staged neighbourhood of its
    // 32 cells (3 planes x 3 rows x 34 columns of 16-byte cells)
    constexpr int TILE = 3 * 3 * (CELLS_PER_WAVE + 2);
    constexpr int SHW = (TILE * 8 > CELLS_PER_WAVE * LDH) ? TILE * 8 : CELLS_PER_WAVE * LDH;
    __shared__ __attribute__((aligned(16))) __bf16 sH[4][SHW];
    __shared__ float sLoss[4];
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63, lr = l & 15, lg = l >> 4;
    // weights: every load of the thread issued before its LDS stores
    static_assert((NH * NINP / 8) % FWD_NT == 0 && (NH * NH / 8) % FWD_NT == 0 && NOUTP * NH / 8 <= FWD_NT, "");
#pragma unroll
    for (int it = 0; it < NH * NINP / 8 / FWD_NT; ++it) {
        const int e = tid + it * FWD_NT, row = e / (NINP / 8), c8 = e - row * (NINP / 8);
        *(bf16x8*)&sW1[row * LD1 + c8 * 8] = ((const bf16x8*)a.w1t)[e];
    }
#pragma unroll
    for (int it = 0; it < NH * NH / 8 / FWD_NT; ++it) {
        const int e = tid + it * FWD_NT;
        *(bf16x8*)&sW2[(e >> 3) * LDH + (e & 7) * 8] = ((const bf16x8*)a.w2t)[e];
    }
    if (tid < NOUTP * NH / 8) *(bf16x8*)&sW3[(tid >> 3) * LDH + (tid & 7) * 8] = ((const bf16x8*)a.w3t)[tid];
    __syncthreads();

    __bf16* H = sH[wave];
    const bf16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
    const long ntiles = (a.n + 4 * CELLS_PER_WAVE - 1) / (4 * CELLS_PER_WAVE);
    float se = 0.f;
    // staged path (a.tiled: nx a multiple of 32, so a wave's 32 cells are one
    // x-row segment): the 306 cells of their 3x3x3 neighbourhoods, 3 planes x 3
    // rows x 34 columns, go through the wave's LDS area.  Each lane holds up to
    // SPL of them for the wave's next tile, loaded while the current one computes.
    constexpr int SPL = (TILE + 63) / 64;
    bf16x8 pre[SPL];
    auto stage_load = [&](long tl) {
        const long b = tl * (4 * CELLS_PER_WAVE) + wave * CELLS_PER_WAVE;
        const long c0 = b < a.n ? b : a.n - CELLS_PER_WAVE;   // a wave past the end repeats the last cells
        const long sb = (c0 / a.n3) * a.n3;
        const int cc = (int)(c0 - sb), nxy = a.nx * a.ny;
        const int k = cc / nxy, j = (cc - k * nxy) / a.nx, x0 = cc - k * nxy - j * a.nx;
#pragma unroll
        for (int it = 0; it < SPL; ++it) {
            const int e = l + 64 * it;
            if (e < TILE) {
                const int r = e / (CELLS_PER_WAVE + 2), xl = e - r * (CELLS_PER_WAVE + 2);
                const int dz = r / 3, dy = r - 3 * dz;
                int kk = k + dz - 1;
                kk = kk < 0 ? 0 : (kk > a.nz - 1 ? a.nz - 1 : kk);
                int jj = j + dy - 1;
                jj = jj < 0 ? jj + a.ny : (jj >= a.ny ? jj - a.ny : jj);
                int xx = x0 + xl - 1;
                xx = xx < 0 ? xx + a.nx : (xx >= a.nx ? xx - a.nx : xx);
                pre[it] = a.grid[sb + (long)kk * nxy + (long)jj * a.nx + xx];
            }
        }
    };
    if (a.tiled && (long)blockIdx.x < ntiles) stage_load(blockIdx.x);
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const long base = tile * (4 * CELLS_PER_WAVE) + wave * CELLS_PER_WAVE;
        bf16x8 av[2][7];
        if (a.tiled) {
            bf16x8* T = (bf16x8*)H;
#pragma unroll
            for (int it = 0; it < SPL; ++it)
                if (l + 64 * it < TILE) T[l + 64 * it] = pre[it];
            if (tile + gridDim.x < ntiles) stage_load(tile + gridDim.x);
            wave_lds_sync();
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int s = 0; s < 7; ++s) {
                    const int nb = s * 4 + lg;
                    const int dz = nb / 9, r9 = nb - 9 * dz, dy = r9 / 3, dx = r9 - 3 * dy;
                    av[rb][s] = nb < 27 ? T[(dz * 3 + dy) * (CELLS_PER_WAVE + 2) + rb * 16 + lr + dx] : zero8;
                }
            wave_lds_sync();   // the staging area becomes H
        } else {
            // all 14 neighbour loads of the wave's two row blocks first
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) {
                long c = base + rb * 16 + lr;
                if (c > a.n - 1) c = a.n - 1;
                const Nbhd nbh(c, a.n3, a.nx, a.ny, a.nz);
#pragma unroll
                for (int s = 0; s < 7; ++s) {
                    const int nb = s * 4 + lg;
                    av[rb][s] = nb < 27 ? a.grid[nbh.at(nb)] : zero8;
                }
            }
        }
        // transposed products (W^T X^T): lane l gets units 4(l/16)..+3 of cell l%16,
        // so each lane writes 4 consecutive units of one cell with one 8-byte store
        f32x4 acc[2][4];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) acc[0][mb] = acc[1][mb] = *(const f32x4*)&a.b1[mb * 16 + lg * 4];
        // layer 1 (the accumulators start at the bias)
#pragma unroll
        for (int s = 0; s < 7; ++s)
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                const bf16x8 w = frag(&sW1[(mb * 16 + lr) * LD1 + s * 32 + lg * 8]);
                acc[0][mb] = mfma(w, av[0][s], acc[0][mb]);
                acc[1][mb] = mfma(w, av[1][s], acc[1][mb]);
            }
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) {
                bf16x4 h;
#pragma unroll
                for (int i = 0; i < 4; ++i) h[i] = (__bf16)gelu(acc[rb][mb][i]);
                *(bf16x4*)&H[(rb * 16 + lr) * LDH + mb * 16 + lg * 4] = h;
            }
        }
        wave_lds_sync();
        // layer 2
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) acc[0][mb] = acc[1][mb] = *(const f32x4*)&a.b2[mb * 16 + lg * 4];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const bf16x8 h0 = frag(&H[lr * LDH + s * 32 + lg * 8]);
            const bf16x8 h1 = frag(&H[(16 + lr) * LDH + s * 32 + lg * 8]);
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                const bf16x8 w = frag(&sW2[(mb * 16 + lr) * LDH + s * 32 + lg * 8]);
                acc[0][mb] = mfma(w, h0, acc[0][mb]);
                acc[1][mb] = mfma(w, h1, acc[1][mb]);
            }
        }
        wave_lds_sync();
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) {
                bf16x4 h;
#pragma unroll
                for (int i = 0; i < 4; ++i) h[i] = (__bf16)gelu(acc[rb][mb][i]);
                *(bf16x4*)&H[(rb * 16 + lr) * LDH + mb * 16 + lg * 4] = h;
            }
        }
        wave_lds_sync();
        // layer 3: lane l gets outputs 4(l/16)..+3 of cell l%16
        f32x4 y[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int s = 0; s < 2; ++s)
                y[rb] = mfma(frag(&sW3[lr * LDH + s * 32 + lg * 8]), frag(&H[(rb * 16 + lr) * LDH + s * 32 + lg * 8]),
                             y[rb]);
        if (lg < 3) {
            const f32x4 b = *(const f32x4*)&a.b3[lg * 4];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) {
                const long c = base + rb * 16 + lr;
                if (c >= a.n) continue;
                const long snap = a.tn ? c / a.n3 : 0;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int o = lg * 4 + i;
                    const float yn = y[rb][i] + b[i];
                    if (a.out[0]) {
                        const double v = (double)yn * a.tsd[o] + a.tmu[o];
                        if (a.out_f32) ((float*)a.out[o])[c] = (float)v;
                        else ((double*)a.out[o])[c] = v;
                    }
                    if (a.tn) {
                        const float e = yn - a.tn[(snap * CL_NOUT + o) * a.n3 + (c - snap * a.n3)];
                        se += e * e;
                    }
                }
            }
        }
        wave_lds_sync();   // H is rewritten by the next tile
    }
    if (a.tn) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) se += __shfl_xor(se, off);
        if (l == 0) sLoss[wave] = se;
        __syncthreads();
        if (tid == 0) a.lossp[blockIdx.x] = (double)sLoss[0] + sLoss[1] + sLoss[2] + sLoss[3];
    }
}

// ---------------------------------------------------------------------------
// training: forward and backward of 32 cells per wave, one wave per workgroup
// ---------------------------------------------------------------------------
struct TrainArgs {
    const bf16x8* grid;
    const float* tn;
    int nx, ny, nz;
    long n3, nsamples;
    const long* cells;             // explicit sample list, or null: hashed samples
    unsigned long long seed;
    long step;
    int batch;
    const __bf16 *w1t, *w2t, *w3t, *w2n, *w3n;
    const float *b1, *b2, *b3;
    float* partial;                // [wave][NPARAM]
    float* lossp;                  // [wave]
};

__device__ __forceinline__ long sample_id(const TrainArgs& a, long r)
{
    if (a.cells) return a.cells[r < a.batch ? r : a.batch - 1];
    const unsigned long long h =
        mix64(a.seed ^ mix64((unsigned long long)a.step * 0x9E3779B97F4A7C15ULL + (unsigned long long)r));
    return (long)(h % (unsigned long long)a.nsamples);
}

__global__ void __launch_bounds__(64) mlp_train_kernel(TrainArgs a)
{
    __shared__ __bf16 XT[NINP * LDC];          // [feature][cell]
    __shared__ __bf16 S0[NH * LDC];            // H1 [cell][LDH], later dZ1T [unit][LDC]
    __shared__ __bf16 S1[NH * LDC];            // H1T [unit][cell]
    __shared__ __bf16 S2[NH * LDC];            // H2 [cell][LDH], later dZ2 [cell][LDH]
    __shared__ __bf16 S3[NH * LDC];            // H2T, later dZ2T [unit][cell]
    __shared__ __bf16 DY[CELLS_PER_WAVE * LDC];   // dL/dy [cell][32] (outputs 12..31 zero)
    __shared__ __bf16 DYT[NOUTP * LDC];           // [output][cell]
    const int l = threadIdx.x, lr = l & 15, lg = l >> 4;
    const long w = blockIdx.x;
    const long r0 = w * CELLS_PER_WAVE;
    float* P = a.partial + w * (long)NPARAM;
    const bf16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
    const float sc = 2.0f / (12.0f * (float)a.batch);

    // ---- forward, layer 1 (features also stored transposed for dW1)
    f32x4 z1[2][4], z2[2][4];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) z1[rb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        const Nbhd nbh(sample_id(a, r0 + rb * 16 + lr), a.n3, a.nx, a.ny, a.nz);
#pragma unroll
        for (int s = 0; s < 7; ++s) {
            const int nbr = s * 4 + lg;
            const bf16x8 av = nbr < 27 ? a.grid[nbh.at(nbr)] : zero8;
#pragma unroll
            for (int j = 0; j < 8; ++j) XT[(s * 32 + lg * 8 + j) * LDC + rb * 16 + lr] = av[j];
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
                z1[rb][nb] = mfma(av, frag(a.w1t + (nb * 16 + lr) * NINP + s * 32 + lg * 8), z1[rb][nb]);
        }
    }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const int col = nb * 16 + lr;
            const float b = a.b1[col];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = rb * 16 + lg * 4 + i;
                z1[rb][nb][i] += b;
                const __bf16 h = (__bf16)gelu(z1[rb][nb][i]);
                S0[row * LDH + col] = h;
                S1[col * LDC + row] = h;
            }
        }
    __syncthreads();
    // ---- layer 2
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) z2[rb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const bf16x8 av = frag(&S0[(rb * 16 + lr) * LDH + s * 32 + lg * 8]);
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
                z2[rb][nb] = mfma(av, frag(a.w2t + (nb * 16 + lr) * NH + s * 32 + lg * 8), z2[rb][nb]);
        }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const int col = nb * 16 + lr;
            const float b = a.b2[col];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = rb * 16 + lg * 4 + i;
                z2[rb][nb][i] += b;
                const __bf16 h = (__bf16)gelu(z2[rb][nb][i]);
                S2[row * LDH + col] = h;
                S3[col * LDC + row] = h;
            }
        }
    __syncthreads();
    // ---- layer 3 and dL/dy
    f32x4 y[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int s = 0; s < 2; ++s)
            y[rb] = mfma(frag(&S2[(rb * 16 + lr) * LDH + s * 32 + lg * 8]), frag(a.w3t + lr * NH + s * 32 + lg * 8),
                         y[rb]);
    const int o = lr;
    float se = 0.f, db = 0.f;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = rb * 16 + lg * 4 + i;
            float dy = 0.f;
            if (o < CL_NOUT && r0 + row < a.batch) {
                const long sid = sample_id(a, r0 + row);
                const long snap = sid / a.n3;
                const float e = y[rb][i] + a.b3[o] - a.tn[(snap * CL_NOUT + o) * a.n3 + (sid - snap * a.n3)];
                se += e * e;
                dy = sc * e;
            }
            db += dy;
            const __bf16 d = (__bf16)dy;
            DY[row * LDC + o] = d;
            DY[row * LDC + 16 + o] = (__bf16)0.f;
            DYT[o * LDC + row] = d;
        }
    db += __shfl_xor(db, 16);
    db += __shfl_xor(db, 32);
    if (lg == 0 && o < CL_NOUT) P[OFF_B3 + o] = db;
    __syncthreads();
    // ---- dW3 = H2^T dY
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        f32x4 g = mfma(frag(&S3[(mb * 16 + lr) * LDC + lg * 8]), frag(&DYT[lr * LDC + lg * 8]),
                       f32x4{0.f, 0.f, 0.f, 0.f});
        if (o < CL_NOUT)
#pragma unroll
            for (int i = 0; i < 4; ++i) P[OFF_W3 + (mb * 16 + lg * 4 + i) * CL_NOUT + o] = g[i];
    }
    __syncthreads();
    // ---- dZ2 = (dY W3^T) gelu'(z2)
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        const bf16x8 av = frag(&DY[(rb * 16 + lr) * LDC + lg * 8]);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const f32x4 d = mfma(av, frag(a.w3n + (nb * 16 + lr) * 32 + lg * 8), f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int i = 0; i < 4; ++i) z2[rb][nb][i] = d[i] * dgelu(z2[rb][nb][i]);   // now dZ2
        }
    }
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        const int col = nb * 16 + lr;
        float s = 0.f;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = rb * 16 + lg * 4 + i;
                const float v = z2[rb][nb][i];
                s += v;
                const __bf16 h = (__bf16)v;
                S2[row * LDH + col] = h;
                S3[col * LDC + row] = h;
            }
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (lg == 0) P[OFF_B2 + col] = s;
    }
    __syncthreads();
    // ---- dW2 = H1^T dZ2
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        const bf16x8 av = frag(&S1[(mb * 16 + lr) * LDC + lg * 8]);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const f32x4 g = mfma(av, frag(&S3[(nb * 16 + lr) * LDC + lg * 8]), f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int i = 0; i < 4; ++i) P[OFF_W2 + (mb * 16 + lg * 4 + i) * NH + nb * 16 + lr] = g[i];
        }
    }
    // ---- dZ1 = (dZ2 W2^T) gelu'(z1)
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            f32x4 d = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 2; ++s)
                d = mfma(frag(&S2[(rb * 16 + lr) * LDH + s * 32 + lg * 8]),
                         frag(a.w2n + (nb * 16 + lr) * NH + s * 32 + lg * 8), d);
#pragma unroll
            for (int i = 0; i < 4; ++i) z1[rb][nb][i] = d[i] * dgelu(z1[rb][nb][i]);   // now dZ1
        }
    __syncthreads();   // S0 (H1) is free: reused for dZ1T
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        const int col = nb * 16 + lr;
        float s = 0.f;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float v = z1[rb][nb][i];
                s += v;
                S0[col * LDC + rb * 16 + lg * 4 + i] = (__bf16)v;
            }
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (lg == 0) P[OFF_B1 + col] = s;
    }
    __syncthreads();
    // ---- dW1 = X^T dZ1 (feature rows 216..223 are padding)
    bf16x8 bz[4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) bz[nb] = frag(&S0[(nb * 16 + lr) * LDC + lg * 8]);
    for (int mb = 0; mb < NINP / 16; ++mb) {
        const bf16x8 av = frag(&XT[(mb * 16 + lr) * LDC + lg * 8]);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const f32x4 g = mfma(av, bz[nb], f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int f = mb * 16 + lg * 4 + i;
                if (f < NIN) P[OFF_W1 + f * NH + nb * 16 + lr] = g[i];
            }
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) se += __shfl_xor(se, off);
    if (l == 0) a.lossp[w] = se;
}

// fixed-order sum of the per-wave gradients (and of the loss partials)
__global__ void __launch_bounds__(256) grad_reduce_kernel(const float* partial, const float* lossp, long nw,
                                                          float* grad, double* loss)
{
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    // the sums run in wave order; unrolled, so a thread keeps 16 loads in flight
    if (p < NPARAM) {
        double s = 0.0;
#pragma unroll 16
        for (long w = 0; w < nw; ++w) s += (double)partial[w * NPARAM + p];
        grad[p] = (float)s;
    } else if (p == NPARAM) {
        double s = 0.0;
#pragma unroll 16
        for (long w = 0; w < nw; ++w) s += (double)lossp[w];
        *loss = s;
    }
}

struct Packed {
    __bf16 *w1t, *w2t, *w3t, *w2n, *w3n;
    float *b1, *b2, *b3;
};

// one parameter into the bf16 layouts the kernels read
__device__ __forceinline__ void pack_one(const Packed& k, int p, float v)
{
    const __bf16 h = (__bf16)v;
    if (p < OFF_B1) {
        const int n = p / NH, c = p - n * NH;
        k.w1t[c * NINP + n] = h;
    } else if (p < OFF_W2) {
        k.b1[p - OFF_B1] = v;
    } else if (p < OFF_B2) {
        const int q = p - OFF_W2, n = q / NH, c = q - n * NH;
        k.w2t[c * NH + n] = h;
        k.w2n[n * NH + c] = h;
    } else if (p < OFF_W3) {
        k.b2[p - OFF_B2] = v;
    } else if (p < OFF_B3) {
        const int q = p - OFF_W3, n = q / CL_NOUT, c = q - n * CL_NOUT;
        k.w3t[c * NH + n] = h;
        k.w3n[n * 32 + c] = h;
    } else {
        k.b3[p - OFF_B3] = v;
    }
}

__global__ void __launch_bounds__(256) pack_kernel(const float* prm, Packed k)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < NPARAM) pack_one(k, p, prm[p]);
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the FP32 master weights
__global__ void __launch_bounds__(256) adam_kernel(float* prm, float* m, float* v, const float* g, float lr, float c1,
                                                   float c2, Packed k)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NPARAM) return;
    const float gi = g[p];
    const float mi = 0.9f * m[p] + 0.1f * gi;
    const float vi = 0.999f * v[p] + 0.001f * gi * gi;
    m[p] = mi;
    v[p] = vi;
    const float x = prm[p] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    prm[p] = x;
    pack_one(k, p, x);
}

struct TendArgs {
    const void *dx[4], *dy[4], *dz[4];
    const double* r;
    void* dots[4];
    long n;
};

template <typename TD, typename TH>
__global__ void __launch_bounds__(256) tend_kernel(TendArgs a)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q >= a.n) return;
    const double ri = 1.0 / a.r[q];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const double div = ((double)((const TD*)a.dx[m])[q] + (double)((const TD*)a.dy[m])[q]) +
                           (double)((const TD*)a.dz[m])[q];
        TH* d = (TH*)a.dots[m];
        d[q] = (TH)(m < 3 ? (double)d[q] - div * ri : (double)d[q] - div);
    }
}

unsigned nblk(long n, int nt = 256) { return (unsigned)((n + nt - 1) / nt); }

}  // namespace
}  // namespace mhd

using namespace mhd;

struct mhd_closure {
    hipStream_t s = nullptr;
    float *prm = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
    Packed k{};
    double* dloss = nullptr;
    double hloss = 0.0;
    long adam_t = 0;
    unsigned long long seed = 0;
    // gradient partials
    float* partial = nullptr;
    float* lossp = nullptr;
    long pcap = 0, lcap = 0;   // waves
    double* flossp = nullptr;
    long fcap = 0;          // forward workgroups
    // dataset: snapshots of one coarse shape
    int cx = 0, cy = 0, cz = 0;
    long n3 = 0;
    int nsnap = 0, cap = 0;
    float* raw = nullptr;   // [snap][20][n3]
    bf16x8* grid = nullptr; // [snap][n3]
    float* tn = nullptr;    // [snap][12][n3]
    int dirty = 1;
    double sum[20] = {0}, sum2[20] = {0};
    long count = 0;
    double norm[40];
    int norm_user = 0;
    // features of a solver state
    bf16x8* sgrid = nullptr;
    long sgrid_n = 0;
};

namespace {

bool ok(hipError_t e) { return e == hipSuccess; }

// z-weights of the coarse-vertex filter: fine plane k at z_k = k / (nz - 1),
// coarse plane K at Z_K = K / (cz - 1).  The wall planes take the fine wall
// plane; an interior plane takes the mean over [Z_K - H/2, Z_K + H/2]
// (H = 1 / (cz - 1)) of the fine profile's linear interpolant, integrated
// exactly interval by interval, the weights normalised to sum to 1.
// Rows: k0[K], nk[K], weights w[K * zmax + i] of planes k0[K] + i.
void closure_zweights(int nz, int cz, std::vector<int>& k0, std::vector<int>& nk, std::vector<double>& w,
                      int& zmax)
{
    std::vector<std::vector<double>> rows(cz);
    k0.assign(cz, 0);
    for (int K = 0; K < cz; ++K) {
        if (K == 0 || K == cz - 1) {
            k0[K] = K == 0 ? 0 : nz - 1;
            rows[K] = {1.0};
            continue;
        }
        const double Z = (double)K / (double)(cz - 1), half = 0.5 / (double)(cz - 1);
        const double a = Z - half, b = Z + half;
        int kl = (int)std::floor(a * (double)(nz - 1)), kh = (int)std::floor(b * (double)(nz - 1));
        kl = kl < 0 ? 0 : kl;
        kh = kh > nz - 2 ? nz - 2 : kh;
        k0[K] = kl;
        std::vector<double> r(kh - kl + 2, 0.0);
        for (int k = kl; k <= kh; ++k) {
            const double zk = (double)k / (double)(nz - 1), zk1 = (double)(k + 1) / (double)(nz - 1);
            const double p = a > zk ? a : zk, q = b < zk1 ? b : zk1;
            if (!(q > p)) continue;
            const double dk = zk1 - zk;
            const double tp = (p - zk) / dk, tq = (q - zk) / dk;
            r[k - kl] += dk * ((tq - tp) - 0.5 * (tq * tq - tp * tp));
            r[k - kl + 1] += dk * (0.5 * (tq * tq - tp * tp));
        }
        double sum = 0.0;
        for (double x : r) sum += x;
        for (double& x : r) x /= sum;
        rows[K] = r;
    }
    zmax = 1;
    for (const auto& r : rows) zmax = (int)r.size() > zmax ? (int)r.size() : zmax;
    nk.assign(cz, 0);
    w.assign((size_t)cz * zmax, 0.0);
    for (int K = 0; K < cz; ++K) {
        nk[K] = (int)rows[K].size();
        for (int i = 0; i < nk[K]; ++i) w[(size_t)K * zmax + i] = rows[K][i];
    }
}

// the fine fields of src filtered into coarse fields and targets (float when
// out_f32) on stream s; the z-weight tables go to `zmem` (device memory of
// closure_ztable_bytes), copied from the host synchronously
size_t closure_ztable_bytes(int cz, int zmax) { return (size_t)cz * (2 * sizeof(int) + zmax * sizeof(double)); }

int launch_targets(const ClSource& src, int nx, int ny, int nz, int f, double fmag, void* const coarse[CL_NF],
                   void* const target[CL_NOUT], int out_f32, void** zmem, hipStream_t s)
{
    const int cx = nx / f, cy = ny / f, cz = nz / f;
    if (cz < 2) return MHD_E_ARG;
    std::vector<int> k0, nk;
    std::vector<double> w;
    int zmax = 1;
    closure_zweights(nz, cz, k0, nk, w, zmax);
    const size_t zb = closure_ztable_bytes(cz, zmax);
    if (!ok(hipMalloc(zmem, zb))) return MHD_E_ALLOC;
    std::vector<unsigned char> h(zb);
    memcpy(h.data(), w.data(), w.size() * sizeof(double));
    memcpy(h.data() + w.size() * sizeof(double), k0.data(), cz * sizeof(int));
    memcpy(h.data() + w.size() * sizeof(double) + cz * sizeof(int), nk.data(), cz * sizeof(int));
    if (!ok(hipMemcpy(*zmem, h.data(), zb, hipMemcpyHostToDevice))) return MHD_E_DEVICE;
    TargetArgs a;
    memset(&a, 0, sizeof(a));
    a.src = src;
    for (int n = 0; n < CL_NF; ++n) a.coarse[n] = coarse[n];
    for (int o = 0; o < CL_NOUT; ++o) a.target[o] = target[o];
    a.out_f32 = out_f32;
    a.nx = nx; a.ny = ny; a.f = f; a.cx = cx; a.cy = cy;
    a.nc = (long)cx * cy * cz;
    a.fmag = fmag;
    a.zw = (const double*)*zmem;
    a.zk0 = (const int*)((const unsigned char*)*zmem + w.size() * sizeof(double));
    a.znk = a.zk0 + cz;
    a.zmax = zmax;
    if (f <= 64 && 64 % f == 0)
        hipLaunchKernelGGL(targets_kernel<true>, dim3(nblk(a.nc * f)), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(targets_kernel<false>, dim3(nblk(a.nc)), dim3(256), 0, s, a);
    return ok(hipGetLastError()) ? 0 : MHD_E_DEVICE;
}

int ensure(void** p, long* cap, long want, size_t elem)
{
    if (*cap >= want) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (!ok(hipMalloc(p, (size_t)want * elem))) return MHD_E_ALLOC;
    *cap = want;
    return 0;
}

void fit_norm(mhd_closure* c)
{
    if (c->norm_user || c->count == 0) return;
    for (int n = 0; n < 20; ++n) {
        const double mu = c->sum[n] / (double)c->count;
        const double var = std::max(c->sum2[n] / (double)c->count - mu * mu, 0.0);
        const double sd = std::sqrt(var);
        const int slot = n < CL_NF ? n : 16 + (n - CL_NF);
        c->norm[slot] = mu;
        c->norm[slot + (n < CL_NF ? 8 : 12)] = sd > 1e-30 ? sd : 1.0;
    }
}

int prepare(mhd_closure* c)
{
    if (!c->dirty) return 0;
    if (c->nsnap == 0) return MHD_E_STATE;
    fit_norm(c);
    PrepArgs a;
    memset(&a, 0, sizeof(a));
    a.raw = c->raw;
    a.n3 = c->n3;
    a.n = c->n3 * c->nsnap;
    for (int n = 0; n < CL_NF; ++n) { a.mu[n] = c->norm[n]; a.sd[n] = c->norm[8 + n]; }
    for (int o = 0; o < CL_NOUT; ++o) { a.tmu[o] = c->norm[16 + o]; a.tsd[o] = c->norm[28 + o]; }
    a.grid = c->grid;
    a.tn = c->tn;
    hipLaunchKernelGGL(prep_kernel, dim3(nblk(a.n)), dim3(256), 0, c->s, a);
    if (!ok(hipGetLastError())) return MHD_E_DEVICE;
    c->dirty = 0;
    return 0;
}

FwdArgs fwd_args(const mhd_closure* c)
{
    FwdArgs a;
    memset(&a, 0, sizeof(a));
    a.w1t = c->k.w1t; a.w2t = c->k.w2t; a.w3t = c->k.w3t;
    a.b1 = c->k.b1; a.b2 = c->k.b2; a.b3 = c->k.b3;
    for (int o = 0; o < CL_NOUT; ++o) { a.tmu[o] = c->norm[16 + o]; a.tsd[o] = c->norm[28 + o]; }
    return a;
}

// workgroups of the persistent forward grid: two per CU (LDS-bound), at most one per tile
long fwd_grid(long n)
{
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
            ncu = 256;
    }
    const long tiles = (n + 4 * CELLS_PER_WAVE - 1) / (4 * CELLS_PER_WAVE);
    return tiles < 2L * ncu ? tiles : 2L * ncu;
}

int launch_fwd(const FwdArgs& a0, hipStream_t s)
{
    FwdArgs a = a0;
    a.tiled = a.nx % CELLS_PER_WAVE == 0 && a.n % CELLS_PER_WAVE == 0;
    hipLaunchKernelGGL(mlp_fwd_kernel, dim3((unsigned)fwd_grid(a.n)), dim3(FWD_NT), 0, s, a);
    return ok(hipGetLastError()) ? 0 : MHD_E_DEVICE;
}

int run_train(mhd_closure* c, int batch, const long* dcells, long step)
{
    const long nw = (batch + CELLS_PER_WAVE - 1) / CELLS_PER_WAVE;
    int e = ensure((void**)&c->partial, &c->pcap, nw, NPARAM * sizeof(float));
    if (!e) e = ensure((void**)&c->lossp, &c->lcap, nw, sizeof(float));
    if (e) return e;
    TrainArgs a;
    memset(&a, 0, sizeof(a));
    a.grid = c->grid;
    a.tn = c->tn;
    a.nx = c->cx; a.ny = c->cy; a.nz = c->cz;
    a.n3 = c->n3;
    a.nsamples = c->n3 * c->nsnap;
    a.cells = dcells;
    a.seed = c->seed;
    a.step = step;
    a.batch = batch;
    a.w1t = c->k.w1t; a.w2t = c->k.w2t; a.w3t = c->k.w3t; a.w2n = c->k.w2n; a.w3n = c->k.w3n;
    a.b1 = c->k.b1; a.b2 = c->k.b2; a.b3 = c->k.b3;
    a.partial = c->partial;
    a.lossp = c->lossp;
    hipLaunchKernelGGL(mlp_train_kernel, dim3((unsigned)nw), dim3(64), 0, c->s, a);
    hipLaunchKernelGGL(grad_reduce_kernel, dim3(nblk(NPARAM + 1)), dim3(256), 0, c->s, c->partial, c->lossp, nw,
                       c->g, c->dloss);
    return ok(hipGetLastError()) ? 0 : MHD_E_DEVICE;
}

// room for one more snapshot of n3 cells in the data set (the store doubles
// its capacity, keeping its contents)
int grow_store(mhd_closure* c, long n3)
{
    if (c->nsnap == c->cap) {
        // grow the snapshot store (double the capacity, keep the contents)
        const int ncap = c->cap ? 2 * c->cap : 4;
        float *raw = nullptr, *tn = nullptr;
        bf16x8* grid = nullptr;
        if (!ok(hipMalloc((void**)&raw, (size_t)ncap * 20 * n3 * 4)) ||
            !ok(hipMalloc((void**)&tn, (size_t)ncap * 12 * n3 * 4)) ||
            !ok(hipMalloc((void**)&grid, (size_t)ncap * n3 * sizeof(bf16x8)))) {
            for (void* p : {(void*)raw, (void*)tn, (void*)grid})
                if (p) (void)hipFree(p);
            return MHD_E_ALLOC;
        }
        if ((c->nsnap &&
             !ok(hipMemcpyAsync(raw, c->raw, (size_t)c->nsnap * 20 * n3 * 4, hipMemcpyDeviceToDevice, c->s))) ||
            !ok(hipStreamSynchronize(c->s))) {
            for (void* p : {(void*)raw, (void*)tn, (void*)grid}) (void)hipFree(p);
            return MHD_E_DEVICE;
        }
        for (void* p : {(void*)c->raw, (void*)c->tn, (void*)c->grid})
            if (p) (void)hipFree(p);
        c->raw = raw;
        c->tn = tn;
        c->grid = grid;
        c->cap = ncap;
    }
    return 0;
}

// per-block (sum, sum of squares) of the 20 arrays of one snapshot
__global__ void __launch_bounds__(256) snap_stats_kernel(const float* __restrict__ raw, long n3, int nblk_per,
                                                         double* __restrict__ part)
{
    const int arr = blockIdx.x / nblk_per, b = blockIdx.x - arr * nblk_per;
    const long lo = n3 * b / nblk_per, hi = n3 * (b + 1) / nblk_per;
    const float* x = raw + (long)arr * n3;
    double s = 0.0, s2 = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += 256) {
        const double v = x[i];
        s += v;
        s2 += v * v;
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, 64);
        s2 += __shfl_xor(s2, off, 64);
    }
    __shared__ double sh[2][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][wave] = s;
        sh[1][wave] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = ((sh[0][0] + sh[0][1]) + sh[0][2]) + sh[0][3];
        part[2 * blockIdx.x + 1] = ((sh[1][0] + sh[1][1]) + sh[1][2]) + sh[1][3];
    }
}

}  // namespace

namespace mhd {

int closure_add_source(mhd_closure* c, const ClSource& src, int nx, int ny, int nz, int f, double fmag)
{
    if (f < 1 || nx % f || ny % f || nz % f || nz / f < 2) return MHD_E_ARG;
    const int cx = nx / f, cy = ny / f, cz = nz / f;
    if (c->nsnap && (cx != c->cx || cy != c->cy || cz != c->cz)) return MHD_E_ARG;
    const long n3 = (long)cx * cy * cz;
    {
        const int e = grow_store(c, n3);
        if (e) return e;
    }
    float* snap = c->raw + (size_t)c->nsnap * 20 * n3;
    void* co[CL_NF];
    void* tg[CL_NOUT];
    for (int n = 0; n < CL_NF; ++n) co[n] = snap + (size_t)n * n3;
    for (int o = 0; o < CL_NOUT; ++o) tg[o] = snap + (size_t)(CL_NF + o) * n3;
    void* zmem = nullptr;
    double* part = nullptr;
    constexpr int NBLK = 64;
    int rc = launch_targets(src, nx, ny, nz, f, fmag, co, tg, 1, &zmem, c->s);
    if (!rc && !ok(hipMalloc((void**)&part, 20 * NBLK * 2 * sizeof(double)))) rc = MHD_E_ALLOC;
    std::vector<double> hp(20 * NBLK * 2);
    if (!rc) {
        hipLaunchKernelGGL(snap_stats_kernel, dim3(20 * NBLK), dim3(256), 0, c->s, (const float*)snap, n3, NBLK,
                           part);
        if (!ok(hipGetLastError()) ||
            !ok(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, c->s)) ||
            !ok(hipStreamSynchronize(c->s)))
            rc = MHD_E_DEVICE;
    }
    if (zmem) (void)hipFree(zmem);
    if (part) (void)hipFree(part);
    if (rc) return rc;
    // the snapshot is in place: the shape and statistics change now
    c->cx = cx; c->cy = cy; c->cz = cz; c->n3 = n3;
    for (int n = 0; n < 20; ++n) {
        double sm = 0.0, sm2 = 0.0;
        for (int b = 0; b < NBLK; ++b) {
            sm += hp[(n * NBLK + b) * 2];
            sm2 += hp[(n * NBLK + b) * 2 + 1];
        }
        c->sum[n] += sm;
        c->sum2[n] += sm2;
    }
    c->count += n3;
    c->nsnap += 1;
    c->dirty = 1;
    return 0;
}

}  // namespace mhd

namespace {
}  // namespace

namespace mhd {

int closure_fluxes(mhd_closure* cl, const ClSource& src, int nx, int ny, int nz, void* const out[CL_NOUT],
                   int out_f32, void* grid, hipStream_t s)
{
    const long n = (long)nx * ny * nz;
    static_assert(sizeof(bf16x8) == 16, "feature grid element");
    bf16x8* fg = (bf16x8*)grid;
    if (!fg) {
        int e = ensure((void**)&cl->sgrid, &cl->sgrid_n, n, sizeof(bf16x8));
        if (e) return e;
        fg = cl->sgrid;
    }
    FeatArgs f;
    memset(&f, 0, sizeof(f));
    f.src = src;
    f.nxy = nx * ny;
    f.n = n;
    for (int q = 0; q < CL_NF; ++q) { f.mu[q] = cl->norm[q]; f.sd[q] = cl->norm[8 + q]; }
    f.grid = fg;
    hipLaunchKernelGGL(feat_kernel, dim3(nblk(n)), dim3(256), 0, s, f);
    FwdArgs a = fwd_args(cl);
    a.grid = fg;
    a.nx = nx; a.ny = ny; a.nz = nz;
    a.n3 = n;
    a.n = n;
    for (int o = 0; o < CL_NOUT; ++o) a.out[o] = out[o];
    a.out_f32 = out_f32;
    return launch_fwd(a, s);
}

int closure_tend(long n, const void* const dx[4], const void* const dy[4], const void* const dz[4], int f32,
                 const double* r, void* const dots[4], int h32, hipStream_t s)
{
    TendArgs a;
    memset(&a, 0, sizeof(a));
    for (int m = 0; m < 4; ++m) {
        a.dx[m] = dx[m];
        a.dy[m] = dy[m];
        a.dz[m] = dz[m];
        a.dots[m] = dots[m];
    }
    a.r = r;
    a.n = n;
    if (f32 && h32) hipLaunchKernelGGL((tend_kernel<float, float>), dim3(nblk(n)), dim3(256), 0, s, a);
    else if (f32) hipLaunchKernelGGL((tend_kernel<float, double>), dim3(nblk(n)), dim3(256), 0, s, a);
    else if (h32) hipLaunchKernelGGL((tend_kernel<double, float>), dim3(nblk(n)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((tend_kernel<double, double>), dim3(nblk(n)), dim3(256), 0, s, a);
    return ok(hipGetLastError()) ? 0 : MHD_E_DEVICE;
}

}  // namespace mhd

extern "C" {

int mhd_closure_nparam(void) { return NPARAM; }

void mhd_closure_destroy(mhd_closure_t* c)
{
    if (!c) return;
    if (c->s) (void)hipStreamSynchronize(c->s);
    for (void* p : {(void*)c->prm, (void*)c->partial, (void*)c->lossp, (void*)c->flossp, (void*)c->raw,
                    (void*)c->grid, (void*)c->tn, (void*)c->sgrid})
        if (p) (void)hipFree(p);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
}

int mhd_closure_create(unsigned seed, mhd_closure_t** out)
{
    if (!out) return MHD_E_ARG;
    *out = nullptr;
    mhd_closure* c = new mhd_closure();
    c->seed = 0x5eed0000ULL + seed;
    // identity normalisation until data arrive: mu = 0, sd = 1
    for (int i = 0; i < 8; ++i) { c->norm[i] = 0.0; c->norm[8 + i] = 1.0; }
    for (int o = 0; o < 12; ++o) { c->norm[16 + o] = 0.0; c->norm[28 + o] = 1.0; }
    if (!ok(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking))) {
        delete c;
        return MHD_E_DEVICE;
    }
    // one allocation: FP32 master, Adam moments, gradient, bf16 layouts, biases, loss
    const size_t f4 = (size_t)NPARAM * 4;
    const size_t nb16 = (size_t)(NH * NINP + NH * NH + NOUTP * NH + NH * NH + NH * 32) * 2;
    const size_t bytes = 4 * ((f4 + 255) & ~(size_t)255) + ((nb16 + 255) & ~(size_t)255) + 256 * 4 + 256;
    char* pool = nullptr;
    if (!ok(hipMalloc((void**)&pool, bytes))) {
        mhd_closure_destroy(c);
        return MHD_E_ALLOC;
    }
    (void)hipMemsetAsync(pool, 0, bytes, c->s);
    size_t off = 0;
    auto take = [&](size_t b) { char* p = pool + off; off += (b + 255) & ~(size_t)255; return p; };
    c->prm = (float*)take(f4);
    c->m = (float*)take(f4);
    c->v = (float*)take(f4);
    c->g = (float*)take(f4);
    char* w = take(nb16);
    c->k.w1t = (__bf16*)w; w += NH * NINP * 2;
    c->k.w2t = (__bf16*)w; w += NH * NH * 2;
    c->k.w3t = (__bf16*)w; w += NOUTP * NH * 2;
    c->k.w2n = (__bf16*)w; w += NH * NH * 2;
    c->k.w3n = (__bf16*)w;
    c->k.b1 = (float*)take(NH * 4);
    c->k.b2 = (float*)take(NH * 4);
    c->k.b3 = (float*)take(NOUTP * 4);
    c->dloss = (double*)take(8);
    // initial weights: normal with variance 1 / fan-in, zero biases
    std::mt19937_64 rng(c->seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<float> h(NPARAM, 0.0f);
    for (int p = OFF_W1; p < OFF_B1; ++p) h[p] = (float)(nd(rng) / std::sqrt((double)NIN));
    for (int p = OFF_W2; p < OFF_B2; ++p) h[p] = (float)(nd(rng) / std::sqrt((double)NH));
    for (int p = OFF_W3; p < OFF_B3; ++p) h[p] = (float)(nd(rng) / std::sqrt((double)NH));
    *out = c;
    return mhd_closure_set_params(c, h.data(), NPARAM);
}

int mhd_closure_set_params(mhd_closure_t* c, const float* p, int n)
{
    if (!c || !p || n != NPARAM) return MHD_E_ARG;
    if (!ok(hipMemcpyAsync(c->prm, p, (size_t)NPARAM * 4, hipMemcpyHostToDevice, c->s))) return MHD_E_DEVICE;
    (void)hipMemsetAsync(c->m, 0, (size_t)NPARAM * 4, c->s);
    (void)hipMemsetAsync(c->v, 0, (size_t)NPARAM * 4, c->s);
    c->adam_t = 0;
    hipLaunchKernelGGL(pack_kernel, dim3(nblk(NPARAM)), dim3(256), 0, c->s, c->prm, c->k);
    if (!ok(hipGetLastError())) return MHD_E_DEVICE;
    return ok(hipStreamSynchronize(c->s)) ? 0 : MHD_E_DEVICE;
}

int mhd_closure_get_params(mhd_closure_t* c, float* p, int n)
{
    if (!c || !p || n != NPARAM) return MHD_E_ARG;
    if (!ok(hipMemcpyAsync(p, c->prm, (size_t)NPARAM * 4, hipMemcpyDeviceToHost, c->s))) return MHD_E_DEVICE;
    return ok(hipStreamSynchronize(c->s)) ? 0 : MHD_E_DEVICE;
}

int mhd_closure_targets(int nx, int ny, int nz, int f, double fmag, const double* const fine[8],
                        double* const coarse[8], double* const target[12])
{
    if (!fine || !coarse || !target || f < 1 || nx < 1 || ny < 1 || nz < 1) return MHD_E_ARG;
    if (nx % f || ny % f || nz % f) return MHD_E_ARG;
    for (int n = 0; n < 8; ++n)
        if (!fine[n] || !coarse[n]) return MHD_E_ARG;
    for (int o = 0; o < 12; ++o)
        if (!target[o]) return MHD_E_ARG;
    const long n3 = (long)nx * ny * nz;
    const int cx = nx / f, cy = ny / f, cz = nz / f;
    if (cz < 2) return MHD_E_ARG;
    const long nc = (long)cx * cy * cz;
    double* d = nullptr;
    void* zmem = nullptr;
    const size_t bytes = (size_t)(8 * n3 + 20 * nc) * sizeof(double);
    if (!ok(hipMalloc((void**)&d, bytes))) return MHD_E_ALLOC;
    int rc = MHD_E_DEVICE;
    do {
        ClSource src;
        memset(&src, 0, sizeof(src));
        bool up = true;
        for (int n = 0; n < 8 && up; ++n) {
            src.p[n] = d + n * n3;
            up = ok(hipMemcpy((void*)src.p[n], fine[n], (size_t)n3 * 8, hipMemcpyHostToDevice));
        }
        if (!up) break;
        void* co[8];
        void* tg[12];
        for (int n = 0; n < 8; ++n) co[n] = d + 8 * n3 + n * nc;
        for (int o = 0; o < 12; ++o) tg[o] = d + 8 * n3 + (8 + o) * nc;
        const int e = launch_targets(src, nx, ny, nz, f, fmag, co, tg, 0, &zmem, nullptr);
        if (e) { rc = e; break; }
        bool dn = true;
        for (int n = 0; n < 8 && dn; ++n) dn = ok(hipMemcpy(coarse[n], co[n], (size_t)nc * 8, hipMemcpyDeviceToHost));
        for (int o = 0; o < 12 && dn; ++o) dn = ok(hipMemcpy(target[o], tg[o], (size_t)nc * 8, hipMemcpyDeviceToHost));
        if (dn) rc = MHD_OK;
    } while (0);
    (void)hipFree(d);
    if (zmem) (void)hipFree(zmem);
    return rc;
}

int mhd_closure_add(mhd_closure_t* c, int cx, int cy, int cz, const double* const coarse[8],
                    const double* const target[12])
{
    if (!c || !coarse || !target || cx < 1 || cy < 1 || cz < 1) return MHD_E_ARG;
    if (c->nsnap && (cx != c->cx || cy != c->cy || cz != c->cz)) return MHD_E_ARG;
    // every input is checked before any state changes: a rejected call leaves
    // the data set, its statistics and its shape as they were
    for (int n = 0; n < 8; ++n)
        if (!coarse[n]) return MHD_E_ARG;
    for (int o = 0; o < 12; ++o)
        if (!target[o]) return MHD_E_ARG;
    const long n3 = (long)cx * cy * cz;
    {
        const int e = grow_store(c, n3);
        if (e) return e;
    }
    std::vector<float> h((size_t)20 * n3);
    double s[20], s2[20];
    for (int n = 0; n < 20; ++n) {
        const double* src = n < 8 ? coarse[n] : target[n - 8];
        s[n] = s2[n] = 0.0;
        for (long q = 0; q < n3; ++q) {
            const float v = (float)src[q];
            h[(size_t)n * n3 + q] = v;
            s[n] += v;
            s2[n] += (double)v * v;
        }
    }
    if (!ok(hipMemcpyAsync(c->raw + (size_t)c->nsnap * 20 * n3, h.data(), h.size() * 4, hipMemcpyHostToDevice, c->s)))
        return MHD_E_DEVICE;
    if (!ok(hipStreamSynchronize(c->s))) return MHD_E_DEVICE;
    // the snapshot is in place: only now do the shape and statistics change
    c->cx = cx; c->cy = cy; c->cz = cz; c->n3 = n3;
    for (int n = 0; n < 20; ++n) {
        c->sum[n] += s[n];
        c->sum2[n] += s2[n];
    }
    c->count += n3;
    c->nsnap += 1;
    c->dirty = 1;
    return 0;
}

int mhd_closure_samples(mhd_closure_t* c, long* n)
{
    if (!c || !n) return MHD_E_ARG;
    *n = c->n3 * c->nsnap;
    return 0;
}

int mhd_closure_get_norm(mhd_closure_t* c, double norm[40])
{
    if (!c || !norm) return MHD_E_ARG;
    fit_norm(c);
    memcpy(norm, c->norm, sizeof(c->norm));
    return 0;
}

int mhd_closure_set_norm(mhd_closure_t* c, const double norm[40])
{
    if (!c || !norm) return MHD_E_ARG;
    for (int i = 0; i < 8; ++i)
        if (!(norm[8 + i] > 0.0)) return MHD_E_ARG;
    for (int o = 0; o < 12; ++o)
        if (!(norm[28 + o] > 0.0)) return MHD_E_ARG;
    memcpy(c->norm, norm, sizeof(c->norm));
    c->norm_user = 1;
    c->dirty = 1;
    return 0;
}

int mhd_closure_train(mhd_closure_t* c, int steps, int batch, double lr, double* loss)
{
    if (!c || steps < 0 || batch < CELLS_PER_WAVE || batch % CELLS_PER_WAVE || !(lr > 0.0)) return MHD_E_ARG;
    int e = prepare(c);
    if (e) return e;
    for (int it = 0; it < steps; ++it) {
        e = run_train(c, batch, nullptr, c->adam_t);
        if (e) return e;
        c->adam_t += 1;
        const float c1 = 1.0f - (float)std::pow(0.9, (double)c->adam_t);
        const float c2 = 1.0f - (float)std::pow(0.999, (double)c->adam_t);
        hipLaunchKernelGGL(adam_kernel, dim3(nblk(NPARAM)), dim3(256), 0, c->s, c->prm, c->m, c->v, c->g, (float)lr,
                           c1, c2, c->k);
        if (!ok(hipGetLastError())) return MHD_E_DEVICE;
    }
    if (loss) {
        double L = 0.0;
        if (!ok(hipMemcpyAsync(&L, c->dloss, 8, hipMemcpyDeviceToHost, c->s))) return MHD_E_DEVICE;
        if (!ok(hipStreamSynchronize(c->s))) return MHD_E_DEVICE;
        *loss = L / (12.0 * (double)batch);
    }
    return ok(hipStreamSynchronize(c->s)) ? 0 : MHD_E_DEVICE;
}

int mhd_closure_grad(mhd_closure_t* c, int ncell, const long* cells, double* loss, float* grad)
{
    if (!c || ncell < 1 || !cells || !grad) return MHD_E_ARG;
    int e = prepare(c);
    if (e) return e;
    const long ns = c->n3 * c->nsnap;
    for (int i = 0; i < ncell; ++i)
        if (cells[i] < 0 || cells[i] >= ns) return MHD_E_ARG;
    long* dc = nullptr;
    if (!ok(hipMalloc((void**)&dc, (size_t)ncell * sizeof(long)))) return MHD_E_ALLOC;
    int rc = MHD_E_DEVICE;
    if (ok(hipMemcpyAsync(dc, cells, (size_t)ncell * sizeof(long), hipMemcpyHostToDevice, c->s)) &&
        run_train(c, ncell, dc, 0) == 0) {
        double L = 0.0;
        if (ok(hipMemcpyAsync(grad, c->g, (size_t)NPARAM * 4, hipMemcpyDeviceToHost, c->s)) &&
            ok(hipMemcpyAsync(&L, c->dloss, 8, hipMemcpyDeviceToHost, c->s)) && ok(hipStreamSynchronize(c->s))) {
            if (loss) *loss = L / (12.0 * (double)ncell);
            rc = 0;
        }
    }
    (void)hipStreamSynchronize(c->s);
    (void)hipFree(dc);
    return rc;
}

int mhd_closure_eval(mhd_closure_t* c, double* loss)
{
    if (!c || !loss) return MHD_E_ARG;
    int e = prepare(c);
    if (e) return e;
    FwdArgs a = fwd_args(c);
    a.grid = c->grid;
    a.nx = c->cx; a.ny = c->cy; a.nz = c->cz;
    a.n3 = c->n3;
    a.n = c->n3 * c->nsnap;
    a.tn = c->tn;
    const long nwg = fwd_grid(a.n);
    e = ensure((void**)&c->flossp, &c->fcap, nwg, sizeof(double));
    if (e) return e;
    a.lossp = c->flossp;
    e = launch_fwd(a, c->s);
    if (e) return e;
    std::vector<double> h(nwg);
    if (!ok(hipMemcpyAsync(h.data(), c->flossp, (size_t)nwg * 8, hipMemcpyDeviceToHost, c->s)) ||
        !ok(hipStreamSynchronize(c->s)))
        return MHD_E_DEVICE;
    double s = 0.0;
    for (double v : h) s += v;
    *loss = s / (12.0 * (double)a.n);
    return 0;
}

int mhd_closure_predict(mhd_closure_t* c, int nx, int ny, int nz, const double* const coarse[8],
                        double* const out[12])
{
    if (!c || !coarse || !out || nx < 1 || ny < 1 || nz < 1) return MHD_E_ARG;
    fit_norm(c);
    const long n = (long)nx * ny * nz;
    double* d = nullptr;
    if (!ok(hipMalloc((void**)&d, (size_t)20 * n * 8))) return MHD_E_ALLOC;
    int rc = MHD_E_DEVICE;
    do {
        ClSource src;
        memset(&src, 0, sizeof(src));
        bool up = true;
        for (int q = 0; q < 8 && up; ++q) {
            if (!coarse[q]) { rc = MHD_E_ARG; up = false; break; }
            src.p[q] = d + q * n;
            up = ok(hipMemcpyAsync((void*)src.p[q], coarse[q], (size_t)n * 8, hipMemcpyHostToDevice, c->s));
        }
        if (!up) break;
        void* o[CL_NOUT];
        for (int q = 0; q < CL_NOUT; ++q) o[q] = d + (8 + q) * n;
        if (closure_fluxes(c, src, nx, ny, nz, o, 0, nullptr, c->s)) break;
        bool dn = true;
        for (int q = 0; q < CL_NOUT && dn; ++q) {
            if (!out[q]) { rc = MHD_E_ARG; dn = false; break; }
            dn = ok(hipMemcpyAsync(out[q], o[q], (size_t)n * 8, hipMemcpyDeviceToHost, c->s));
        }
        if (dn && ok(hipStreamSynchronize(c->s))) rc = 0;
    } while (0);
    (void)hipStreamSynchronize(c->s);
    (void)hipFree(d);
    return rc;
}

// kernel times on the dataset: one forward pass over every sample and one
// training step of `batch` samples, averaged over `reps` launches each
int mhd_closure_time(mhd_closure_t* c, int batch, int reps, double* ms_fwd, double* ms_train, long* n_fwd)
{
    if (!c || reps < 1 || batch < CELLS_PER_WAVE || batch % CELLS_PER_WAVE) return MHD_E_ARG;
    int e = prepare(c);
    if (e) return e;
    FwdArgs a = fwd_args(c);
    a.grid = c->grid;
    a.nx = c->cx; a.ny = c->cy; a.nz = c->cz;
    a.n3 = c->n3;
    a.n = c->n3 * c->nsnap;
    const long n = a.n;
    float* o = nullptr;
    if (!ok(hipMalloc((void**)&o, (size_t)12 * n * 4))) return MHD_E_ALLOC;
    for (int q = 0; q < CL_NOUT; ++q) a.out[q] = o + (long)q * n;
    a.out_f32 = 1;
    hipEvent_t ev[4];
    for (auto& x : ev) (void)hipEventCreate(&x);
    int rc = MHD_E_DEVICE;
    do {
        if (launch_fwd(a, c->s)) break;   // warm-up
        (void)hipEventRecord(ev[0], c->s);
        bool good = true;
        for (int r = 0; r < reps && good; ++r) good = launch_fwd(a, c->s) == 0;
        (void)hipEventRecord(ev[1], c->s);
        if (!good || run_train(c, batch, nullptr, 0)) break;
        (void)hipEventRecord(ev[2], c->s);
        for (int r = 0; r < reps && good; ++r) good = run_train(c, batch, nullptr, r) == 0;
        (void)hipEventRecord(ev[3], c->s);
        if (!good || !ok(hipEventSynchronize(ev[3]))) break;
        float t0 = 0.f, t1 = 0.f;
        (void)hipEventElapsedTime(&t0, ev[0], ev[1]);
        (void)hipEventElapsedTime(&t1, ev[2], ev[3]);
        if (ms_fwd) *ms_fwd = t0 / reps;
        if (ms_train) *ms_train = t1 / reps;
        if (n_fwd) *n_fwd = n;
        rc = 0;
    } while (0);
    (void)hipStreamSynchronize(c->s);
    for (auto& x : ev) (void)hipEventDestroy(x);
    (void)hipFree(o);
    return rc;
}

}  // extern "C"
