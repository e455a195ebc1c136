#pragma once
// spectral_impl.hpp -- horizontal transform kernels for gfx950 (template
// definitions; spectral_g*.hip instantiate them per transform length).
//
// One kernel template covers the four horizontal operator families of the
// reference (src/fft.F90): spectral derivatives (d1bydx/d1bydy/d2bydx/d2bydy),
// the 2/3 dealias, and the two halves of the 2-D real transforms
// (forwardfft: r2hc per line; inversefft: hc2r per line).
//
// A workgroup owns 2*TPW real lines (rows for dir=x, columns for dir=y) of one
// job and packs line pairs (a, b) into TPW complex lines a + i b held in LDS.
// Because every operator here multiplies each mode by a Hermitian-symmetric
// factor m(N-k) = conj(m(k)), one complex FFT, one multiply and one inverse
// FFT return the operator applied to a in the real part and to b in the
// imaginary part; real<->halfcomplex conversions separate / merge the pair
// explicitly.  Each line is read once from HBM and each output written once.
//
// Rows (dir=x) are contiguous, so a workgroup reads 2*TPW*N consecutive
// elements.  Columns (dir=y) are read as tiles of 2*TPW adjacent columns,
// i.e. 2*TPW*sizeof(T)-byte segments per row.
#include "fft_lds.hpp"
#include "spectral.hpp"

namespace mhd {

template <typename TC>
__device__ __forceinline__ TC load_job(const SpecJob& J, long i, int k)
{
    if (J.load_op == LD_PLAIN) {
        if (J.bg) return (TC)(ldv<double>(J.in[0], i, J.in_f32) - J.bg[k]);
        return ldv<TC>(J.in[0], i, J.in_f32);
    } else if (J.load_op == LD_PROD) {
        TC a = ldv<TC>(J.in[0], i, J.in_f32), b = ldv<TC>(J.in[1], i, J.in_f32);
        return ((TC)J.scale * a) * b;
    } else {
        TC a = ldv<TC>(J.in[0], i, J.in_f32), b = ldv<TC>(J.in[1], i, J.in_f32);
        TC c = ldv<TC>(J.in[2], i, J.in_f32), d = ldv<TC>(J.in[3], i, J.in_f32);
        return a * b - c * d;
    }
}

// spectral multiplier of mode k (full complex spectrum index), folded with 1/N
template <typename TC, int N>
__device__ __forceinline__ cpx<TC> apply_mul(int kind, cpx<TC> z, int k, const double* __restrict__ ak1,
                                             const double* __restrict__ ak2, int icut)
{
    constexpr double invN = 1.0 / N;
    if (kind == MUL_D1) {
        if (k == 0 || k == N / 2) return {(TC)0, (TC)0};
        const TC a = (TC)(ak1[k] * invN);
        // +i*a for k < N/2; the negative frequency N-k carries -i*a
        if (k < N / 2) return {-a * z.y, a * z.x};
        return {a * z.y, -a * z.x};
    } else if (kind == MUL_D2) {
        const TC a = (TC)(ak2[k] * invN);
        return {a * z.x, a * z.y};
    } else {
        const int m = (k < N - k) ? k : N - k;
        if (m < N / 2 - icut) return {z.x * (TC)invN, z.y * (TC)invN};
        return {(TC)0, (TC)0};
    }
}

template <int DIR, int N, int TPW>
struct LineMap {
    // element e of a workgroup's block -> (line in WG, position n, global index, plane, valid)
    static __device__ __forceinline__ void map(const Geom& g, int e, int& line, int& n, long& gi, int& k,
                                               bool& ok)
    {
        constexpr int LINES = 2 * TPW;
        if constexpr (DIR == 0) {
            line = e / N;
            n = e - line * N;
            const long gl = (long)blockIdx.x * LINES + line;
            ok = gl < (long)g.nz * g.ny;
            gi = gl * N + n;
            k = (int)(gl / g.ny);
        } else {
            n = e / LINES;
            line = e - n * LINES;
            const int tpp = g.nx / LINES;
            k = blockIdx.x / tpp;
            const int x0 = (blockIdx.x - k * tpp) * LINES;
            gi = (long)k * g.nxy + (long)n * g.nx + x0 + line;
            ok = true;
        }
    }
};

template <typename TC, int N, int DIR, int TPW, int MODE, int NT>
__global__ void __launch_bounds__(NT) spec_kernel(SpecLaunch L)
{
    constexpr int NOUT = (MODE == SM_DERIV2) ? 2 : 1;
    constexpr int S = (DIR == 0) ? N : N + 1;
    constexpr int LINES = 2 * TPW;
    constexpr int NE = LINES * N;
    constexpr int NBIN = TPW * N;
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<TC>* buf = reinterpret_cast<cpx<TC>*>(smem);
    TC* rbuf = reinterpret_cast<TC*>(smem);
    const SpecJob& J = L.job[blockIdx.y];
    const Geom g = L.g;
    const int tid = threadIdx.x;
    const cpx<TC>* __restrict__ tw = reinterpret_cast<const cpx<TC>*>(L.tw);

    // ---- load 2*TPW real lines into TPW complex lines --------------------
#pragma unroll 4
    for (int e = tid; e < NE; e += NT) {
        int line, n, k;
        long gi;
        bool ok;
        LineMap<DIR, N, TPW>::map(g, e, line, n, gi, k, ok);
        TC v = (TC)0;
        if (ok) v = (MODE == SM_HC2R) ? ldv<TC>(J.in[0], gi, J.in_f32) : load_job<TC>(J, gi, k);
        rbuf[((line >> 1) * S + n) * 2 + (line & 1)] = v;
    }
    __syncthreads();

    if constexpr (MODE == SM_HC2R) {
        // merge two halfcomplex lines (a in .x, b in .y) into Z = A + iB
        constexpr int PER = (NBIN + NT - 1) / NT;
        cpx<TC> z[PER];
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int b = tid + p * NT;
            if (NBIN % NT == 0 || b < NBIN) {
                const int t = b / N, kk = b - (b / N) * N;
                const cpx<TC> u = buf[t * S + kk];
                const cpx<TC> m = buf[t * S + ((N - kk) % N)];
                cpx<TC> A, B;
                if (kk == 0 || kk == N / 2) { A = {u.x, (TC)0}; B = {u.y, (TC)0}; }
                else if (kk < N / 2) { A = {u.x, m.x}; B = {u.y, m.y}; }
                else { A = {m.x, -u.x}; B = {m.y, -u.y}; }
                z[p] = {A.x - B.y, A.y + B.x};
            }
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int b = tid + p * NT;
            if (NBIN % NT == 0 || b < NBIN) buf[(b / N) * S + (b - (b / N) * N)] = z[p];
        }
        __syncthreads();
        LdsFFT<TC, N, TPW, NT, S, true>::run(buf, tw, tid);
    } else {
        LdsFFT<TC, N, TPW, NT, S, false>::run(buf, tw, tid);
    }

    if constexpr (MODE == SM_DERIV1 || MODE == SM_DERIV2) {
        for (int b = tid; b < NBIN; b += NT) {
            const int t = b / N, kk = b - (b / N) * N;
            const cpx<TC> z = buf[t * S + kk];
            if constexpr (NOUT == 2) buf[TPW * S + t * S + kk] = apply_mul<TC, N>(J.mul[1], z, kk, L.ak1, L.ak2, L.icut);
            buf[t * S + kk] = apply_mul<TC, N>(J.mul[0], z, kk, L.ak1, L.ak2, L.icut);
        }
        __syncthreads();
        LdsFFT<TC, N, NOUT * TPW, NT, S, true>::run(buf, tw, tid);
    }

    // ---- store -----------------------------------------------------------
#pragma unroll 2
    for (int e = tid; e < NE; e += NT) {
        int line, n, k;
        long gi;
        bool ok;
        LineMap<DIR, N, TPW>::map(g, e, line, n, gi, k, ok);
        if (!ok) continue;
        if constexpr (MODE == SM_R2HC) {
            const int t = line >> 1;
            const int kk = (n <= N / 2) ? n : N - n;
            const cpx<TC> zk = buf[t * S + kk];
            const cpx<TC> zn = buf[t * S + ((N - kk) % N)];
            TC v;
            if ((line & 1) == 0) v = (n <= N / 2) ? (TC)0.5 * (zk.x + zn.x) : (TC)0.5 * (zk.y - zn.y);
            else v = (n <= N / 2) ? (TC)0.5 * (zk.y + zn.y) : (TC)0.5 * (zn.x - zk.x);
            stv<TC>(J.out[0], gi, J.out_f32, v);
        } else {
#pragma unroll
            for (int o = 0; o < NOUT; ++o) {
                TC v = rbuf[((o * TPW + (line >> 1)) * S + n) * 2 + (line & 1)];
                if constexpr (MODE == SM_HC2R) v = v * (TC)(1.0 / N);
                if (J.addk) v = v + (TC)J.addk[k];
                if (J.adds != 0.0) v = v + (TC)J.adds;
                stv<TC>(J.out[o], gi, J.out_f32, v);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host dispatch
// ---------------------------------------------------------------------------
template <int N>
constexpr int x_tpw() { return (1024 / N) < 2 ? 2 : (1024 / N) > 16 ? 16 : (1024 / N); }

template <typename TC, int N, int DIR, int TPW, int MODE>
static int launch_one(const SpecLaunch& L, hipStream_t s)
{
    constexpr int NT = 256;
    constexpr int NOUT = (MODE == SM_DERIV2) ? 2 : 1;
    constexpr int S = (DIR == 0) ? N : N + 1;
    const size_t lds = (size_t)NOUT * TPW * S * sizeof(cpx<TC>);
    auto kern = spec_kernel<TC, N, DIR, TPW, MODE, NT>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return -21;
    }
    dim3 grid;
    if (DIR == 0) grid.x = (unsigned)(((long)L.g.nz * L.g.ny + 2 * TPW - 1) / (2 * TPW));
    else grid.x = (unsigned)(L.g.nz * (L.g.nx / (2 * TPW)));
    grid.y = L.njobs;
    hipLaunchKernelGGL(kern, grid, dim3(NT), lds, s, L);
    return hipGetLastError() == hipSuccess ? 0 : -22;
}

template <typename TC, int N, int DIR, int TPW>
static int launch_mode(const SpecLaunch& L, hipStream_t s)
{
    switch (L.mode) {
    case SM_DERIV1: return launch_one<TC, N, DIR, TPW, SM_DERIV1>(L, s);
    case SM_DERIV2: return launch_one<TC, N, DIR, TPW, SM_DERIV2>(L, s);
    case SM_R2HC: return launch_one<TC, N, DIR, TPW, SM_R2HC>(L, s);
    case SM_HC2R: return launch_one<TC, N, DIR, TPW, SM_HC2R>(L, s);
    }
    return -23;
}

// Column tiles: the widest tile (2*TPW columns) whose LDS footprint fits
// 160 KB and whose width divides nx.  FP64 tiles are 16 columns (128 B per
// row), FP32 tiles 32 columns.
template <typename TC, int N>
constexpr int y_tpw_max()
{
    constexpr int cap = sizeof(TC) == 8 ? 8 : 16;
    constexpr int fit = (160 * 1024) / (2 * (N + 1) * (int)sizeof(cpx<TC>));
    return fit >= cap ? cap : fit >= 8 ? 8 : fit >= 4 ? 4 : fit >= 2 ? 2 : 1;
}

template <typename TC, int N>
static int launch_n(const SpecLaunch& L, hipStream_t s)
{
    if (L.dir == 0) return launch_mode<TC, N, 0, x_tpw<N>()>(L, s);
    constexpr int T0 = y_tpw_max<TC, N>();
    if (L.g.nx % (2 * T0) == 0) return launch_mode<TC, N, 1, T0>(L, s);
    if constexpr (T0 > 4) {
        if (L.g.nx % 8 == 0) return launch_mode<TC, N, 1, 4>(L, s);
    }
    return launch_mode<TC, N, 1, 1>(L, s);
}

}  // namespace mhd

// Defines launch_spectral_<group>(n, L, f32, s) for the lengths in SIZES.
#define MHD_SPECTRAL_GROUP(name, SIZES)                                        \
    namespace mhd {                                                            \
    int name(int n, const SpecLaunch& L, int f32, hipStream_t s)               \
    {                                                                          \
        switch (n) {                                                           \
            SIZES(MHD_SPEC_CASE)                                               \
        }                                                                      \
        return -20;                                                            \
    }                                                                          \
    }
#define MHD_SPEC_CASE(n_)                                                      \
    case n_:                                                                   \
        return f32 ? launch_n<float, n_>(L, s) : launch_n<double, n_>(L, s);
