// fft_reg.hpp -- register-blocked complex FFTs for gfx950.
//
// A transform of length N = R_0 R_1 ... R_{P-1} runs as P self-sorting
// (Stockham) passes.  In every pass a thread owns one radix-R butterfly with
// all R points in registers; between passes the points move through LDS
// once (write, barrier, read, barrier).  Pass 0 takes its inputs
// x[j + r*N/R_0] straight from memory and the last pass leaves
// X[j + r*N/R_{P-1}] in registers, so a spectral multiply and the inverse
// transform -- run with the radix plan reversed, whose first pass consumes
// exactly those registers -- need no LDS round trip between them, and the
// inverse's last pass again holds x[j + r*N/R_0], ready to store.
//
// LDS indices are swizzled (pad()), which keeps the stride-R writes of the
// passes on distinct banks.
#pragma once

#include "fft_lds.hpp"

namespace mhd {

// radix plans (forward order); every radix is one of 2, 3, 4, 5, 8, 16
template <int N>
struct FftPlan;
#define MHD_PLAN(N_, ...)                                        \
    template <>                                                  \
    struct FftPlan<N_> {                                         \
        static constexpr int N = N_;                             \
        static constexpr int R[] = {__VA_ARGS__};                \
        static constexpr int P = sizeof(R) / sizeof(int);        \
    };
// Radix-4-heavy plans: with the pipelined kernel, small butterflies (few
// registers, high occupancy) beat radix-16 ones although they add LDS passes
// (measured at 256x256x160 FP64: x d1+d2 243 us with 4x4x4x4 vs 265 us with
// 16x16; x d1 143 us vs 162 us).
MHD_PLAN(8, 8)
MHD_PLAN(12, 4, 3)
MHD_PLAN(16, 4, 4)
MHD_PLAN(20, 4, 5)
MHD_PLAN(24, 8, 3)
MHD_PLAN(30, 2, 3, 5)
MHD_PLAN(32, 4, 8)
MHD_PLAN(40, 8, 5)
MHD_PLAN(48, 4, 4, 3)
MHD_PLAN(60, 4, 3, 5)
MHD_PLAN(64, 4, 4, 4)
MHD_PLAN(80, 4, 4, 5)
MHD_PLAN(96, 4, 8, 3)
MHD_PLAN(120, 8, 3, 5)
MHD_PLAN(128, 4, 4, 8)
MHD_PLAN(160, 8, 4, 5)
MHD_PLAN(192, 4, 4, 4, 3)
MHD_PLAN(240, 4, 4, 3, 5)
MHD_PLAN(256, 4, 4, 4, 4)
MHD_PLAN(320, 4, 4, 4, 5)
MHD_PLAN(384, 4, 4, 8, 3)
MHD_PLAN(480, 8, 4, 3, 5)
MHD_PLAN(512, 4, 4, 4, 8)
MHD_PLAN(640, 4, 4, 8, 5)
MHD_PLAN(768, 4, 4, 4, 4, 3)
MHD_PLAN(1024, 4, 4, 4, 4, 4)
#undef MHD_PLAN

// An explicit plan (the autotuner's alternatives use these).
template <int N_, int... Rs>
struct Plan {
    static constexpr int N = N_;
    static constexpr int R[] = {Rs...};
    static constexpr int P = sizeof...(Rs);
};

// Alternative plan per length for the autotuner (type void: none).  Each
// trades LDS passes against butterfly size (registers, occupancy).
template <int N>
struct AltPlan {
    using type = void;
};
template <> struct AltPlan<64> { using type = Plan<64, 8, 8>; };
template <> struct AltPlan<128> { using type = Plan<128, 8, 16>; };
template <> struct AltPlan<160> { using type = Plan<160, 4, 8, 5>; };
template <> struct AltPlan<256> { using type = Plan<256, 4, 8, 8>; };
template <> struct AltPlan<512> { using type = Plan<512, 8, 8, 8>; };
// a second alternative (SpecLaunch::variant = 2): two radix-16 passes, 16
// threads per transform, so a workgroup of 16 FP32 column pairs has 256
// threads instead of 1024 and several fit on a CU
template <int N>
struct AltPlan2 {
    using type = void;
};
template <> struct AltPlan2<256> { using type = Plan<256, 16, 16>; };

template <typename PL>
constexpr int plan_radix(int p, bool inv)
{
    return inv ? PL::R[PL::P - 1 - p] : PL::R[p];
}
// product of the radices before pass p
template <typename PL>
constexpr int plan_ns(int p, bool inv)
{
    int ns = 1;
    for (int q = 0; q < p; ++q) ns *= plan_radix<PL>(q, inv);
    return ns;
}
// threads per transform: the largest butterfly count of any pass
template <typename PL>
constexpr int plan_tpt()
{
    int t = 1;
    for (int p = 0; p < PL::P; ++p)
        if (PL::N / PL::R[p] > t) t = PL::N / PL::R[p];
    return t;
}
template <typename PL>
constexpr int plan_rmax()
{
    int r = 1;
    for (int p = 0; p < PL::P; ++p)
        if (PL::R[p] > r) r = PL::R[p];
    return r;
}

// W_8^k and W_16^k twiddle multiplies (forward = exp(-i...))
template <typename T, bool INV>
__device__ __forceinline__ cpx<T> rot(cpx<T> a, T c, T s)
{
    // a * (c - i s) forward, a * (c + i s) inverse
    if (INV) return {a.x * c - a.y * s, a.y * c + a.x * s};
    return {a.x * c + a.y * s, a.y * c - a.x * s};
}

template <typename T, bool INV>
__device__ __forceinline__ void dft8(cpx<T>* v)
{
    cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
    butterfly<T, 4, INV>(e);
    butterfly<T, 4, INV>(o);
    const T h = (T)0.70710678118654752440084436210485;
    o[1] = rot<T, INV>(o[1], h, h);
    o[2] = cmul_i<T, INV ? 1 : -1>(o[2]);
    o[3] = rot<T, INV>(o[3], -h, h);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = cadd(e[k], o[k]);
        v[k + 4] = csub(e[k], o[k]);
    }
}

template <typename T, bool INV>
__device__ __forceinline__ void dft16(cpx<T>* v)
{
    cpx<T> e[8], o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        e[k] = v[2 * k];
        o[k] = v[2 * k + 1];
    }
    dft8<T, INV>(e);
    dft8<T, INV>(o);
    const T c1 = (T)0.92387953251128675612818318939679;  // cos(pi/8)
    const T s1 = (T)0.38268343236508977172845998403040;  // sin(pi/8)
    const T h = (T)0.70710678118654752440084436210485;
    o[1] = rot<T, INV>(o[1], c1, s1);
    o[2] = rot<T, INV>(o[2], h, h);
    o[3] = rot<T, INV>(o[3], s1, c1);
    o[4] = cmul_i<T, INV ? 1 : -1>(o[4]);
    o[5] = rot<T, INV>(o[5], -s1, c1);
    o[6] = rot<T, INV>(o[6], -h, h);
    o[7] = rot<T, INV>(o[7], -c1, s1);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        v[k] = cadd(e[k], o[k]);
        v[k + 8] = csub(e[k], o[k]);
    }
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft_r(cpx<T>* v)
{
    if constexpr (R == 8) dft8<T, INV>(v);
    else if constexpr (R == 16) dft16<T, INV>(v);
    else butterfly<T, R, INV>(v);
}

// WAVE: every transform's butterflies lie within one wavefront (and each
// transform has its own LDS area), so the exchanges only need the wave's own
// LDS operations ordered, not a workgroup barrier.
template <typename T, typename PL, bool WAVE = false>
struct RegFFT {
    static constexpr int N = PL::N;
    static constexpr int P = PL::P;
    static constexpr int TPT = plan_tpt<PL>();
    static constexpr int RMAX = plan_rmax<PL>();
    static constexpr int R0 = PL::R[0];
    static constexpr int RL = PL::R[P - 1];
    static constexpr int NB0 = N / R0;   // butterflies of the first forward pass
    static constexpr int NBL = N / RL;   // butterflies of the last forward pass
    static constexpr int PADSH = sizeof(T) == 8 ? 4 : 5;   // one pad per 256 bytes
    // LDS stride per transform
    static constexpr int SP = (N % 16 == 0) ? N + 6 : N + (N >> PADSH) + 1;

    // LDS slot of element i of a transform.  Lengths that are multiples of 16
    // XOR the low four index bits with 5 (i >> 4): the exchanges' strided writes
    // (stride Ns = 1, 4, 16, ... of the radix-4 and -8 passes) then fall on
    // distinct banks within a 16-lane group, and contiguous reads stay
    // conflict-free; other lengths pad one element per 256 bytes.
    // orders the LDS accesses around an exchange: a workgroup barrier, or, for
    // wave-local transforms, a compiler fence (a wave's LDS operations are
    // performed in order)
    static __device__ __forceinline__ void sync()
    {
        if constexpr (WAVE) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            __syncthreads();
        }
    }
    static __device__ __forceinline__ int pad(int i)
    {
        if constexpr (N % 16 == 0) return i ^ (((i >> 4) * 5) & 15);
        else return i + (i >> PADSH);
    }

    // Twiddles in LDS, one table per pass: pass p (of the forward plan, or of
    // the reversed plan for the inverse) keeps W_N^(r kk N/(Ns R)) for its
    // butterflies' kk = j mod Ns at entry kk (R-1) + r-1 of its table.  The
    // lanes of a wave then read entries R-1 apart; read from one table of
    // W_N^i they were r N/(Ns R) apart, on few LDS banks.  A plan equal to its
    // reverse shares the tables of both directions.
    static constexpr int tw_count(int p, bool inv)
    {
        return plan_ns<PL>(p, inv) > 1 ? plan_ns<PL>(p, inv) * (plan_radix<PL>(p, inv) - 1) : 0;
    }
    static constexpr bool palindrome()
    {
        for (int p = 0; p < P; ++p)
            if (plan_radix<PL>(p, false) != plan_radix<PL>(p, true)) return false;
        return true;
    }
    static constexpr int tw_off(int p, bool inv)
    {
        int o = 0;
        if (inv && !palindrome())
            for (int q = 0; q < P; ++q) o += tw_count(q, false);
        for (int q = 0; q < p; ++q) o += tw_count(q, inv);
        return o;
    }
    static constexpr int TWN = tw_off(P, !palindrome());   // table entries in all
    template <bool INV, int p>
    static __device__ __forceinline__ void stage_pass(cpx<T>* tw, const cpx<T>* __restrict__ twg, int tid, int nt)
    {
        constexpr int R = plan_radix<PL>(p, INV);
        constexpr int Ns = plan_ns<PL>(p, INV);
        if constexpr (Ns > 1) {
            constexpr int off = tw_off(p, INV), cnt = Ns * (R - 1), st = N / (Ns * R);
            for (int e = tid; e < cnt; e += nt) {
                const int kk = e / (R - 1), r = e - kk * (R - 1) + 1;
                tw[off + e] = twg[r * kk * st];
            }
        }
        if constexpr (p + 1 < P) stage_pass<INV, p + 1>(tw, twg, tid, nt);
    }
    // fill the tables (TWN entries) from W_N^i (i < N) in global memory
    static __device__ __forceinline__ void stage_twiddles(cpx<T>* tw, const cpx<T>* __restrict__ twg, int tid,
                                                          int nt)
    {
        stage_pass<false, 0>(tw, twg, tid, nt);
        if constexpr (!palindrome()) stage_pass<true, 0>(tw, twg, tid, nt);
    }

    // butterfly of pass p on v (thread's butterfly index j)
    template <int p, bool INV>
    static __device__ __forceinline__ void pass(cpx<T>* v, int j, const cpx<T>* __restrict__ tw)
    {
        constexpr int R = plan_radix<PL>(p, INV);
        constexpr int Ns = plan_ns<PL>(p, INV);
        if constexpr (Ns > 1) {
            const int kk = j % Ns;
            const cpx<T>* __restrict__ twp = tw + tw_off(p, INV) + kk * (R - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const cpx<T> w = twp[r - 1];
                v[r] = INV ? cmulc(v[r], w) : cmul(v[r], w);
            }
        }
        dft_r<T, R, INV>(v);
    }

    // outputs of pass p -> LDS -> inputs of pass p+1
    template <int p, bool INV>
    static __device__ __forceinline__ void exchange(cpx<T>* v, int j, cpx<T>* lds)
    {
        constexpr int R = plan_radix<PL>(p, INV);
        constexpr int Ns = plan_ns<PL>(p, INV);
        constexpr int NB = N / R;
        if (j < NB) {
            const int base = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
            for (int r = 0; r < R; ++r) lds[pad(base + r * Ns)] = v[r];
        }
        sync();
        constexpr int R2 = plan_radix<PL>(p + 1, INV);
        constexpr int NB2 = N / R2;
        if (j < NB2) {
#pragma unroll
            for (int r = 0; r < R2; ++r) v[r] = lds[pad(j + r * NB2)];
        }
        sync();
    }

    template <int p, bool INV>
    static __device__ __forceinline__ void run_from(cpx<T>* v, int j, cpx<T>* lds, const cpx<T>* __restrict__ tw)
    {
        if constexpr (p < P) {
            constexpr int NB = N / plan_radix<PL>(p, INV);
            if (j < NB) pass<p, INV>(v, j, tw);
            if constexpr (p + 1 < P) exchange<p, INV>(v, j, lds);
            run_from<p + 1, INV>(v, j, lds, tw);
        }
    }

    // forward: v holds x[j + r*NB0] (j < NB0) -> X[j + r*NBL] (j < NBL)
    static __device__ __forceinline__ void forward(cpx<T>* v, int j, cpx<T>* lds, const cpx<T>* __restrict__ tw)
    {
        run_from<0, false>(v, j, lds, tw);
    }
    // backward (unnormalised): v holds X[j + r*NBL] -> x[j + r*NB0]
    static __device__ __forceinline__ void inverse(cpx<T>* v, int j, cpx<T>* lds, const cpx<T>* __restrict__ tw)
    {
        run_from<0, true>(v, j, lds, tw);
    }
};

}  // namespace mhd
