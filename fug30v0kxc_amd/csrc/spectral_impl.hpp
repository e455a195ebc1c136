#pragma once
// spectral_impl.hpp -- horizontal transform kernels for gfx950 (template
// definitions; spectral_g*.hip instantiate them per transform length).
//
// One kernel template covers the four horizontal operator families of the
// reference (src/fft.F90): spectral derivatives (d1bydx/d1bydy/d2bydx/d2bydy),
// the 2/3 dealias, and the two halves of the 2-D real transforms
// (forwardfft: r2hc per line; inversefft: hc2r per line).
//
// A workgroup owns B transforms of one job; each transform packs two real
// lines (rows for dir=x, adjacent columns for dir=y) as a + i b.  Because
// every operator here multiplies each mode by a Hermitian-symmetric factor
// m(N-k) = conj(m(k)), one complex FFT, one multiply and one inverse FFT
// return the operator applied to a in the real part and to b in the
// imaginary part; real<->halfcomplex conversions separate / merge the pair
// through one extra LDS round trip.  Each line is read once from HBM and each
// output written once, straight from registers (see fft_reg.hpp).
//
// Rows (dir=x): the TPT threads of a transform read TPT consecutive elements
// of each of its two rows per load.  Columns (dir=y): consecutive threads
// belong to consecutive column pairs, so one load reads 2*B adjacent columns
// of one row (2*B*sizeof(T) bytes).
#include <type_traits>

#include "fft_reg.hpp"
#include "launch.hpp"
#include "spectral.hpp"

namespace mhd {

// Adjacent pairs (i, i+1) with i even: the two columns of a y transform.
// One 8-byte (float) or 16-byte (double) access instead of two.
template <typename TC>
using vec2_t = typename std::conditional<sizeof(TC) == 4, float2, double2>::type;
template <typename TC>
__device__ __forceinline__ cpx<TC> ld_pair(const TC* __restrict__ p, long i)
{
    const vec2_t<TC> v = *reinterpret_cast<const vec2_t<TC>*>(p + i);
    return {(TC)v.x, (TC)v.y};
}
template <typename TC>
__device__ __forceinline__ void st_pair(TC* __restrict__ p, long i, TC a, TC b)
{
    vec2_t<TC> v;
    v.x = a;
    v.y = b;
    *reinterpret_cast<vec2_t<TC>*>(p + i) = v;
}
// derivative outputs are read back only by later kernels, after the step has
// streamed other fields through the caches: stored non-temporally
template <typename TC>
__device__ __forceinline__ void st_pair_nt(TC* __restrict__ p, long i, TC a, TC b)
{
    using v2 = TC __attribute__((ext_vector_type(2)));
    v2 v;
    v.x = a;
    v.y = b;
    __builtin_nontemporal_store(v, reinterpret_cast<v2*>(p + i));
}
template <typename TC>
__device__ __forceinline__ void st_nt(TC* __restrict__ p, long i, TC a)
{
    __builtin_nontemporal_store(a, p + i);
}

// Input loaders, fixed at compile time per launch.  Inputs have the compute
// type TC, except LD_BG, which reads the FP64 state and subtracts a per-plane
// background in FP64 before rounding to TC.
template <typename TC, int LOP>
struct Loader {
    const TC* __restrict__ a;
    const TC* __restrict__ b;
    const TC* __restrict__ c;
    const TC* __restrict__ d;
    const double* __restrict__ s64;
    TC scale;
    double bgk;
    TC m0, m1, m2, m3;   // per-plane means of the factors (0 when absent)
    const TC* __restrict__ jj;   // LD_CROSS: current density component, or null
    TC jc;                       // its per-plane coefficient
    __device__ __forceinline__ Loader(const SpecJob& J, int k)
    {
        a = (const TC*)J.in[0];
        b = (const TC*)J.in[1];
        c = (const TC*)J.in[2];
        d = (const TC*)J.in[3];
        s64 = (const double*)J.in[0];
        scale = (TC)J.scale;
        bgk = (LOP == LD_BG && J.bg) ? J.bg[k] : 0.0;
        m0 = m1 = m2 = m3 = (TC)0;
        jj = nullptr;
        jc = (TC)0;
        if constexpr (LOP == LD_CROSS) {
            if (J.jin) {
                jj = (const TC*)J.jin;
                jc = (TC)J.jcoef[k];
            }
        }
        if constexpr (LOP == LD_PROD || LOP == LD_CROSS || LOP == LD_MAXW) {
            if (J.mean[0]) m0 = (TC)J.mean[0][k];
            if (J.mean[1]) m1 = (TC)J.mean[1][k];
            if (J.mean[2]) m2 = (TC)J.mean[2][k];
            if (J.mean[3]) m3 = (TC)J.mean[3][k];
        }
    }
    // elements i and i+1 (i even)
    __device__ __forceinline__ cpx<TC> pair(long i) const
    {
        if constexpr (LOP == LD_PLAIN) return ld_pair(a, i);
        else if constexpr (LOP == LD_BG) {
            const cpx<double> v = ld_pair(s64, i);
            return {(TC)(v.x - bgk), (TC)(v.y - bgk)};
        } else if constexpr (LOP == LD_PROD) {
            const cpx<TC> x = ld_pair(a, i), y = ld_pair(b, i);
            return {(scale * x.x) * y.x + scale * (x.x * m1 + m0 * y.x),
                    (scale * x.y) * y.y + scale * (x.y * m1 + m0 * y.y)};
        } else if constexpr (LOP == LD_MAXW) {
            const cpx<TC> x = ld_pair(a, i), y = ld_pair(b, i), z = ld_pair(c, i);
            return {scale * x.x * x.x - scale * y.x * y.x - scale * z.x * z.x +
                        (TC)2 * scale * (m0 * x.x - m1 * y.x - m2 * z.x),
                    scale * x.y * x.y - scale * y.y * y.y - scale * z.y * z.y +
                        (TC)2 * scale * (m0 * x.y - m1 * y.y - m2 * z.y)};
        } else {
            return cross_pair(ld_pair(a, i), ld_pair(b, i), ld_pair(c, i), ld_pair(d, i), i);
        }
    }
    // PLAIN, PROD or MAXW from factor values already loaded (x = in0, y = in1,
    // z = in2; both elements of a pair), the arithmetic of pair()
    __device__ __forceinline__ cpx<TC> combine(cpx<TC> x, cpx<TC> y, cpx<TC> z) const
    {
        if constexpr (LOP == LD_PROD) {
            return {(scale * x.x) * y.x + scale * (x.x * m1 + m0 * y.x),
                    (scale * x.y) * y.y + scale * (x.y * m1 + m0 * y.y)};
        } else if constexpr (LOP == LD_MAXW) {
            return {scale * x.x * x.x - scale * y.x * y.x - scale * z.x * z.x +
                        (TC)2 * scale * (m0 * x.x - m1 * y.x - m2 * z.x),
                    scale * x.y * x.y - scale * y.y * y.y - scale * z.y * z.y +
                        (TC)2 * scale * (m0 * x.y - m1 * y.y - m2 * z.y)};
        } else {
            (void)y;
            (void)z;
            return x;
        }
    }
    // LD_CROSS from factor values already loaded (x = in0, y = in1, z = in2,
    // w = in3 at element i), the extra term loaded here
    __device__ __forceinline__ TC cross(TC x, TC y, TC z, TC w, long i) const
    {
        TC r = x * y - z * w + (x * m1 + m0 * y) - (z * m3 + m2 * w);
        if (jj) r = r - jc * jj[i];
        return r;
    }
    __device__ __forceinline__ cpx<TC> cross_pair(cpx<TC> x, cpx<TC> y, cpx<TC> z, cpx<TC> w, long i) const
    {
        cpx<TC> r = {x.x * y.x - z.x * w.x + (x.x * m1 + m0 * y.x) - (z.x * m3 + m2 * w.x),
                     x.y * y.y - z.y * w.y + (x.y * m1 + m0 * y.y) - (z.y * m3 + m2 * w.y)};
        if (jj) {
            const cpx<TC> q = ld_pair(jj, i);
            r.x = r.x - jc * q.x;
            r.y = r.y - jc * q.y;
        }
        return r;
    }
    __device__ __forceinline__ TC operator()(long i) const
    {
        if constexpr (LOP == LD_PLAIN) return a[i];
        else if constexpr (LOP == LD_BG) return (TC)(s64[i] - bgk);
        else if constexpr (LOP == LD_PROD) {
            // (s*a)*b is the reference's product; the mean terms drop its constant
            const TC x = a[i], y = b[i];
            return (scale * x) * y + scale * (x * m1 + m0 * y);
        } else if constexpr (LOP == LD_MAXW) {
            const TC x = a[i], y = b[i], z = c[i];
            return scale * x * x - scale * y * y - scale * z * z +
                   (TC)2 * scale * (m0 * x - m1 * y - m2 * z);
        } else {
            return cross(a[i], b[i], c[i], d[i], i);
        }
    }
};

// spectral multiplier of mode k (full complex spectrum index), folded with
// 1/N.  The wavenumbers are those of src/fft.F90:196-221: mode m = min(k,N-k)
// has akx1 = m*cx and akx2 = -(m*m*cx*cx), Nyquist included in akx2 only.
template <typename TC, int N>
__device__ __forceinline__ cpx<TC> apply_mul(int kind, cpx<TC> z, int k, double cx, int icut,
                                              const double* mc)
{
    constexpr double invN = 1.0 / N;
    const int m = (k < N - k) ? k : N - k;
    if (kind == MUL_D1) {
        if (k == 0 || k == N / 2) return {(TC)0, (TC)0};
        const TC a = (TC)(((double)m * cx) * invN);
        // +i*a for k < N/2; the negative frequency N-k carries -i*a
        if (k < N / 2) return {-a * z.y, a * z.x};
        return {a * z.y, -a * z.x};
    } else if (kind == MUL_D2 || kind == MUL_D2C) {
        double f = -((double)m * (double)m * cx * cx);
        if (kind == MUL_D2C) f *= (k == N / 2) ? mc[0] : mc[0] + mc[1];
        const TC a = (TC)(f * invN);
        return {a * z.x, a * z.y};
    } else {
        if (m < N / 2 - icut) return {z.x * (TC)invN, z.y * (TC)invN};
        return {(TC)0, (TC)0};
    }
}

template <typename PL, int DIR, int B>
struct Lines {
    static constexpr int N = PL::N;
    static constexpr int TPT = plan_tpt<PL>();
    // (transform t, butterfly j) of this thread
    static __device__ __forceinline__ void lane(int tid, int& t, int& j)
    {
        if constexpr (DIR == 0) { j = tid % TPT; t = tid / TPT; }
        else { t = tid % B; j = tid / B; }
    }
    // line groups per job: B complex lines each
    static __host__ __device__ __forceinline__ long groups(const Geom& g)
    {
        if constexpr (DIR == 0) return ((long)g.nz * g.ny / 2 + B - 1) / B;
        else return (long)g.nz * (g.nx / (2 * B));
    }
    // global index of element 0 of the transform's lines a and b in group grp
    static __device__ __forceinline__ bool base(const Geom& g, long grp, int t, long& ia, long& ib, long& step,
                                                int& k)
    {
        if constexpr (DIR == 0) {
            const long cl = grp * B + t;
            const long row = 2 * cl;
            k = (int)(row / g.ny);
            ia = row * N;
            ib = ia + N;
            step = 1;
            return cl < (long)g.nz * g.ny / 2;
        } else {
            const int tpp = g.nx / (2 * B);
            k = (int)(grp / tpp);
            const int x0 = (int)(grp - (long)k * tpp) * 2 * B;
            ia = (long)k * g.nxy + x0 + 2 * t;
            ib = ia + 1;
            step = g.nx;
            return true;
        }
    }
};

// Persistent, software-pipelined kernel: workgroup w handles the items
// (job, line group) w, w + gridDim.x, ...  The pass-0 inputs of the next item
// are loaded into registers before the current item is transformed, and the
// transform touches only registers and LDS (twiddles are staged in LDS), so
// the loads of item i+1 are in flight while item i is computed and stored.
// With L.group_major (a persistent grid smaller than the items), item i of
// round r = i / (G njobs) is job (i mod G njobs) / G of line group
// r G + i mod G: a workgroup then does every job of one line group back to
// back, so jobs that load the same fields (the cross products of u and B)
// find the lines in their CU's caches instead of re-reading HBM.
//
// LOP = LD_CROSS3 (SM_R2HC, three jobs forming one cross product) runs
// body_cross3 instead: an item is one line group, whose six factor lines are
// loaded once and combined into the three jobs' inputs.
template <typename TC, typename PL, int DIR, int B, int MODE, int LOP>
struct SpecK {
    using Args = SpecLaunch;
    static constexpr int NT = B * plan_tpt<PL>();
    // each transform within one wavefront: rows with at most 64 butterflies
    // per transform, or a workgroup of one wave
    static constexpr bool WAVE_LOCAL =
        NT <= 64 || (DIR == 0 && plan_tpt<PL>() <= 64 && 64 % plan_tpt<PL>() == 0);
    using F = RegFFT<TC, PL, WAVE_LOCAL>;
    static __device__ __forceinline__ void body(const SpecLaunch& L, uint3 bid, uint3 gdim)
    {
        if constexpr (LOP == LD_CROSS3) body_cross3(L, bid, gdim);
        else if constexpr (LOP == LD_BSTRESS) body_bstress(L, bid, gdim);
        else body_jobs(L, bid, gdim);
    }
    static __device__ __forceinline__ void body_bstress(const SpecLaunch& L, uint3 bid, uint3 gdim);
    static __device__ __forceinline__ void body_jobs(const SpecLaunch& L, uint3 bid, uint3 gdim);
    static __device__ __forceinline__ void body_cross3(const SpecLaunch& L, uint3 bid, uint3 gdim);
    // Column groups narrower than a 128-byte line (y, 2 B sizeof(TC) < 128
    // bytes) share each line with the next group.  Within every 16 items, item
    // c (< 8) takes group 2c and item c + 8 group 2c + 1: workgroups w and w + 8
    // run on the same XCD (they are dealt round-robin over 8) at the same time,
    // so a line is fetched once.  Identity otherwise.
    static __device__ __forceinline__ long paired_group(long item, long G, long ngrp)
    {
        constexpr bool half_line = DIR == 1 && 2 * B * (int)sizeof(TC) < 128;
        if constexpr (half_line) {
            if (G % 16 == 0 && ngrp % 16 == 0) {
                const long c = item & 15;
                return item - c + (c < 8 ? 2 * c : 2 * (c - 8) + 1);
            }
        }
        return item;
    }
    // forward spectrum in v -> the two halfcomplex lines of the pair, stored
    static __device__ __forceinline__ void r2hc_store(cpx<TC>* v, int j, cpx<TC>* lds, bool ok, TC* __restrict__ out0,
                                                      long ia, long ib, long st);
};

template <typename TC, typename PL, int DIR, int B, int MODE, int LOP>
__device__ __forceinline__ void SpecK<TC, PL, DIR, B, MODE, LOP>::r2hc_store(cpx<TC>* v, int j, cpx<TC>* lds, bool ok,
                                                                             TC* __restrict__ out0, long ia, long ib,
                                                                             long st)
{
    constexpr int N = PL::N;
    constexpr int NB0 = F::NB0, NBL = F::NBL, R0 = F::R0, RL = F::RL;
    // spectrum -> LDS, then separate the pair into two halfcomplex lines
    if (j < NBL) {
#pragma unroll
        for (int r = 0; r < RL; ++r) lds[F::pad(j + r * NBL)] = v[r];
    }
    F::sync();
    if (ok && j < NB0) {
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int n = j + r * NB0;
            const int kk = (n <= N / 2) ? n : N - n;
            const cpx<TC> zk = lds[F::pad(kk)];
            const cpx<TC> zn = lds[F::pad((N - kk) % N)];
            TC a, b;
            if (n <= N / 2) { a = (TC)0.5 * (zk.x + zn.x); b = (TC)0.5 * (zk.y + zn.y); }
            else { a = (TC)0.5 * (zk.y - zn.y); b = (TC)0.5 * (zn.x - zk.x); }
            if constexpr (DIR == 1) st_pair(out0, ia + n * st, a, b);
            else {
                out0[ia + n * st] = a;
                out0[ib + n * st] = b;
            }
        }
    }
    F::sync();
}

template <typename TC, typename PL, int DIR, int B, int MODE, int LOP>
__device__ __forceinline__ void SpecK<TC, PL, DIR, B, MODE, LOP>::body_cross3(const SpecLaunch& L, uint3 bid,
                                                                              uint3 gdim)
{
    using LN = Lines<PL, DIR, B>;
    constexpr int NB0 = F::NB0, R0 = F::R0;
    const Geom g = L.g;
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<TC>* twl = reinterpret_cast<cpx<TC>*>(smem);
    const cpx<TC>* __restrict__ twg = reinterpret_cast<const cpx<TC>*>(L.tw);
    F::stage_twiddles(twl, twg, threadIdx.x, NT);
    int t, j;
    LN::lane(threadIdx.x, t, j);
    cpx<TC>* lds = twl + F::TWN + t * F::SP;
    const long ngrp = LN::groups(g);
    const long G = gdim.x;
    // the factors: job q loads {a_{q+1}, b_{q+2}, a_{q+2}, b_{q+1}}
    const TC* __restrict__ a0 = (const TC*)L.job[2].in[0];
    const TC* __restrict__ a1 = (const TC*)L.job[0].in[0];
    const TC* __restrict__ a2 = (const TC*)L.job[1].in[0];
    const TC* __restrict__ b0 = (const TC*)L.job[1].in[1];
    const TC* __restrict__ b1 = (const TC*)L.job[2].in[1];
    const TC* __restrict__ b2 = (const TC*)L.job[0].in[1];
    __syncthreads();

    // the three jobs' pass-0 inputs of line group grp
    auto fetch = [&](long grp, cpx<TC> (*dst)[R0]) {
        long ia, ib, st;
        int k;
        const bool ok = LN::base(g, grp, t, ia, ib, st, k);
        const Loader<TC, LD_CROSS> l0(L.job[0], ok ? k : 0), l1(L.job[1], ok ? k : 0), l2(L.job[2], ok ? k : 0);
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int n = j + r * NB0;
            dst[0][r] = dst[1][r] = dst[2][r] = {(TC)0, (TC)0};
            if (ok && j < NB0) {
                if constexpr (DIR == 1) {
                    const long i = ia + n * st;
                    const cpx<TC> x0 = ld_pair(a0, i), x1 = ld_pair(a1, i), x2 = ld_pair(a2, i);
                    const cpx<TC> y0 = ld_pair(b0, i), y1 = ld_pair(b1, i), y2 = ld_pair(b2, i);
                    dst[0][r] = l0.cross_pair(x1, y2, x2, y1, i);
                    dst[1][r] = l1.cross_pair(x2, y0, x0, y2, i);
                    dst[2][r] = l2.cross_pair(x0, y1, x1, y0, i);
                } else {
                    const long i = ia + n * st, e = ib + n * st;
                    const TC x0 = a0[i], x1 = a1[i], x2 = a2[i], y0 = b0[i], y1 = b1[i], y2 = b2[i];
                    const TC u0 = a0[e], u1 = a1[e], u2 = a2[e], w0 = b0[e], w1 = b1[e], w2 = b2[e];
                    dst[0][r] = {l0.cross(x1, y2, x2, y1, i), l0.cross(u1, w2, u2, w1, e)};
                    dst[1][r] = {l1.cross(x2, y0, x0, y2, i), l1.cross(u2, w0, u0, w2, e)};
                    dst[2][r] = {l2.cross(x0, y1, x1, y0, i), l2.cross(u0, w1, u1, w0, e)};
                }
            }
        }
    };

    cpx<TC> nxt[3][R0];
    long item = bid.x;
    if (item < ngrp) fetch(paired_group(item, G, ngrp), nxt);
    for (; item < ngrp; item += G) {
        const long grp = paired_group(item, G, ngrp);
        cpx<TC> cur[3][R0];
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int r = 0; r < R0; ++r) cur[q][r] = nxt[q][r];
        long ia, ib, st;
        int k;
        const bool ok = LN::base(g, grp, t, ia, ib, st, k);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            // the next group's loads fly during the last job's transform
            if (q == 2 && item + G < ngrp) fetch(paired_group(item + G, G, ngrp), nxt);
            cpx<TC> v[F::RMAX];
#pragma unroll
            for (int r = 0; r < R0; ++r) v[r] = cur[q][r];
            F::forward(v, j, lds, twl);
            r2hc_store(v, j, lds, ok, (TC*)L.job[q].out[0], ia, ib, st);
        }
    }
}

// The first derivatives of the three Maxwell-stress loads of a direction: an
// item is one line group; the three field components b0, b1, b2 (job 2's
// inputs) are loaded once, each job's input formed from them with the job's
// own means and scale (the arithmetic of the per-job loader), and the three
// transforms run back to back.
template <typename TC, typename PL, int DIR, int B, int MODE, int LOP>
__device__ __forceinline__ void SpecK<TC, PL, DIR, B, MODE, LOP>::body_bstress(const SpecLaunch& L, uint3 bid,
                                                                               uint3 gdim)
{
    constexpr int N = PL::N;
    using LN = Lines<PL, DIR, B>;
    constexpr int NB0 = F::NB0, NBL = F::NBL, R0 = F::R0, RL = F::RL;
    const Geom g = L.g;
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<TC>* twl = reinterpret_cast<cpx<TC>*>(smem);
    const cpx<TC>* __restrict__ twg = reinterpret_cast<const cpx<TC>*>(L.tw);
    F::stage_twiddles(twl, twg, threadIdx.x, NT);
    int t, j;
    LN::lane(threadIdx.x, t, j);
    cpx<TC>* lds = twl + F::TWN + t * F::SP;
    const long ngrp = LN::groups(g);
    const long G = gdim.x;
    const double cx = L.cx;
    const int icut = L.icut;
    const TC* __restrict__ f0 = (const TC*)L.job[2].in[0];
    const TC* __restrict__ f1 = (const TC*)L.job[2].in[1];
    const TC* __restrict__ f2 = (const TC*)L.job[2].in[2];
    const bool pat = L.bpat != 0;   // job 0 is PROD(b0, b1) rather than PROD(b1, b0)
    __syncthreads();

    auto fetch = [&](long grp, cpx<TC> (*dst)[R0]) {
        long ia, ib, st;
        int k;
        const bool ok = LN::base(g, grp, t, ia, ib, st, k);
        const Loader<TC, LD_PROD> l0(L.job[0], ok ? k : 0), l1(L.job[1], ok ? k : 0);
        const Loader<TC, LD_MAXW> l2(L.job[2], ok ? k : 0);
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int n = j + r * NB0;
            dst[0][r] = dst[1][r] = dst[2][r] = {(TC)0, (TC)0};
            if (ok && j < NB0) {
                cpx<TC> a, b, c;
                if constexpr (DIR == 1) {
                    const long i = ia + n * st;
                    a = ld_pair(f0, i);
                    b = ld_pair(f1, i);
                    c = ld_pair(f2, i);
                } else {
                    const long i = ia + n * st, e = ib + n * st;
                    a = {f0[i], f0[e]};
                    b = {f1[i], f1[e]};
                    c = {f2[i], f2[e]};
                }
                dst[0][r] = pat ? l0.combine(a, b, c) : l0.combine(b, a, c);
                dst[1][r] = l1.combine(c, a, b);
                dst[2][r] = l2.combine(a, b, c);
            }
        }
    };

    cpx<TC> nxt[3][R0];
    long item = bid.x;
    if (item < ngrp) fetch(paired_group(item, G, ngrp), nxt);
    for (; item < ngrp; item += G) {
        const long grp = paired_group(item, G, ngrp);
        cpx<TC> cur[3][R0];
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int r = 0; r < R0; ++r) cur[q][r] = nxt[q][r];
        long ia, ib, st;
        int k;
        const bool ok = LN::base(g, grp, t, ia, ib, st, k);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const SpecJob& J = L.job[q];
            // the next group's loads fly during the last job's transform
            if (q == 2 && item + G < ngrp) fetch(paired_group(item + G, G, ngrp), nxt);
            cpx<TC> v[F::RMAX];
#pragma unroll
            for (int r = 0; r < R0; ++r) v[r] = cur[q][r];
            F::forward(v, j, lds, twl);
            if (j < NBL) {
#pragma unroll
                for (int r = 0; r < RL; ++r) v[r] = apply_mul<TC, N>(J.mul[0], v[r], j + r * NBL, cx, icut, J.mc);
            }
            F::inverse(v, j, lds, twl);
            if (ok && j < NB0) {
                TC* __restrict__ out = (TC*)J.out[0];
                const TC add = (J.addk ? (TC)J.addk[k] : (TC)0) + (TC)J.adds;
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    const int n = j + r * NB0;
                    if constexpr (DIR == 1) st_pair_nt(out, ia + n * st, v[r].x + add, v[r].y + add);
                    else {
                        st_nt(out, ia + n * st, v[r].x + add);
                        st_nt(out, ib + n * st, v[r].y + add);
                    }
                }
            }
        }
    }
}

template <typename TC, typename PL, int DIR, int B, int MODE, int LOP>
__device__ __forceinline__ void SpecK<TC, PL, DIR, B, MODE, LOP>::body_jobs(const SpecLaunch& L, uint3 bid,
                                                                            uint3 gdim)
{
    constexpr int N = PL::N;
    using LN = Lines<PL, DIR, B>;
    constexpr int NB0 = F::NB0, NBL = F::NBL, R0 = F::R0, RL = F::RL;
    const Geom g = L.g;
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<TC>* twl = reinterpret_cast<cpx<TC>*>(smem);
    const cpx<TC>* __restrict__ twg = reinterpret_cast<const cpx<TC>*>(L.tw);
    F::stage_twiddles(twl, twg, threadIdx.x, NT);
    int t, j;
    LN::lane(threadIdx.x, t, j);
    cpx<TC>* lds = twl + F::TWN + t * F::SP;
    const long ngrp = LN::groups(g);
    const long G = gdim.x;
    const bool gmaj = L.group_major != 0;
    const long nitems = gmaj ? ((ngrp + G - 1) / G) * G * L.njobs : ngrp * L.njobs;
    const double cx = L.cx;
    const int icut = L.icut;
    __syncthreads();

    // Column groups narrower than a 128-byte line (y direction, 2 B elements of
    // sizeof(TC) < 128 bytes: the 512-point lengths) share each line with the
    // next group.  Within every 16 items, item c (< 8) takes group 2c and item
    // c + 8 group 2c + 1: workgroups w and w + 8 run on the same XCD (they are
    // dealt round-robin over 8) at the same time, so a line is fetched once.
    constexpr bool half_line = DIR == 1 && 2 * B * (int)sizeof(TC) < 128;
    const bool pair = half_line && !gmaj && (G % 16 == 0) && (ngrp % 16 == 0);
    // (job, line group) of an item; false for the padding of the last round
    auto decode = [&](long item, int& jb, long& grp) -> bool {
        if (!gmaj) {
            if (pair) {
                const long c = item & 15;
                item = item - c + (c < 8 ? 2 * c : 2 * (c - 8) + 1);
            }
            jb = (int)(item / ngrp);
            grp = item - (long)jb * ngrp;
            return true;
        }
        const long per = G * L.njobs;
        const long r = item / per;
        const long rem = item - r * per;
        jb = (int)(rem / G);
        grp = r * G + (rem - (long)jb * G);
        return grp < ngrp;
    };

    // pass-0 inputs of one item (zeros for lines past the end)
    auto fetch = [&](long item, cpx<TC>* dst) {
        int jb;
        long grp;
        if (!decode(item, jb, grp)) return;
        const SpecJob& J = L.job[jb];
        long ia, ib, st;
        int k;
        const bool ok = LN::base(g, grp, t, ia, ib, st, k);
        if constexpr (MODE == SM_HC2R) {
            const TC* __restrict__ src = (const TC*)J.in[0];
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                const int n = j + r * NB0;
                dst[r] = {(TC)0, (TC)0};
                if (ok && j < NB0) {
                    if constexpr (DIR == 1) dst[r] = ld_pair(src, ia + n * st);
                    else dst[r] = {src[ia + n * st], src[ib + n * st]};
                }
            }
        } else {
            auto load_with = [&](auto op) {
                constexpr int LO = decltype(op)::value;
                const Loader<TC, LO> ld(J, ok ? k : 0);
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    const int n = j + r * NB0;
                    dst[r] = {(TC)0, (TC)0};
                    if (ok && j < NB0) {
                        if constexpr (DIR == 1) dst[r] = ld.pair(ia + n * st);
                        else dst[r] = {ld(ia + n * st), ld(ib + n * st)};
                    }
                }
            };
            if constexpr (LOP == LD_ANY) {
                // the job's load operation, uniform over the workgroup
                if (J.load_op == LD_PROD) load_with(std::integral_constant<int, LD_PROD>());
                else if (J.load_op == LD_MAXW) load_with(std::integral_constant<int, LD_MAXW>());
                else load_with(std::integral_constant<int, LD_PLAIN>());
            } else {
                load_with(std::integral_constant<int, LOP>());
            }
        }
    };

    cpx<TC> nxt[R0];
    long item = bid.x;
    if (item < nitems) fetch(item, nxt);
    for (; item < nitems; item += G) {
        cpx<TC> v[F::RMAX];
#pragma unroll
        for (int r = 0; r < R0; ++r) v[r] = nxt[r];
        if (item + G < nitems) fetch(item + G, nxt);

        int jb;
        long grp;
        if (!decode(item, jb, grp)) continue;   // uniform over the workgroup
        const SpecJob& J = L.job[jb];
        long ia, ib, st;
        int k;
        const bool ok = LN::base(g, grp, t, ia, ib, st, k);
        TC* __restrict__ out0 = (TC*)J.out[0];
        TC* __restrict__ out1 = (TC*)J.out[1];
        // epilogue constant of this line pair (one plane): the mean parts of B
        const TC add = (J.addk && ok ? (TC)J.addk[k] : (TC)0) + (TC)J.adds;

        if constexpr (MODE == SM_HC2R) {
            // raw halfcomplex lines -> LDS (natural order), then Z = A + iB
            if (j < NB0) {
#pragma unroll
                for (int r = 0; r < R0; ++r) lds[F::pad(j + r * NB0)] = v[r];
            }
            F::sync();
            if (j < NBL) {
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    const int kk = j + r * NBL;
                    const cpx<TC> u = lds[F::pad(kk)];
                    const cpx<TC> m = lds[F::pad((N - kk) % N)];
                    cpx<TC> A, Bv;
                    if (kk == 0 || kk == N / 2) { A = {u.x, (TC)0}; Bv = {u.y, (TC)0}; }
                    else if (kk < N / 2) { A = {u.x, m.x}; Bv = {u.y, m.y}; }
                    else { A = {m.x, -u.x}; Bv = {m.y, -u.y}; }
                    v[r] = {A.x - Bv.y, A.y + Bv.x};
                }
            }
            F::sync();
            F::inverse(v, j, lds, twl);
            if (ok && j < NB0) {
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    const int n = j + r * NB0;
                    const TC xa = v[r].x * (TC)(1.0 / N) + add, xb = v[r].y * (TC)(1.0 / N) + add;
                    if constexpr (DIR == 1) st_pair(out0, ia + n * st, xa, xb);
                    else {
                        out0[ia + n * st] = xa;
                        out0[ib + n * st] = xb;
                    }
                }
            }
        } else {
            F::forward(v, j, lds, twl);
            if constexpr (MODE == SM_R2HC) {
                r2hc_store(v, j, lds, ok, out0, ia, ib, st);
            } else {
                constexpr int NOUT = (MODE == SM_DERIV2) ? 2 : 1;
                cpx<TC> z[RL];
                if constexpr (NOUT == 2) {
#pragma unroll
                    for (int r = 0; r < RL; ++r) z[r] = v[r];
                }
#pragma unroll
                for (int o = 0; o < NOUT; ++o) {
                    if (o == 1 && out1 == nullptr) break;   // a first-derivative job (uniform)
                    if (j < NBL) {
#pragma unroll
                        for (int r = 0; r < RL; ++r)
                            v[r] = apply_mul<TC, N>(J.mul[o], o == 0 ? v[r] : z[r], j + r * NBL, cx, icut, J.mc);
                    }
                    F::inverse(v, j, lds, twl);
                    if (ok && j < NB0) {
                        TC* __restrict__ out = (o == 0) ? out0 : out1;
#pragma unroll
                        for (int r = 0; r < R0; ++r) {
                            const int n = j + r * NB0;
                            if constexpr (DIR == 1) st_pair_nt(out, ia + n * st, v[r].x + add, v[r].y + add);
                            else {
                                st_nt(out, ia + n * st, v[r].x + add);
                                st_nt(out, ib + n * st, v[r].y + add);
                            }
                        }
                    }
                }
            }
        }
    }
}


// ---------------------------------------------------------------------------
// host dispatch
// ---------------------------------------------------------------------------
// transforms per workgroup: about 256 threads along x; along y, 8 (FP64) or
// 16 (FP32) column pairs = 128-byte row segments.  LDS per workgroup: up to
// 64 KB, and up to 160 KB (gfx950's limit) for FP64 along y, so that the
// 512-point FP64 lengths keep full-line column groups (about 78 KB; the 4-field
// pass at 512^2 x 256 went from 2.15 to 1.48 ms).  In FP32 the 64 KB groups
// (half lines, paired onto one XCD) measured faster than full-line ones.
constexpr int SPEC_LDS_X = 65536, SPEC_LDS_Y = 163840;
template <typename TC, typename PL, int DIR>
constexpr int spec_b()
{
    constexpr int tpt = plan_tpt<PL>();
    constexpr int sp = RegFFT<TC, PL>::SP * (int)sizeof(cpx<TC>);
    constexpr int twb = RegFFT<TC, PL>::TWN * (int)sizeof(cpx<TC>);
    constexpr int lim = (DIR == 1 && sizeof(TC) == 8) ? SPEC_LDS_Y : SPEC_LDS_X;
    // FP64 y groups of 8 pairs (128-byte rows); 16 (256-byte rows) where a
    // transform takes at most 16 threads, so the group stays at 256 threads
    int b = DIR == 0 ? (128 / tpt > 0 ? 128 / tpt : 1) : (sizeof(TC) == 8 ? (tpt <= 16 ? 16 : 8) : 16);
    while (b > 1 && (b * tpt > 1024 || twb + b * sp > lim)) b /= 2;
    return b;
}

template <typename TC, typename PL, int B>
constexpr size_t spec_lds_bytes()
{
    return (size_t)(RegFFT<TC, PL>::TWN + B * RegFFT<TC, PL>::SP) * sizeof(cpx<TC>);
}

template <typename TC, typename PL, int DIR, int B, int MODE, int LOP>
static int launch_one(const SpecLaunch& L, hipStream_t s)
{
    using K = SpecK<TC, PL, DIR, B, MODE, LOP>;
    constexpr int NT = K::NT;
    constexpr size_t lds = spec_lds_bytes<TC, PL, B>();
    static_assert(lds <= (size_t)SPEC_LDS_Y, "transform workgroup exceeds 160 KB of LDS");
    auto kern = single_kernel<K>;
    // LD_CROSS3, LD_BSTRESS: one item per line group for all three jobs
    const long items = Lines<PL, DIR, B>::groups(L.g) * (LOP == LD_CROSS3 || LOP == LD_BSTRESS ? 1 : L.njobs);
    // more than 64 KB of dynamic LDS is opted into once per kernel
    static int lds_ok = lds <= 65536 ? 1 : 0;
    if (!lds_ok) {
        if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
            hipSuccess)
            return -22;
        lds_ok = 1;
    }
    // persistent grid: the workgroups that fit on the device at once
    static int per_cu = 0;
    if (per_cu == 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, NT, lds) != hipSuccess || n < 1) n = 1;
        per_cu = n;
    }
    long grid = items;
    const long resident = (long)per_cu * L.ncu * (L.grid_mult > 1 ? L.grid_mult : 1);
    if (L.ncu > 0 && grid > resident) grid = resident;
    if (L.max_groups > 0 && grid > L.max_groups) grid = L.max_groups;
    // line group first only for the FP64 EMF pass, whose three jobs load the same
    // six fields (at 256^2 x 160: 235 -> 191 us); with independent inputs, and in
    // FP32, the job-major order measured 5-10 % faster
    SpecLaunch La = L;
    La.group_major = grid < items && L.njobs > 1 && LOP == LD_CROSS && sizeof(TC) == 8;
    return mhd_launch<K>(dim3((unsigned)grid), dim3(NT), lds, s, La);
}

// (mode, load op) pairs the solver uses; every job of L has the same load op
template <typename TC, typename PL, int DIR, int B>
static int launch_mode(const SpecLaunch& L, hipStream_t s)
{
    if (L.any_load) return launch_one<TC, PL, DIR, B, SM_DERIV1, LD_ANY>(L, s);
    if (L.bstress) return launch_one<TC, PL, DIR, B, SM_DERIV1, LD_BSTRESS>(L, s);
    if constexpr (DIR == 0) {
        if (L.cross3) return launch_one<TC, PL, DIR, B, SM_R2HC, LD_CROSS3>(L, s);
    }
    const int lop = L.job[0].load_op;
    switch (L.mode * 8 + lop) {
    case SM_DERIV1 * 8 + LD_PLAIN: return launch_one<TC, PL, DIR, B, SM_DERIV1, LD_PLAIN>(L, s);
    case SM_DERIV1 * 8 + LD_PROD: return launch_one<TC, PL, DIR, B, SM_DERIV1, LD_PROD>(L, s);
    case SM_DERIV1 * 8 + LD_BG: return launch_one<TC, PL, DIR, B, SM_DERIV1, LD_BG>(L, s);
    case SM_DERIV1 * 8 + LD_MAXW: return launch_one<TC, PL, DIR, B, SM_DERIV1, LD_MAXW>(L, s);
    case SM_DERIV2 * 8 + LD_PLAIN: return launch_one<TC, PL, DIR, B, SM_DERIV2, LD_PLAIN>(L, s);
    case SM_DERIV2 * 8 + LD_BG: return launch_one<TC, PL, DIR, B, SM_DERIV2, LD_BG>(L, s);
    case SM_R2HC * 8 + LD_PLAIN: return launch_one<TC, PL, DIR, B, SM_R2HC, LD_PLAIN>(L, s);
    case SM_R2HC * 8 + LD_CROSS: return launch_one<TC, PL, DIR, B, SM_R2HC, LD_CROSS>(L, s);
    case SM_HC2R * 8 + LD_PLAIN: return launch_one<TC, PL, DIR, B, SM_HC2R, LD_PLAIN>(L, s);
    }
    return -23;
}

// the single-output derivative passes along y (one input, one output per
// line) with half the column pairs per workgroup (512 threads instead of 1024
// at 256 points: two workgroups per CU fit where one did)
template <typename TC, typename PL, int B>
static int launch_y_deriv1(const SpecLaunch& L, hipStream_t s)
{
    if (L.any_load) return launch_one<TC, PL, 1, B, SM_DERIV1, LD_ANY>(L, s);
    if (L.bstress) return launch_one<TC, PL, 1, B, SM_DERIV1, LD_BSTRESS>(L, s);
    switch (L.job[0].load_op) {
    case LD_PLAIN: return launch_one<TC, PL, 1, B, SM_DERIV1, LD_PLAIN>(L, s);
    case LD_PROD: return launch_one<TC, PL, 1, B, SM_DERIV1, LD_PROD>(L, s);
    case LD_MAXW: return launch_one<TC, PL, 1, B, SM_DERIV1, LD_MAXW>(L, s);
    }
    return launch_mode<TC, PL, 1, B>(L, s);
}

template <typename TC, typename PL>
static int launch_plan(const SpecLaunch& L, hipStream_t s)
{
    if (L.dir == 0) return launch_mode<TC, PL, 0, spec_b<TC, PL, 0>()>(L, s);
    constexpr int B0 = spec_b<TC, PL, 1>();
    // FP32 only, where a full group would be a 1024-thread workgroup (one per
    // CU): at 256 points 8 pairs beat 16 for these passes (76, 72, 45 us
    // against 81, 83, 48); in FP64 4 pairs lose to 8.  The other FP32 modes keep
    // 16 pairs (8 pairs: d1+d2 155 against 137 us, hc2r 125 against 109, r2hc
    // 63 against 58)
    constexpr int BH = (sizeof(TC) == 4 && B0 >= 8 && B0 * plan_tpt<PL>() >= 1024) ? B0 / 2 : B0;
    if (BH != B0 && L.mode == SM_DERIV1 && L.g.nx % (2 * BH) == 0) return launch_y_deriv1<TC, PL, BH>(L, s);
    // FP64 Maxwell-stress triple (about 150 VGPRs): half groups, 3 workgroups
    // per CU instead of 1 (at 256^2 x 160: 153 against 171 us; the groups
    // sharing a line are paired onto one XCD, so each line is fetched once)
    if constexpr (sizeof(TC) == 8 && B0 >= 8) {
        if (L.bstress && L.g.nx % B0 == 0) return launch_one<TC, PL, 1, B0 / 2, SM_DERIV1, LD_BSTRESS>(L, s);
    }
    if (L.g.nx % (2 * B0) == 0) return launch_mode<TC, PL, 1, B0>(L, s);
    return launch_mode<TC, PL, 1, 1>(L, s);
}

template <typename TC, int N>
static int launch_n(const SpecLaunch& L, hipStream_t s)
{
    if constexpr (!std::is_void<typename AltPlan<N>::type>::value) {
        if (L.variant == 1) return launch_plan<TC, typename AltPlan<N>::type>(L, s);
    }
    if constexpr (!std::is_void<typename AltPlan2<N>::type>::value) {
        if (L.variant == 2) return launch_plan<TC, typename AltPlan2<N>::type>(L, s);
    }
    return launch_plan<TC, FftPlan<N>>(L, s);
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
