// spectral_generic.hip -- the horizontal transform kernels for any even
// length that has no register-blocked plan (fft_reg.hpp), as the reference
// plans FFTW for any even nx, ny (src/fft.F90:101-136).
//
// A workgroup holds `nlb` complex lines (two real lines each, as in
// spectral_impl.hpp) in LDS and runs a self-sorting (Stockham) FFT whose
// radices are the length's factors, found at run time: 4s, then 2, 3, 5, 7
// and any remaining prime.  Each pass computes every output point directly,
//   dst[o] = sum_r src[j + r N/R] W_N^(r (N / (Ns R)) (kk + m Ns)),
// with o = (j / Ns) Ns R + kk + m Ns, kk = j mod Ns, m < R, from one twiddle
// table W_N^t (LDS), ping-ponging between two LDS buffers.  This costs N R
// multiply-adds per line and pass (R a prime factor), against the
// register-blocked kernels' fixed radix-2/3/4/5/8 butterflies, which serve
// the usual lengths; loads (every LoadOp), spectral multipliers, the
// halfcomplex packing and the stores are those of spec_kernel.
#include "spectral_impl.hpp"

namespace mhd {

namespace {

constexpr int GNT = 256;

template <typename TC>
__device__ __forceinline__ cpx<TC> apply_mul_rt(int kind, cpx<TC> z, int k, int N, double cx, int icut,
                                                 const double* mc)
{
    const double invN = 1.0 / N;
    const int m = (k < N - k) ? k : N - k;
    if (kind == MUL_D1) {
        if (k == 0 || 2 * k == N) return {(TC)0, (TC)0};
        const TC a = (TC)(((double)m * cx) * invN);
        if (2 * k < N) return {-a * z.y, a * z.x};
        return {a * z.y, -a * z.x};
    } else if (kind == MUL_D2 || kind == MUL_D2C) {
        double f = -((double)m * (double)m * cx * cx);
        if (kind == MUL_D2C) f *= (2 * k == N) ? mc[0] : mc[0] + mc[1];
        const TC a = (TC)(f * invN);
        return {a * z.x, a * z.y};
    } else {
        if (m < N / 2 - icut) return {z.x * (TC)invN, z.y * (TC)invN};
        return {(TC)0, (TC)0};
    }
}

// one pass over the workgroup's nlb lines, src -> dst
template <typename TC, bool INV>
__device__ __forceinline__ void gen_pass(const cpx<TC>* __restrict__ src, cpx<TC>* __restrict__ dst,
                                         const cpx<TC>* __restrict__ tw, int N, int nlb, int Ns, int R)
{
    const int NB = N / R;
    const int tstride = N / (Ns * R);
    const int NsR = Ns * R;
    for (int q = threadIdx.x; q < nlb * N; q += GNT) {
        const int b = q / N, o = q - b * N;
        const int blk = o / NsR, rem = o - blk * NsR;
        const int m = rem / Ns, kk = rem - m * Ns;
        const int j = blk * Ns + kk;
        const int estep = (tstride * (kk + m * Ns)) % N;
        const cpx<TC>* s = src + b * N;
        cpx<TC> acc = s[j];
        int e = 0;
        for (int r = 1; r < R; ++r) {
            e += estep;
            if (e >= N) e -= N;
            const cpx<TC> x = s[j + r * NB];
            acc = cadd(acc, INV ? cmulc(x, tw[e]) : cmul(x, tw[e]));
        }
        dst[b * N + o] = acc;
    }
    __syncthreads();
}

// the transform of the lines in *a (b is scratch of the same size); returns
// the buffer holding the result
template <typename TC, bool INV>
__device__ __forceinline__ cpx<TC>* gen_fft(cpx<TC>* a, cpx<TC>* b, const cpx<TC>* tw, const SpecLaunch& L,
                                            int N, int nlb)
{
    int Ns = 1;
    for (int p = 0; p < L.gen_npass; ++p) {
        const int R = L.gen_radix[p];
        gen_pass<TC, INV>(a, b, tw, N, nlb, Ns, R);
        cpx<TC>* t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

template <int DIR>
struct GenLines {
    // line groups of nlb complex lines per job
    static __host__ __device__ long groups(const Geom& g, int nlb)
    {
        if constexpr (DIR == 0) return ((long)g.nz * g.ny / 2 + nlb - 1) / nlb;
        else return (long)g.nz * ((g.nx / 2 + nlb - 1) / nlb);
    }
    // element 0 of the lines a, b of line b_ in group grp; false past the end
    static __device__ __forceinline__ bool base(const Geom& g, int nlb, long grp, int b_, long& ia, long& ib,
                                                long& st, int& k)
    {
        if constexpr (DIR == 0) {
            const long cl = grp * nlb + b_;
            const long row = 2 * cl;
            k = (int)(row / g.ny);
            ia = row * g.nx;
            ib = ia + g.nx;
            st = 1;
            return cl < (long)g.nz * g.ny / 2;
        } else {
            const int gpp = (g.nx / 2 + nlb - 1) / nlb;
            k = (int)(grp / gpp);
            const int cp = (int)(grp - (long)k * gpp) * nlb + b_;
            ia = (long)k * g.nxy + 2 * cp;
            ib = ia + 1;
            st = g.nx;
            return 2 * cp < g.nx;
        }
    }
};

template <typename TC, int DIR, int MODE, int LOP>
struct GenK {
    using Args = SpecLaunch;
    static constexpr int NT = GNT;
    static __device__ __forceinline__ void body(const SpecLaunch& L, uint3 bid, uint3 gdim);
};

template <typename TC, int DIR, int MODE, int LOP>
__device__ __forceinline__ void GenK<TC, DIR, MODE, LOP>::body(const SpecLaunch& L, uint3 bid, uint3 gdim)
{
    const Geom g = L.g;
    const int N = DIR == 0 ? g.nx : g.ny;
    const int nlb = L.gen_nlb;
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<TC>* twl = reinterpret_cast<cpx<TC>*>(smem);
    cpx<TC>* buf0 = twl + N;
    cpx<TC>* buf1 = buf0 + nlb * N;
    cpx<TC>* buf2 = buf1 + nlb * N;   // the kept spectrum of two-output passes
    const cpx<TC>* __restrict__ twg = reinterpret_cast<const cpx<TC>*>(L.tw);
    for (int i = threadIdx.x; i < N; i += GNT) twl[i] = twg[i];
    const long ngrp = GenLines<DIR>::groups(g, nlb);
    const long nitems = ngrp * L.njobs;
    const int P = nlb * N;
    __syncthreads();

    for (long item = bid.x; item < nitems; item += gdim.x) {
        const int jb = (int)(item / ngrp);
        const long grp = item - (long)jb * ngrp;
        const SpecJob& J = L.job[jb];
        // point q of the workgroup: line b, element n (rows: n fastest;
        // columns: the line fastest, so a row of adjacent column pairs is one
        // contiguous load)
        auto point = [&](int q, int& b, int& n) {
            if constexpr (DIR == 0) { b = q / N; n = q - b * N; }
            else { n = q / nlb; b = q - n * nlb; }
        };
        // load
        for (int q = threadIdx.x; q < P; q += GNT) {
            int b, n;
            point(q, b, n);
            long ia, ib, st;
            int k;
            cpx<TC> v = {(TC)0, (TC)0};
            if (GenLines<DIR>::base(g, nlb, grp, b, ia, ib, st, k)) {
                if constexpr (MODE == SM_HC2R) {
                    const TC* __restrict__ src = (const TC*)J.in[0];
                    v = {src[ia + n * st], src[ib + n * st]};
                } else {
                    const Loader<TC, LOP> ld(J, k);
                    v = {ld(ia + n * st), ld(ib + n * st)};
                }
            }
            buf0[b * N + n] = v;
        }
        __syncthreads();

        TC* __restrict__ out0 = (TC*)J.out[0];
        TC* __restrict__ out1 = (TC*)J.out[1];
        auto store = [&](const cpx<TC>* res, TC* __restrict__ out, TC scale) {
            for (int q = threadIdx.x; q < P; q += GNT) {
                int b, n;
                point(q, b, n);
                long ia, ib, st;
                int k;
                if (!GenLines<DIR>::base(g, nlb, grp, b, ia, ib, st, k)) continue;
                const TC add = (J.addk ? (TC)J.addk[k] : (TC)0) + (TC)J.adds;
                const cpx<TC> v = res[b * N + n];
                out[ia + n * st] = v.x * scale + add;
                out[ib + n * st] = v.y * scale + add;
            }
        };

        if constexpr (MODE == SM_HC2R) {
            // two halfcomplex lines -> the spectrum of z = a + i b
            for (int q = threadIdx.x; q < P; q += GNT) {
                const int b = q / N, kk = q - b * N;
                const cpx<TC> u = buf0[b * N + kk];
                const cpx<TC> m = buf0[b * N + (N - kk) % N];
                cpx<TC> A, Bv;
                if (kk == 0 || 2 * kk == N) { A = {u.x, (TC)0}; Bv = {u.y, (TC)0}; }
                else if (2 * kk < N) { A = {u.x, m.x}; Bv = {u.y, m.y}; }
                else { A = {m.x, -u.x}; Bv = {m.y, -u.y}; }
                buf1[b * N + kk] = {A.x - Bv.y, A.y + Bv.x};
            }
            __syncthreads();
            const cpx<TC>* res = gen_fft<TC, true>(buf1, buf2, twl, L, N, nlb);
            store(res, out0, (TC)(1.0 / N));
        } else {
            cpx<TC>* spec = gen_fft<TC, false>(buf0, buf1, twl, L, N, nlb);
            cpx<TC>* other = spec == buf0 ? buf1 : buf0;
            if constexpr (MODE == SM_R2HC) {
                for (int q = threadIdx.x; q < P; q += GNT) {
                    int b, n;
                    point(q, b, n);
                    long ia, ib, st;
                    int k;
                    if (!GenLines<DIR>::base(g, nlb, grp, b, ia, ib, st, k)) continue;
                    const int kk = (2 * n <= N) ? n : N - n;
                    const cpx<TC> zk = spec[b * N + kk];
                    const cpx<TC> zn = spec[b * N + (N - kk) % N];
                    TC a, bb;
                    if (2 * n <= N) { a = (TC)0.5 * (zk.x + zn.x); bb = (TC)0.5 * (zk.y + zn.y); }
                    else { a = (TC)0.5 * (zk.y - zn.y); bb = (TC)0.5 * (zn.x - zk.x); }
                    out0[ia + n * st] = a;
                    out0[ib + n * st] = bb;
                }
            } else {
                constexpr int NOUT = (MODE == SM_DERIV2) ? 2 : 1;
                for (int o = 0; o < NOUT; ++o) {
                    if (o == 1 && out1 == nullptr) break;   // a first-derivative job (uniform)
                    // multiplied spectrum into `other`, inverse through buf2
                    for (int q = threadIdx.x; q < P; q += GNT) {
                        const int b = q / N, kk = q - b * N;
                        other[q] = apply_mul_rt<TC>(J.mul[o], spec[q], kk, N, L.cx, L.icut, J.mc);
                        (void)b;
                    }
                    __syncthreads();
                    const cpx<TC>* res = gen_fft<TC, true>(other, buf2, twl, L, N, nlb);
                    store(res, o == 0 ? out0 : out1, (TC)1);
                    __syncthreads();   // `other` and buf2 are reused by the next output
                }
            }
        }
        __syncthreads();   // LDS reused by the next item
    }
}

size_t gen_lds_bytes(int N, int nlb, size_t csz) { return (size_t)(N + 3 * nlb * N) * csz; }

template <typename TC, int DIR, int MODE, int LOP>
int gen_launch(const SpecLaunch& L, hipStream_t s)
{
    const size_t lds = gen_lds_bytes(L.dir == 0 ? L.g.nx : L.g.ny, L.gen_nlb, sizeof(cpx<TC>));
    using K = GenK<TC, DIR, MODE, LOP>;
    auto kern = single_kernel<K>;
    // more than 64 KB of dynamic LDS is opted into once per kernel
    static int lds_max = 65536;
    if ((int)lds > lds_max) {
        if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, 163840) !=
            hipSuccess)
            return -22;
        lds_max = 163840;
    }
    const long items = GenLines<DIR>::groups(L.g, L.gen_nlb) * L.njobs;
    long grid = items;
    const long cap = (long)(L.ncu > 0 ? L.ncu : 256) * 8;
    if (grid > cap) grid = cap;
    return mhd_launch<K>(dim3((unsigned)grid), dim3(GNT), lds, s, L);
}

template <typename TC, int DIR>
int gen_mode(const SpecLaunch& L, hipStream_t s)
{
    switch (L.mode * 8 + L.job[0].load_op) {
    case SM_DERIV1 * 8 + LD_PLAIN: return gen_launch<TC, DIR, SM_DERIV1, LD_PLAIN>(L, s);
    case SM_DERIV1 * 8 + LD_PROD: return gen_launch<TC, DIR, SM_DERIV1, LD_PROD>(L, s);
    case SM_DERIV1 * 8 + LD_BG: return gen_launch<TC, DIR, SM_DERIV1, LD_BG>(L, s);
    case SM_DERIV1 * 8 + LD_MAXW: return gen_launch<TC, DIR, SM_DERIV1, LD_MAXW>(L, s);
    case SM_DERIV2 * 8 + LD_PLAIN: return gen_launch<TC, DIR, SM_DERIV2, LD_PLAIN>(L, s);
    case SM_DERIV2 * 8 + LD_BG: return gen_launch<TC, DIR, SM_DERIV2, LD_BG>(L, s);
    case SM_R2HC * 8 + LD_PLAIN: return gen_launch<TC, DIR, SM_R2HC, LD_PLAIN>(L, s);
    case SM_R2HC * 8 + LD_CROSS: return gen_launch<TC, DIR, SM_R2HC, LD_CROSS>(L, s);
    case SM_HC2R * 8 + LD_PLAIN: return gen_launch<TC, DIR, SM_HC2R, LD_PLAIN>(L, s);
    }
    return -23;
}

}  // namespace

bool generic_plan(int n, int radix[SPEC_GEN_MAX_PASSES], int* npass)
{
    if (n < 2 || n % 2 || n > SPEC_GEN_MAX_N) return false;
    int p = 0, m = n;
    while (m % 4 == 0 && p < SPEC_GEN_MAX_PASSES) { radix[p++] = 4; m /= 4; }
    for (int f = 2; m > 1 && p < SPEC_GEN_MAX_PASSES; ++f)
        while (m % f == 0 && p < SPEC_GEN_MAX_PASSES) { radix[p++] = f; m /= f; }
    *npass = p;
    return m == 1;
}

int launch_spectral_generic(const SpecLaunch& L0, int f32, hipStream_t s)
{
    SpecLaunch L = L0;
    const int N = L.dir == 0 ? L.g.nx : L.g.ny;
    if (!generic_plan(N, L.gen_radix, &L.gen_npass)) return -20;
    // lines per workgroup: about 1024 points, at most 64 KB of LDS where one
    // line fits in that
    const size_t csz = f32 ? sizeof(cpx<float>) : sizeof(cpx<double>);
    int nlb = 1024 / N;
    if (nlb < 1) nlb = 1;
    if (nlb > 64) nlb = 64;
    while (nlb > 1 && gen_lds_bytes(N, nlb, csz) > 65536) --nlb;
    L.gen_nlb = nlb;
    if (f32) return L.dir == 0 ? gen_mode<float, 0>(L, s) : gen_mode<float, 1>(L, s);
    return L.dir == 0 ? gen_mode<double, 0>(L, s) : gen_mode<double, 1>(L, s);
}

}  // namespace mhd
