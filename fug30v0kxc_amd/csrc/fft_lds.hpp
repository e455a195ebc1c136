// fft_lds.hpp -- batched complex FFTs held in LDS, for gfx950.
//
// Self-sorting (Stockham) mixed-radix passes with radices 4, 2, 3, 5, fully
// unrolled at compile time for a transform length N and a batch of BATCH
// transforms per workgroup of NT threads.  Each pass reads R points with
// stride N/R, applies twiddles from a table W_N^t = exp(-2 pi i t / N) and an
// R-point butterfly, and writes them back self-sorted; reads and writes are
// separated by a workgroup barrier so the pass works in place.  Transform b
// lives at buf[b*S .. b*S+N-1] (S >= N; padding keeps the column-tile loads of
// the y kernels off a single LDS bank).
//
// Sign convention: forward = exp(-i...), as in FFTW's REAL_TO_COMPLEX; the
// backward transform is unnormalised.
#pragma once

#include "common.hpp"

namespace mhd {

constexpr int radix_of(int n)
{
    return (n % 4 == 0) ? 4 : (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : (n % 5 == 0) ? 5 : n;
}

constexpr bool fft_len_ok(int n)
{
    if (n < 2) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    while (n % 5 == 0) n /= 5;
    return n == 1;
}

// R-point DFT of v[0..R-1] in registers; INV selects exp(+i...)
template <typename T, int R, bool INV>
__device__ __forceinline__ void butterfly(cpx<T>* v)
{
    constexpr int SG = INV ? 1 : -1;
    if constexpr (R == 2) {
        cpx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        cpx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        cpx<T> c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        cpx<T> id = cmul_i<T, SG>(d);
        v[0] = cadd(a, c);
        v[1] = cadd(b, id);
        v[2] = csub(a, c);
        v[3] = csub(b, id);
    } else if constexpr (R == 3) {
        const T h = (T)0.86602540378443864676372317075294;  // sqrt(3)/2
        cpx<T> s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        cpx<T> base = {v[0].x - (T)0.5 * s.x, v[0].y - (T)0.5 * s.y};
        cpx<T> hd = {h * d.x, h * d.y};
        cpx<T> rot = cmul_i<T, SG>(hd);
        v[0] = cadd(v[0], s);
        v[1] = cadd(base, rot);
        v[2] = csub(base, rot);
    } else if constexpr (R == 5) {
        const T c1 = (T)0.30901699437494742410229341718282;   // cos(2pi/5)
        const T c2 = (T)-0.80901699437494742410229341718282;  // cos(4pi/5)
        const T s1 = (T)0.95105651629515357211643933337938;   // sin(2pi/5)
        const T s2 = (T)0.58778525229247312916870595463907;   // sin(4pi/5)
        cpx<T> a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
        cpx<T> b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
        cpx<T> p1 = {v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
        cpx<T> p2 = {v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
        cpx<T> q1 = {s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
        cpx<T> q2 = {s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
        cpx<T> iq1 = cmul_i<T, SG>(q1), iq2 = cmul_i<T, SG>(q2);
        v[0] = cadd(v[0], cadd(a1, a2));
        v[1] = cadd(p1, iq1);
        v[4] = csub(p1, iq1);
        v[2] = cadd(p2, iq2);
        v[3] = csub(p2, iq2);
    }
}

template <typename T, int N, int BATCH, int NT, int S, bool INV>
struct LdsFFT {
    static_assert(fft_len_ok(N), "transform length must be 2^a 3^b 5^c");

    template <int Ns>
    static __device__ __forceinline__ void pass(cpx<T>* buf, const cpx<T>* __restrict__ tw, int tid)
    {
        if constexpr (Ns < N) {
            constexpr int R = radix_of(N / Ns);
            constexpr int NB = N / R;         // butterflies per transform
            constexpr int TOT = BATCH * NB;   // butterflies per pass
            constexpr int PER = (TOT + NT - 1) / NT;
            cpx<T> v[PER][R];
            int dst[PER];
#pragma unroll
            for (int p = 0; p < PER; ++p) {
                const int idx = tid + p * NT;
                dst[p] = -1;
                if (TOT % NT == 0 || idx < TOT) {
                    const int b = idx / NB, j = idx - (idx / NB) * NB;
                    const int kk = j % Ns;
                    const cpx<T>* src = buf + b * S;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        cpx<T> a = src[j + r * NB];
                        if (Ns > 1 && r > 0) {
                            const cpx<T> w = tw[r * kk * (N / (Ns * R))];
                            a = INV ? cmulc(a, w) : cmul(a, w);
                        }
                        v[p][r] = a;
                    }
                    butterfly<T, R, INV>(v[p]);
                    dst[p] = b * S + (j / Ns) * Ns * R + kk;
                }
            }
            __syncthreads();
#pragma unroll
            for (int p = 0; p < PER; ++p) {
                if (TOT % NT == 0 || dst[p] >= 0) {
#pragma unroll
                    for (int r = 0; r < R; ++r) buf[dst[p] + r * Ns] = v[p][r];
                }
            }
            __syncthreads();
            pass<Ns * R>(buf, tw, tid);
        }
    }

    // buf must be visible to all NT threads (LDS); call with all threads.
    static __device__ __forceinline__ void run(cpx<T>* buf, const cpx<T>* __restrict__ tw, int tid)
    {
        pass<1>(buf, tw, tid);
    }
};

}  // namespace mhd
