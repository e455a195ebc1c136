// zstencil.hpp -- vertical 4th-order finite differences (device side).
//
// Row-by-row restatement of the serial branches of the reference's d1bydz,
// d1upbydz and d2bydz (src/dbydz.F90:1685-1732, 2353-2423, 3010-3082) built
// from its stencil subroutines (src/dbydz.F90:371-1558).  k is 0-based;
// F(kk) returns the input at plane kk for the calling thread's (i,j) column.
// h = ho12 = (nz-1)/12, s = sc1 = (nz-1)^2/12.  Flags are validated on the
// host (unknown flag -> error 7 before any launch).
#pragma once

#include "common.hpp"

namespace mhd {

enum ZKind { ZK_D1 = 0, ZK_D1UP = 1, ZK_D2 = 2 };

// d1bydz rows: topdz / topdz_low / interior / lastdz_up / lastdz
template <typename T, typename Fn>
__device__ __forceinline__ T dz1(int k, int nz, int flag, T h, Fn F)
{
    if (k == 0) {
        if (flag == 1) return h * ((T)14 * F(0) + (T)36 * F(1) - (T)18 * F(2) + (T)4 * F(3));
        if (flag == 5) return h * ((T)-22 * F(0) + (T)36 * F(1) - (T)18 * F(2) + (T)4 * F(3));
        return (T)0;
    }
    if (k == 1) return h * ((T)-4 * F(0) - (T)6 * F(1) + (T)12 * F(2) - (T)2 * F(3));
    if (k == nz - 2)
        return h * ((T)4 * F(nz - 1) + (T)6 * F(nz - 2) - (T)12 * F(nz - 3) + (T)2 * F(nz - 4));
    if (k == nz - 1) {
        if (flag == 1)
            return h * ((T)-14 * F(nz - 1) - (T)36 * F(nz - 2) + (T)18 * F(nz - 3) - (T)4 * F(nz - 4));
        if (flag == 5)
            return h * ((T)22 * F(nz - 1) - (T)36 * F(nz - 2) + (T)18 * F(nz - 3) - (T)4 * F(nz - 4));
        return (T)0;
    }
    return h * (F(k - 2) - (T)8 * F(k - 1) + (T)8 * F(k + 1) - F(k + 2));
}

// d1upbydz rows: as dz1 at k = 0, 1, nz-2, nz-1; dup_up at k = 2,
// dup_low at k = nz-3, dupinterior between; branch on w >= 0
template <typename T, typename Fn>
__device__ __forceinline__ T dz1up(int k, int nz, int flag, T h, T wv, Fn F)
{
    if (k <= 1 || k >= nz - 2) return dz1<T>(k, nz, flag, h, F);
    const bool up = wv >= (T)0;
    if (k == 2) {
        if (up) return h * ((T)2 * F(0) - (T)12 * F(1) + (T)6 * F(2) + (T)4 * F(3));
        return h * (F(5) - (T)6 * F(4) + (T)18 * F(3) - (T)10 * F(2) - (T)3 * F(1));
    }
    if (k == nz - 3) {
        if (up) return h * (-F(k - 3) + (T)6 * F(k - 2) - (T)18 * F(k - 1) + (T)10 * F(k) + (T)3 * F(k + 1));
        return h * ((T)-2 * F(k + 2) + (T)12 * F(k + 1) - (T)6 * F(k) - (T)4 * F(k - 1));
    }
    if (up) return h * (-F(k - 3) + (T)6 * F(k - 2) - (T)18 * F(k - 1) + (T)10 * F(k) + (T)3 * F(k + 1));
    return h * (F(k + 3) - (T)6 * F(k + 2) + (T)18 * F(k + 1) - (T)10 * F(k) - (T)3 * F(k - 1));
}

// d2bydz rows: d2topdz / d2topdz_low / d2interior / d2lastdz_up / d2lastdz
template <typename T, typename Fn>
__device__ __forceinline__ T dz2(int k, int nz, int flag, T s, Fn F)
{
    if (k == 0) {
        if (flag == 3)
            return s * ((T)45 * F(0) - (T)154 * F(1) + (T)214 * F(2) - (T)156 * F(3) + (T)61 * F(4) - (T)10 * F(5));
        return (T)0;
    }
    if (k == 1)
        return s * ((T)10 * F(0) - (T)15 * F(1) - (T)4 * F(2) + (T)14 * F(3) - (T)6 * F(4) + F(5));
    if (k == nz - 2)
        return s * ((T)10 * F(nz - 1) - (T)15 * F(nz - 2) - (T)4 * F(nz - 3) + (T)14 * F(nz - 4) -
                    (T)6 * F(nz - 5) + F(nz - 6));
    if (k == nz - 1) {
        if (flag == 3)
            return s * ((T)45 * F(nz - 1) - (T)154 * F(nz - 2) + (T)214 * F(nz - 3) - (T)156 * F(nz - 4) +
                        (T)61 * F(nz - 5) - (T)10 * F(nz - 6));
        return (T)0;
    }
    return s * (-F(k - 2) + (T)16 * F(k - 1) - (T)30 * F(k) + (T)16 * F(k + 1) - F(k + 2));
}

inline bool zflag_ok(int kind, int flag)
{
    if (kind == ZK_D2) return flag == 0 || flag == 3;
    return flag == 0 || flag == 1 || flag == 2 || flag == 5;
}

}  // namespace mhd
