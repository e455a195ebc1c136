// hydro_col.hip -- the hydrodynamic right-hand side of src/hydro_deriv.f90
// as one column-walking kernel for gfx950.
//
// A thread owns one (i,j) column and a run of zc planes.  Away from the
// boundary rows (4 <= k <= nz-5) it keeps sliding register windows of the
// fields whose z-derivatives the reference takes, so each such field is read
// from HBM once per plane: T, u, v, rho over k-3..k+3 (upwind and centred
// d1, d2), w over k-3..k+4, and over k-2..k+2 the products F bz bx, F bz by,
// the composite h = scko3 (ux + vy) + F/2 (bz^2 - bx^2 - by^2) (line 391),
// g = d1bydz(w, 1) and the x/y derivatives wx, wy.  The eight planes nearest
// the walls use the reference's boundary stencils (zstencil.hpp) with direct
// loads.
//
// The composite fluxes the reference differentiates whole are split
// linearly, which is exact up to rounding:
//   d/dx[scko3 divu + F/2 (bx^2-by^2-bz^2)]         (lines 181, 258)
//     = scko3 (d/dx d/dx u + d/dy(vx) + d1bydz(wx, 1)) + dxm
//   d/dy[scko3 divu + F/2 (by^2-bx^2-bz^2)]         (lines 182, 300)
//     = scko3 (d/dx(uy) + d/dy d/dy v + d1bydz(wy, 1)) + dym
//   d1bydz(h, 2) and d1bydz(4 scko3 g, 5)            (lines 183-184, 391, 434)
// with divu = ux + vy + d1bydz(w, 1).  The d/dx d/dx u and d/dy d/dy v parts
// ride in the x / y passes' second outputs (MUL_D2C), since they differ from
// d2bydx / d2bydy at the Nyquist mode.
#include <type_traits>

#include "launch.hpp"
#include "physics.hpp"
#include "zstencil.hpp"

namespace mhd {

namespace {

constexpr int HNT = 256;

// z-derivatives and window-held values at one point
struct ZD {
    double tzu, tzc, tzz, wzu, wzc, uzu, uzz, vzu, vzz, zzx, zzy, rzu, rzc, dzwx, dzwy, c3, c4;
    double rr, tt, u0, v0, w0, wx, wy;
};

template <typename T>
__device__ __forceinline__ double d1i(const T* f, int c, double h)
{
    return h * (f[c - 2] - 8.0 * f[c - 1] + 8.0 * f[c + 1] - f[c + 2]);
}
template <typename T>
__device__ __forceinline__ double d2i(const T* f, int c, double s)
{
    return s * (-f[c - 2] + 16.0 * f[c - 1] - 30.0 * f[c] + 16.0 * f[c + 1] - f[c + 2]);
}
template <typename T>
__device__ __forceinline__ double dupi(const T* f, int c, double h, bool up)
{
    if (up) return h * (-f[c - 3] + 6.0 * f[c - 2] - 18.0 * f[c - 1] + 10.0 * f[c] + 3.0 * f[c + 1]);
    return h * (f[c + 3] - 6.0 * f[c + 2] + 18.0 * f[c + 1] - 10.0 * f[c] - 3.0 * f[c - 1]);
}

// WALL = false: interior planes [4, nz-4), zc per thread, register windows;
// WALL = true: one boundary row per thread (planes 0..3 and nz-4..nz-1).
template <typename TW, bool MIX, bool WALL>
struct HydroColKernel {
    using Args = HydroColArgs;
    static constexpr int NT = HNT;
    static __device__ __forceinline__ void body(const HydroColArgs& a, uint3 bid, uint3 gdim);
};

template <typename TW, bool MIX, bool WALL>
__device__ __forceinline__ void HydroColKernel<TW, MIX, WALL>::body(const HydroColArgs& a, uint3 bid, uint3 gdim)
{
    (void)bid;
    (void)gdim;
    using TS = typename std::conditional<MIX, float, double>::type;
    const int nxy = a.g.nxy, nz = a.g.nz;
    const long col = (long)bid.x * HNT + threadIdx.x;
    if (col >= nxy) return;
    const double h = a.p.ho12, s = a.p.sc1;
    const double F = a.p.f, fo2 = a.p.fo2, scko3 = a.p.scko3;

    auto at = [&](const void* ptr, long i) -> double { return (double)((const TW*)ptr)[i]; };
    auto idx = [&](int kk) -> long { return col + (long)kk * nxy; };
    // full values at plane kk of this column
    auto T = [&](int kk) -> double {
        const double v = (double)((const TS*)a.t)[idx(kk)];
        if constexpr (MIX) return a.t0[kk] + v;
        else return v;
    };
    auto R = [&](int kk) -> double {
        const double v = (double)((const TS*)a.r)[idx(kk)];
        if constexpr (MIX) return a.r0[kk] + v;
        else return v;
    };
    auto U = [&](int kk) -> double { return (double)((const TS*)a.u)[idx(kk)]; };
    auto V = [&](int kk) -> double { return (double)((const TS*)a.v)[idx(kk)]; };
    auto W = [&](int kk) -> double { return (double)((const TS*)a.w)[idx(kk)]; };
    // B at plane kk -> F bz bx, F bz by, h
    auto bplane = [&](int kk, double& zx, double& zy, double& hh) {
        const long i = idx(kk);
        double bxv = at(a.bx, i), byv = at(a.by, i), bzv = at(a.bz, i);
        if constexpr (MIX) {
            bxv += a.bxm[kk];
            byv += a.bym[kk];
            bzv += a.bzm;
        }
        zx = (F * bzv) * bxv;
        zy = (F * bzv) * byv;
        const double t1 = fo2 * bxv * bxv, t2 = fo2 * byv * byv, t3 = fo2 * bzv * bzv;
        hh = scko3 * at(a.ux, i) + scko3 * at(a.vy, i) + t3 - t1 - t2;
    };

    // ---- point-wise assembly at plane k, in the reference's order of accumulation
    auto finish = [&](int k, const ZD& d) __attribute__((always_inline)) {
        const Phys& p = a.p;
        const long q = idx(k);
        const double rr = d.rr, tt = d.tt, u0 = d.u0, v0 = d.v0, w0 = d.w0;
        const double kap = a.kappa[k], dkap = a.dkappa[k];
        // 45-53, 82-110
        const double rx = at(a.rx, q), ry = at(a.ry, q), tx = at(a.tx, q), ty = at(a.ty, q);
        double rdot = -u0 * rx - v0 * ry;
        const double exv = -tt * rx, eyv = -tt * ry;
        double tdot = -rr * (u0 * tx + v0 * ty + w0 * d.tzu) +
                      p.gck * kap * (at(a.txx, q) + at(a.tyy, q) + d.tzz) + p.gck * dkap * d.tzu +
                      p.fzckg * (a.etar ? a.etar[k] : 1.0) * at(a.jsq, q);
        double udot = exv - rr * tx;
        double vdot = eyv - rr * ty;
        double wdot = -rr * d.tzc;
        // 139-168
        const double z1 = at(a.ux, q), z2 = at(a.vy, q), z3 = d.wzu, z4 = d.wzc;
        const double divu = z1 + z2 + z4;
        rdot = rdot - rr * divu;
        tdot = tdot - p.gm1 * tt * divu * rr +
               2 * p.sckg * (z1 * z1 + z2 * z2 + z4 * z4 - (1.0 / 3.0) * divu * divu);
        udot = udot - rr * u0 * z1;
        vdot = vdot - rr * v0 * z2;
        wdot = wdot - rr * w0 * z3;
        // 201-219
        {
            const double w1 = at(a.uy, q), w2 = at(a.vx, q);
            udot = udot - rr * v0 * w1;
            vdot = vdot - rr * u0 * w2;
            tdot = tdot + p.sckg * (w1 + w2) * (w1 + w2);
        }
        // 245-261
        udot = udot + at(a.ybb, q);
        vdot = vdot + at(a.xbb, q);
        {
            const double w3 = d.wx, w4 = d.uzu;
            udot = udot - rr * w0 * w4;
            wdot = wdot - rr * u0 * w3;
            tdot = tdot + p.sckg * (w3 + w4) * (w3 + w4);
        }
        // 287-309
        udot = udot + d.zzx;
        wdot = wdot + at(a.xzx, q) + p.thmplus1 * rr;
        {
            const double w1 = d.wy, w2 = d.vzu;
            vdot = vdot - rr * w0 * w2;
            wdot = wdot - rr * v0 * w1;
            tdot = tdot + p.sckg * (w1 + w2) * (w1 + w2);
        }
        udot = udot - rr * p.cor * (w0 * p.cpsi + v0 * p.spsi);
        vdot = vdot - rr * p.cor * (-u0 * p.spsi);
        wdot = wdot - rr * p.cor * (-u0 * p.cpsi);
        // 331-338
        vdot = vdot + d.zzy;
        wdot = wdot + at(a.yzy, q);
        // 360-432 (uxxc and vyyc carry sck*d2 plus the scko3 d/dx d/dx u and
        // d/dy d/dy v parts of the composite fluxes)
        udot = udot + p.sck * d.uzz;
        vdot = vdot + p.sck * d.vzz;
        udot = udot + at(a.uxxc, q);
        vdot = vdot + p.sck * at(a.vxx, q);
        wdot = wdot + p.sck * at(a.wxx, q);
        udot = udot + p.sck * at(a.uyy, q);
        vdot = vdot + at(a.vyyc, q);
        wdot = wdot + p.sck * at(a.wyy, q);
        // composite flux derivatives (181-184)
        udot = udot + (scko3 * (at(a.vxy, q) + d.dzwx) + at(a.dxm, q));
        vdot = vdot + (scko3 * (at(a.uxy, q) + d.dzwy) + at(a.dym, q));
        wdot = wdot + d.c3 + d.c4;
        // 455-470: d(rho)/dz on the walls from dw/dt = 0
        double rzu = d.rzu, rzc = d.rzc;
        if (k == 0 || k == nz - 1) {
            rzu = wdot / tt;
            rzc = wdot / tt;
        }
        rdot = rdot - w0 * rzu;
        wdot = wdot - tt * rzc;
        // 487-497
        const double ri = 1.0 / rr;
        const TW fu = (TW)(udot * ri), fv = (TW)(vdot * ri), fw = (TW)(wdot * ri), fr = (TW)rdot,
                 ft = (TW)(tdot * ri);
        if (!a.ab3) {
            ((TW*)a.rdot)[q] = fr;
            ((TW*)a.tdot)[q] = ft;
            ((TW*)a.udot)[q] = fu;
            ((TW*)a.vdot)[q] = fv;
            ((TW*)a.wdot)[q] = fw;
            return;
        }
        // AB3, src/main.F90:413-443, in the order x + c0 f3 + c1 f2 + c2 f1
        auto upd = [&](int i, double xo, TW f) {
            return xo + a.c0 * (double)f + a.c1 * (double)((const TW*)a.h2[i])[q] +
                   a.c2 * (double)((const TW*)a.h1[i])[q];
        };
        double xu, xv, xw, xr, xt;
        if constexpr (MIX) {
            xu = upd(0, a.x64[0][q], fu);
            xv = upd(1, a.x64[1][q], fv);
            xw = upd(2, a.x64[2][q], fw);
            xr = upd(3, a.x64[3][q], fr);
            xt = upd(4, a.x64[4][q], ft);
        } else {
            xu = upd(0, u0, fu);
            xv = upd(1, v0, fv);
            xw = upd(2, w0, fw);
            xr = upd(3, rr, fr);
            xt = upd(4, tt, ft);
        }
        TW rd = fr;
        if (xr < 0.05) {
            // density floor: clamp and zero its history
            xr = 0.05;
            rd = (TW)0;
            ((TW*)a.h3z[0])[q] = (TW)0;
            ((TW*)a.h3z[1])[q] = (TW)0;
        }
        // the RHS levels and the new state are next read a step or more later,
        // after the step has streamed other fields: stored non-temporally
        __builtin_nontemporal_store(fu, (TW*)a.udot + q);
        __builtin_nontemporal_store(fv, (TW*)a.vdot + q);
        __builtin_nontemporal_store(fw, (TW*)a.wdot + q);
        __builtin_nontemporal_store(rd, (TW*)a.rdot + q);
        __builtin_nontemporal_store(ft, (TW*)a.tdot + q);
        __builtin_nontemporal_store(xu, a.xnew[0] + q);
        __builtin_nontemporal_store(xv, a.xnew[1] + q);
        __builtin_nontemporal_store(xw, a.xnew[2] + q);
        __builtin_nontemporal_store(xr, a.xnew[3] + q);
        __builtin_nontemporal_store(xt, a.xnew[4] + q);
        if constexpr (MIX) {
            __builtin_nontemporal_store((float)xu, a.snew[0] + q);
            __builtin_nontemporal_store((float)xv, a.snew[1] + q);
            __builtin_nontemporal_store((float)xw, a.snew[2] + q);
            __builtin_nontemporal_store((float)(xr - a.r0[k]), a.snew[3] + q);
            __builtin_nontemporal_store((float)(xt - a.t0[k]), a.snew[4] + q);
        }
    };

    // boundary rows: the reference's row stencils with direct loads
    auto slow = [&](int k) {
        ZD d;
        const long i = idx(k);
        d.rr = R(k);
        d.tt = T(k);
        d.u0 = U(k);
        d.v0 = V(k);
        d.w0 = W(k);
        d.wx = at(a.wx, i);
        d.wy = at(a.wy, i);
        auto ZX = [&](int kk) { double zx, zy, hh; bplane(kk, zx, zy, hh); return zx; };
        auto ZY = [&](int kk) { double zx, zy, hh; bplane(kk, zx, zy, hh); return zy; };
        auto HH = [&](int kk) { double zx, zy, hh; bplane(kk, zx, zy, hh); return hh; };
        auto WX = [&](int kk) { return at(a.wx, idx(kk)); };
        auto WY = [&](int kk) { return at(a.wy, idx(kk)); };
        auto G = [&](int kk) { return dz1<double>(kk, nz, 1, h, W); };
        d.tzu = dz1up<double>(k, nz, 5, h, d.w0, T);
        d.tzc = dz1<double>(k, nz, 5, h, T);
        d.tzz = dz2<double>(k, nz, 0, s, T);
        d.wzu = dz1up<double>(k, nz, 1, h, d.w0, W);
        d.wzc = dz1<double>(k, nz, 1, h, W);
        d.uzu = dz1up<double>(k, nz, 2, h, d.w0, U);
        d.uzz = dz2<double>(k, nz, 3, s, U);
        d.vzu = dz1up<double>(k, nz, 2, h, d.w0, V);
        d.vzz = dz2<double>(k, nz, 3, s, V);
        d.zzx = dz1<double>(k, nz, 5, h, ZX);
        d.zzy = dz1<double>(k, nz, 5, h, ZY);
        d.rzu = dz1up<double>(k, nz, 0, h, d.w0, R);
        d.rzc = dz1<double>(k, nz, 0, h, R);
        d.dzwx = dz1<double>(k, nz, 1, h, WX);
        d.dzwy = dz1<double>(k, nz, 1, h, WY);
        d.c3 = dz1<double>(k, nz, 2, h, HH);
        d.c4 = 4.0 * scko3 * dz1<double>(k, nz, 5, h, G);
        finish(k, d);
    };

    if constexpr (WALL) {
        const int nb = nz < 8 ? nz : 8;
        const int pr = (int)bid.y;
        slow(pr < 4 ? pr : nz - (nb - pr));
        return;
    }
    // planes [4, nz-4) use the interior stencils of every derivative
    const int kf0 = 4 + (int)bid.y * a.zc;
    const int kf1 = (kf0 + a.zc < nz - 4) ? kf0 + a.zc : nz - 4;
    if (kf0 < kf1) {
        // windows: centre index 3 of the 7-wide, 4 of the 9-wide (w), 2 of the 5-wide
        double tw[7], uw[7], vw[7], rw[7], ww[9];
        double zxw[5], zyw[5], hw[5], wxw[5], wyw[5];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            tw[j + 1] = T(kf0 - 3 + j);
            uw[j + 1] = U(kf0 - 3 + j);
            vw[j + 1] = V(kf0 - 3 + j);
            rw[j + 1] = R(kf0 - 3 + j);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) ww[j + 1] = W(kf0 - 4 + j);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            bplane(kf0 - 2 + j, zxw[j + 1], zyw[j + 1], hw[j + 1]);
            wxw[j + 1] = at(a.wx, idx(kf0 - 2 + j));
            wyw[j + 1] = at(a.wy, idx(kf0 - 2 + j));
        }
        for (int k = kf0; k < kf1; ++k) {
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                tw[j] = tw[j + 1];
                uw[j] = uw[j + 1];
                vw[j] = vw[j + 1];
                rw[j] = rw[j + 1];
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) ww[j] = ww[j + 1];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                zxw[j] = zxw[j + 1];
                zyw[j] = zyw[j + 1];
                hw[j] = hw[j + 1];
                wxw[j] = wxw[j + 1];
                wyw[j] = wyw[j + 1];
            }
            tw[6] = T(k + 3);
            uw[6] = U(k + 3);
            vw[6] = V(k + 3);
            rw[6] = R(k + 3);
            ww[8] = W(k + 4);
            bplane(k + 2, zxw[4], zyw[4], hw[4]);
            wxw[4] = at(a.wx, idx(k + 2));
            wyw[4] = at(a.wy, idx(k + 2));

            ZD d;
            d.rr = rw[3];
            d.tt = tw[3];
            d.u0 = uw[3];
            d.v0 = vw[3];
            d.w0 = ww[4];
            d.wx = wxw[2];
            d.wy = wyw[2];
            const bool up = d.w0 >= 0.0;
            d.tzu = dupi(tw, 3, h, up);
            d.tzc = d1i(tw, 3, h);
            d.tzz = d2i(tw, 3, s);
            d.wzu = dupi(ww, 4, h, up);
            // g = d1bydz(w, 1) at k-2..k+2 (interior rows)
            const double g0 = d1i(ww, 2, h), g1 = d1i(ww, 3, h), g2 = d1i(ww, 4, h);
            const double g3 = d1i(ww, 5, h), g4 = d1i(ww, 6, h);
            d.wzc = g2;
            d.c4 = 4.0 * scko3 * (h * (g0 - 8.0 * g1 + 8.0 * g3 - g4));
            d.uzu = dupi(uw, 3, h, up);
            d.uzz = d2i(uw, 3, s);
            d.vzu = dupi(vw, 3, h, up);
            d.vzz = d2i(vw, 3, s);
            d.zzx = d1i(zxw, 2, h);
            d.zzy = d1i(zyw, 2, h);
            d.rzu = dupi(rw, 3, h, up);
            d.rzc = d1i(rw, 3, h);
            d.dzwx = d1i(wxw, 2, h);
            d.dzwy = d1i(wyw, 2, h);
            d.c3 = d1i(hw, 2, h);
            finish(k, d);
        }
    }
}

}  // namespace

int launch_hydro_col(const HydroColArgs& a, int mixed, hipStream_t s)
{
    if (a.zc <= 0) return -10;
    const unsigned bx = (unsigned)((a.g.nxy + HNT - 1) / HNT);
    const int nin = a.g.nz - 8;                   // interior planes
    if (nin > 0) {
        const dim3 grid(bx, (unsigned)((nin + a.zc - 1) / a.zc));
        const int e = mixed ? mhd_launch<HydroColKernel<float, true, false>>(grid, dim3(HNT), 0, s, a)
                            : mhd_launch<HydroColKernel<double, false, false>>(grid, dim3(HNT), 0, s, a);
        if (e) return e;
    }
    const dim3 gw(bx, (unsigned)(a.g.nz < 8 ? a.g.nz : 8));
    return mixed ? mhd_launch<HydroColKernel<float, true, true>>(gw, dim3(HNT), 0, s, a)
                 : mhd_launch<HydroColKernel<double, false, true>>(gw, dim3(HNT), 0, s, a);
}

}  // namespace mhd
