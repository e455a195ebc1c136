// physics.hip -- vertical derivatives, right-hand-side assembly, magnetic
// spectral operations, AB3 update, boundary conditions and reductions for
// gfx950.  Point-wise kernels use one thread per grid point with x fastest,
// so every wavefront reads 64 consecutive elements of each operand.
#include "launch.hpp"
#include "physics.hpp"
#include "spectral.hpp"
#include "zstencil.hpp"

namespace mhd {

static constexpr int NT = 256;

static inline unsigned nblk(long n, int nt = NT) { return (unsigned)((n + nt - 1) / nt); }

// ---------------------------------------------------------------------------
// batched z-derivatives
// ---------------------------------------------------------------------------
// A thread owns one (i,j) column and KC consecutive planes.  Chunks clear of
// the boundary rows hold planes k0-3 .. k0+KC+2 in a register window and
// apply the interior stencils at compile-time offsets; the chunks touching
// the first or last three planes apply the row-specific stencils.
constexpr int ZKC = 16;

// TC: compute/output type; TIN: input element type; KIND: ZKind; PROD: the
// input is scale*in*in1.  w (upwind selector) has the element type TW.
template <typename TC, typename TIN, typename TW, int KIND, bool PROD>
struct ZKernel {
    using Args = ZLaunch;
    static constexpr int NT = 256;
    static __device__ __forceinline__ void body(const ZLaunch& L, uint3 bid, uint3 gdim);
};

template <typename TC, typename TIN, typename TW, int KIND, bool PROD>
__device__ __forceinline__ void ZKernel<TC, TIN, TW, KIND, PROD>::body(const ZLaunch& L, uint3 bid, uint3 gdim)
{
    (void)bid;
    (void)gdim;
    const ZJob& J = L.job[bid.z];
    const int nxy = L.g.nxy, nz = L.g.nz;
    const long col = (long)bid.x * NT + threadIdx.x;
    if (col >= nxy) return;
    const TIN* __restrict__ in = (const TIN*)J.in;
    const TIN* __restrict__ in1 = (const TIN*)J.in1;
    const TC scale = (TC)J.scale;
    TC* __restrict__ out = (TC*)J.out;
    const int flag = J.flag;
    const TW* __restrict__ wp = (const TW*)L.w;
    const TC h = (TC)L.ho12, s = (TC)L.sc1;
    auto ld = [=](int kk) -> TC {
        const long i = col + (long)kk * nxy;
        TC v = (TC)in[i];
        if constexpr (PROD) v = (scale * v) * (TC)in1[i];
        return v;
    };
    const int k0 = bid.y * ZKC;
    if (k0 >= 3 && k0 + ZKC <= nz - 3) {
        constexpr int LO = (KIND == ZK_D1UP) ? 0 : 1;          // first window slot used
        constexpr int HI = (KIND == ZK_D1UP) ? ZKC + 6 : ZKC + 5;
        TC win[ZKC + 6];
#pragma unroll
        for (int t = LO; t < HI; ++t) win[t] = ld(k0 - 3 + t);
#pragma unroll
        for (int t = 0; t < ZKC; ++t) {
            const int c = t + 3;
            const long o = col + (long)(k0 + t) * nxy;
            TC d;
            if constexpr (KIND == ZK_D1) {
                d = h * (win[c - 2] - (TC)8 * win[c - 1] + (TC)8 * win[c + 1] - win[c + 2]);
            } else if constexpr (KIND == ZK_D2) {
                d = s * (-win[c - 2] + (TC)16 * win[c - 1] - (TC)30 * win[c] + (TC)16 * win[c + 1] - win[c + 2]);
            } else {
                const TC wv = (TC)wp[o];
                d = (wv >= (TC)0)
                        ? h * (-win[c - 3] + (TC)6 * win[c - 2] - (TC)18 * win[c - 1] + (TC)10 * win[c] +
                               (TC)3 * win[c + 1])
                        : h * (win[c + 3] - (TC)6 * win[c + 2] + (TC)18 * win[c + 1] - (TC)10 * win[c] -
                               (TC)3 * win[c - 1]);
            }
            out[o] = d;
        }
    } else {
        const int k1 = (k0 + ZKC < nz) ? k0 + ZKC : nz;
        for (int k = k0; k < k1; ++k) {
            const long o = col + (long)k * nxy;
            TC d;
            if constexpr (KIND == ZK_D1) d = dz1<TC>(k, nz, flag, h, ld);
            else if constexpr (KIND == ZK_D2) d = dz2<TC>(k, nz, flag, s, ld);
            else d = dz1up<TC>(k, nz, flag, h, (TC)wp[o], ld);
            out[o] = d;
        }
    }
}

template <typename TC, typename TIN, typename TW, int KIND, bool PROD>
static int launch_z_group(ZLaunch& G, hipStream_t s)
{
    dim3 grid(nblk(G.g.nxy), (unsigned)((G.g.nz + ZKC - 1) / ZKC), G.njobs);
    return mhd_launch<ZKernel<TC, TIN, TW, KIND, PROD>>(grid, dim3(NT), 0, s, G);
}

template <typename TC, typename TIN, typename TW>
static int launch_z_kind(ZLaunch& G, int kind, bool prod, hipStream_t s)
{
    switch (kind * 2 + (prod ? 1 : 0)) {
    case 0: return launch_z_group<TC, TIN, TW, ZK_D1, false>(G, s);
    case 1: return launch_z_group<TC, TIN, TW, ZK_D1, true>(G, s);
    case 2: return launch_z_group<TC, TIN, TW, ZK_D1UP, false>(G, s);
    case 3: return launch_z_group<TC, TIN, TW, ZK_D1UP, true>(G, s);
    case 4: return launch_z_group<TC, TIN, TW, ZK_D2, false>(G, s);
    case 5: return launch_z_group<TC, TIN, TW, ZK_D2, true>(G, s);
    }
    return -23;
}

// Jobs are grouped by (input type, kind, product load); one launch per group.
int launch_z(const ZLaunch& L, hipStream_t s)
{
    if (L.njobs <= 0) return 0;
    for (int i = 0; i < L.njobs; ++i)
        if (!zflag_ok(L.job[i].kind, L.job[i].flag)) return 7;
    bool done[Z_MAX_JOBS] = {};
    for (int i = 0; i < L.njobs; ++i) {
        if (done[i]) continue;
        const ZJob& a = L.job[i];
        ZLaunch G = L;
        G.njobs = 0;
        for (int j = i; j < L.njobs; ++j) {
            const ZJob& b = L.job[j];
            if (!done[j] && b.in_f32 == a.in_f32 && b.kind == a.kind && b.load_op == a.load_op &&
                b.out_f32 == a.out_f32) {
                G.job[G.njobs++] = b;
                done[j] = true;
            }
        }
        const bool prod = a.load_op == LD_PROD;
        int e;
        if (a.out_f32) {
            if (L.w_f32) e = a.in_f32 ? launch_z_kind<float, float, float>(G, a.kind, prod, s)
                                      : launch_z_kind<float, double, float>(G, a.kind, prod, s);
            else e = a.in_f32 ? launch_z_kind<float, float, double>(G, a.kind, prod, s)
                              : launch_z_kind<float, double, double>(G, a.kind, prod, s);
        } else {
            if (a.in_f32) return -24;   // FP32 input with FP64 output is not used
            e = L.w_f32 ? launch_z_kind<double, double, float>(G, a.kind, prod, s)
                        : launch_z_kind<double, double, double>(G, a.kind, prod, s);
        }
        if (e) return e;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// magnetic part on the double-halfcomplex layout (src/magnetic.f90)
// ---------------------------------------------------------------------------
// Workgroups are dealt to the 8 XCDs round-robin.  Remapping block b to tile
// (b % 8) * (nb / 8) + b / 8 lets each XCD walk a contiguous eighth of the tiles
// (whole planes), so the index-swap kernels find a plane's partner rows in
// their own L2.  The tail beyond 8 * (nb / 8) keeps the identity.
__device__ __forceinline__ long xcd_tile(long b, long nb)
{
    const long per = nb / 8;
    if (b >= 8 * per) return b;
    return (b % 8) * per + b / 8;
}

// spectral d/dx at (p,q): index swap along x, src/magnetic.f90:141-150
template <typename TA>
__device__ __forceinline__ double sdx(const TA* a, long base, int p, int q, int nx, const double* akx1)
{
    if (p == 0 || p == nx / 2) return 0.0;
    const double v = (double)a[base + (long)q * nx + (nx - p)] * akx1[p];
    return (p < nx / 2) ? -v : v;
}
// spectral d/dy at (p,q): index swap along y, src/magnetic.f90:155-164
template <typename TA>
__device__ __forceinline__ double sdy(const TA* a, long base, int p, int q, int nx, int ny, const double* aky1)
{
    if (q == 0 || q == ny / 2) return 0.0;
    const double v = (double)a[base + (long)(ny - q) * nx + p] * aky1[q];
    return (q < ny / 2) ? -v : v;
}
// 2/3 truncation band, src/magnetic.f90:666-717
__device__ __forceinline__ bool in_band(int p, int q, int nx, int ny, int icx, int icy)
{
    return (p >= nx / 2 - icx && p <= nx / 2 + icx) || (q >= ny / 2 - icy && q <= ny / 2 + icy);
}

// src/magnetic.f90:362-407: cc = dEx/dy - dEy/dx, d = dEx/dx + dEy/dy by index
// swaps, one thread per mirror quad as in bandj_spec_kernel
template <typename TW>
struct MagSpec1Kernel {
    using Args = MagArgs;
    static constexpr int NT = 256;
    static __device__ __forceinline__ void body(const MagArgs& a, uint3 bid, uint3 gdim);
};

template <typename TW>
__device__ __forceinline__ void MagSpec1Kernel<TW>::body(const MagArgs& a, uint3 bid, uint3 gdim)
{
    (void)bid;
    (void)gdim;
    const int nx = a.g.nx, ny = a.g.ny, nxy = a.g.nxy;
    const int hx = nx / 2 + 1, hy = ny / 2 + 1;
    const long perk = (long)hx * hy;
    const long t = xcd_tile(bid.x, gdim.x) * NT + threadIdx.x;
    if (t >= perk * a.g.nz) return;
    const int k = (int)(t / perk);
    const int rem = (int)(t - (long)k * perk);
    const int jj = rem / hx, ii = rem - jj * hx;
    const long base = (long)k * nxy;
    const bool xm = !(ii == 0 || ii == nx / 2), ym = !(jj == 0 || jj == ny / 2);
    const int col[2] = {ii, nx - ii}, row[2] = {jj, ny - jj};
    const TW* ex = (const TW*)a.exh;
    const TW* ey = (const TW*)a.eyh;
    double X[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, Y[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int sy = 0; sy < 2; ++sy)
#pragma unroll
        for (int sx = 0; sx < 2; ++sx) {
            if ((sy && !ym) || (sx && !xm)) continue;
            const long q = base + (long)row[sy] * nx + col[sx];
            X[sy][sx] = (double)ex[q];
            Y[sy][sx] = (double)ey[q];
        }
#pragma unroll
    for (int sy = 0; sy < 2; ++sy)
#pragma unroll
        for (int sx = 0; sx < 2; ++sx) {
            if ((sy && !ym) || (sx && !xm)) continue;
            const int p = col[sx], q = row[sy];
            double dxex = 0.0, dxey = 0.0, dyex = 0.0, dyey = 0.0;
            if (xm) {
                const double sg = (p < nx / 2) ? -1.0 : 1.0;
                dxex = sg * (X[sy][1 - sx] * a.akx1[p]);
                dxey = sg * (Y[sy][1 - sx] * a.akx1[p]);
            }
            if (ym) {
                const double sg = (q < ny / 2) ? -1.0 : 1.0;
                dyex = sg * (X[1 - sy][sx] * a.aky1[q]);
                dyey = sg * (Y[1 - sy][sx] * a.aky1[q]);
            }
            const long o = base + (long)q * nx + p;
            ((TW*)a.cc)[o] = (TW)(dyex - dxey);
            ((TW*)a.d)[o] = (TW)(dxex + dyey);
        }
}

// mean-field equations, src/magnetic.f90:328-352 (one workgroup)
template <typename TW>
struct MeanEmfKernel {
    using Args = MagArgs;
    static constexpr int NT = 1024;
    static __device__ __forceinline__ void body(const MagArgs& a, uint3 bid, uint3 gdim);
};

template <typename TW>
__device__ __forceinline__ void MeanEmfKernel<TW>::body(const MagArgs& a, uint3 bid, uint3 gdim)
{
    (void)bid;
    (void)gdim;
    extern __shared__ double sh[];
    const int nz = a.g.nz;
    double* w1d = sh;
    double* w2d = sh + nz;
    double* bxb = sh + 2 * nz;
    double* byb = sh + 3 * nz;
    for (int k = threadIdx.x; k < nz; k += blockDim.x) {
        const long b = (long)k * a.g.nxy;
        w1d[k] = a.p.xn * (double)((const TW*)a.exh)[b];
        w2d[k] = a.p.xn * (double)((const TW*)a.eyh)[b];
        bxb[k] = a.bxbar[k];
        byb[k] = a.bybar[k];
    }
    __syncthreads();
    const double h = a.p.ho12, s = a.p.sc1;
    for (int k = threadIdx.x; k < nz; k += blockDim.x) {
        const double dbx = dz1<double>(k, nz, 5, h, [&](int kk) { return w2d[kk]; });
        const double dby = dz1<double>(k, nz, 5, h, [&](int kk) { return w1d[kk]; });
        const double d2x = dz2<double>(k, nz, 3, s, [&](int kk) { return bxb[kk]; });
        const double d2y = dz2<double>(k, nz, 3, s, [&](int kk) { return byb[kk]; });
        a.bxbardot[k] = a.p.zck * d2x - dbx;
        a.bybardot[k] = a.p.zck * d2y + dby;
    }
}

// poldot, tordot, src/magnetic.f90:429-479.  Column walk: a thread owns one
// (i,j) column and a.zc planes; pol and tor stay in 5-plane register windows
// so each plane is read once; the two planes next to each wall use the
// one-sided d2bydz rows (flag 3) with direct loads.  D_z of d (line 419,
// d1bydz flag 5) comes from a third window, in the working precision (as a
// separate z pass would compute and store it), instead of a stored array.
template <typename TW>
struct MagRhsKernel {
    using Args = MagArgs;
    static constexpr int NT = 256;
    static __device__ __forceinline__ void body(const MagArgs& a, uint3 bid, uint3 gdim);
};

template <typename TW>
__device__ __forceinline__ void MagRhsKernel<TW>::body(const MagArgs& a, uint3 bid, uint3 gdim)
{
    (void)bid;
    (void)gdim;
    using TW2 = TW __attribute__((ext_vector_type(2)));
    using D2 = double __attribute__((ext_vector_type(2)));
    const int nx = a.g.nx, nxy = a.g.nxy, nz = a.g.nz;
    // a thread owns two neighbouring columns of one row (nx is even): every
    // load and store moves two elements
    const long col = 2 * ((long)bid.x * NT + threadIdx.x);
    if (col >= nxy) return;
    const int k0 = bid.y * a.zc;
    const int k1 = (k0 + a.zc < nz) ? k0 + a.zc : nz;
    const int jq = (int)(col / nx), ip0 = (int)(col - (long)jq * nx);
    double k2[2];
    bool dc[2], band[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        k2[c] = a.akx2[ip0 + c] + a.aky2[jq];
        dc[c] = (ip0 + c == 0 && jq == 0);
        band[c] = in_band(ip0 + c, jq, nx, a.g.ny, a.icutx, a.icuty);
    }
    const double* pol = a.pol;
    const double* tor = a.tor;
    const TW* __restrict__ dd = (const TW*)a.d;
    const double zck = a.p.zck, s = a.p.sc1;
    const TW hz = (TW)a.p.ho12;
    auto ld2 = [](const void* p, long q) { return *(const TW2*)((const TW*)p + q); };
    auto ldd = [](const double* p, long q) { return *(const D2*)(p + q); };
    // interior planes [2, nz-2) use the centred stencils from the windows
    const int ki0 = k0 > 2 ? k0 : 2, ki1 = k1 < nz - 2 ? k1 : nz - 2;
    double pw[2][5], tw[2][5];
    TW dw[2][5];
    if (ki0 < ki1) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long q = col + (long)(ki0 - 2 + j) * nxy;
            const D2 p2 = ldd(pol, q), t2 = ldd(tor, q);
            const TW2 d2 = ld2(dd, q);
            pw[0][j + 1] = p2.x; pw[1][j + 1] = p2.y;
            tw[0][j + 1] = t2.x; tw[1][j + 1] = t2.y;
            dw[0][j + 1] = d2.x; dw[1][j + 1] = d2.y;
        }
    }
    for (int k = k0; k < k1; ++k) {
        const long q = col + (long)k * nxy;
        double pzz[2], tzz[2], p0[2], t0[2], dzv[2];
        if (k >= ki0 && k < ki1) {
            const D2 p2 = ldd(pol, q + 2L * nxy), t2 = ldd(tor, q + 2L * nxy);
            const TW2 d2 = ld2(dd, q + 2L * nxy);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    pw[c][j] = pw[c][j + 1];
                    tw[c][j] = tw[c][j + 1];
                    dw[c][j] = dw[c][j + 1];
                }
            }
            pw[0][4] = p2.x; pw[1][4] = p2.y;
            tw[0][4] = t2.x; tw[1][4] = t2.y;
            dw[0][4] = d2.x; dw[1][4] = d2.y;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                pzz[c] = s * (-pw[c][0] + 16.0 * pw[c][1] - 30.0 * pw[c][2] + 16.0 * pw[c][3] - pw[c][4]);
                tzz[c] = s * (-tw[c][0] + 16.0 * tw[c][1] - 30.0 * tw[c][2] + 16.0 * tw[c][3] - tw[c][4]);
                dzv[c] = (double)(hz * (dw[c][0] - (TW)8 * dw[c][1] + (TW)8 * dw[c][3] - dw[c][4]));
                p0[c] = pw[c][2];
                t0[c] = tw[c][2];
            }
        } else {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                pzz[c] = dz2<double>(k, nz, 3, s, [&](int kk) { return pol[col + c + (long)kk * nxy]; });
                tzz[c] = dz2<double>(k, nz, 3, s, [&](int kk) { return tor[col + c + (long)kk * nxy]; });
                dzv[c] = (double)dz1<TW>(k, nz, 5, hz, [&](int kk) { return dd[col + c + (long)kk * nxy]; });
                p0[c] = pol[q + c];
                t0[c] = tor[q + c];
            }
        }
        const TW2 cc2 = ld2(a.cc, q), ez2 = ld2(a.ezh, q);
        const double ccv[2] = {(double)cc2.x, (double)cc2.y}, ezv[2] = {(double)ez2.x, (double)ez2.y};
        TW pd[2], td[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const double il_c = dc[c] ? 0.0 : ccv[c] / k2[c];
            const double il_d = dc[c] ? 0.0 : dzv[c] / k2[c];
            pd[c] = (TW)(zck * p0[c] * k2[c] + zck * pzz[c] + il_c);
            td[c] = (TW)(zck * t0[c] * k2[c] + zck * tzz[c] - il_d + ezv[c]);
        }
        TW2 pdv, tdv;
        pdv.x = pd[0]; pdv.y = pd[1];
        tdv.x = td[0]; tdv.y = td[1];
        *(TW2*)((TW*)a.poldot + q) = pdv;
        *(TW2*)((TW*)a.tordot + q) = tdv;
        if (a.ab3) {
            // x + c0 f3 + c1 f2 + c2 f1, then the 2/3 band of pol and tor cleared
            const TW2 ph2 = ld2(a.ph2, q), ph1 = ld2(a.ph1, q), th2 = ld2(a.th2, q), th1 = ld2(a.th1, q);
            const double ph2v[2] = {(double)ph2.x, (double)ph2.y}, ph1v[2] = {(double)ph1.x, (double)ph1.y};
            const double th2v[2] = {(double)th2.x, (double)th2.y}, th1v[2] = {(double)th1.x, (double)th1.y};
            D2 pn, tn;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                double pv = p0[c] + a.c0 * (double)pd[c] + a.c1 * ph2v[c] + a.c2 * ph1v[c];
                double tv = t0[c] + a.c0 * (double)td[c] + a.c1 * th2v[c] + a.c2 * th1v[c];
                if (band[c]) {
                    pv = 0.0;
                    tv = 0.0;
                }
                if (c == 0) { pn.x = pv; tn.x = tv; }
                else { pn.y = pv; tn.y = tv; }
            }
            *(D2*)(a.polnew + q) = pn;
            *(D2*)(a.tornew + q) = tn;
        }
    }
}

// spectral B and J pieces of bandj, src/magnetic.f90:519-617.  The index swaps
// pair the points (j, i), (j, nx-i), (ny-j, i), (ny-j, nx-i) of a plane and
// nothing else: one thread takes such a quad (j <= ny/2, i <= nx/2; fewer
// points on the lines i = 0, nx/2 and j = 0, ny/2, which are their own
// mirrors), reads pol, tor and d pol/dz once at each of its points and writes
// all six outputs there, with the arithmetic of the per-point formulas.
template <typename TW>
struct BandjSpecKernel {
    using Args = MagArgs;
    static constexpr int NT = 256;
    static __device__ __forceinline__ void body(const MagArgs& a, uint3 bid, uint3 gdim);
};

template <typename TW>
__device__ __forceinline__ void BandjSpecKernel<TW>::body(const MagArgs& a, uint3 bid, uint3 gdim)
{
    (void)bid;
    (void)gdim;
    const int nx = a.g.nx, ny = a.g.ny, nxy = a.g.nxy;
    const int hx = nx / 2 + 1, hy = ny / 2 + 1;
    const long perk = (long)hx * hy;
    const long t = xcd_tile(bid.x, gdim.x) * NT + threadIdx.x;
    if (t >= perk * a.g.nz) return;
    const int k = (int)(t / perk);
    const int rem = (int)(t - (long)k * perk);
    const int jj = rem / hx, ii = rem - jj * hx;
    const long base = (long)k * nxy;
    const bool xm = !(ii == 0 || ii == nx / 2), ym = !(jj == 0 || jj == ny / 2);
    const int col[2] = {ii, nx - ii}, row[2] = {jj, ny - jj};
    const TW* pz = (const TW*)a.polz;
    double P[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, T[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, Z[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int sy = 0; sy < 2; ++sy)
#pragma unroll
        for (int sx = 0; sx < 2; ++sx) {
            if ((sy && !ym) || (sx && !xm)) continue;
            const long q = base + (long)row[sy] * nx + col[sx];
            P[sy][sx] = a.pol[q];
            T[sy][sx] = a.tor[q];
            Z[sy][sx] = (double)pz[q];
        }
#pragma unroll
    for (int sy = 0; sy < 2; ++sy)
#pragma unroll
        for (int sx = 0; sx < 2; ++sx) {
            if ((sy && !ym) || (sx && !xm)) continue;
            const int p = col[sx], q = row[sy];
            const double k2 = a.akx2[p] + a.aky2[q];
            // d/dx and d/dy by index swap (src/magnetic.f90:141-164)
            double dxz = 0.0, dxt = 0.0, dxb = 0.0, dyz = 0.0, dyt = 0.0, dyb = 0.0;
            if (xm) {
                const int pp = col[1 - sx];
                const double sg = (p < nx / 2) ? -1.0 : 1.0;
                dxz = sg * (Z[sy][1 - sx] * a.akx1[p]);
                dxt = sg * (T[sy][1 - sx] * a.akx1[p]);
                dxb = sg * ((-1.0 * P[sy][1 - sx] * (a.akx2[pp] + a.aky2[q])) * a.akx1[p]);
            }
            if (ym) {
                const int qq = row[1 - sy];
                const double sg = (q < ny / 2) ? -1.0 : 1.0;
                dyz = sg * (Z[1 - sy][sx] * a.aky1[q]);
                dyt = sg * (T[1 - sy][sx] * a.aky1[q]);
                dyb = sg * ((-1.0 * P[1 - sy][sx] * (a.akx2[p] + a.aky2[qq])) * a.aky1[q]);
            }
            const long o = base + (long)q * nx + p;
            ((TW*)a.bxh)[o] = (TW)(dxz + dyt);
            ((TW*)a.byh)[o] = (TW)(dyz - dxt);
            ((TW*)a.bzh)[o] = (TW)(-1.0 * P[sy][sx] * k2);
            ((TW*)a.jzh)[o] = (TW)(-1.0 * T[sy][sx] * k2);
            ((TW*)a.dybzh)[o] = (TW)dyb;
            ((TW*)a.dxbzh)[o] = (TW)dxb;
        }
}

// J^2, src/magnetic.f90:630-645 (dybzh/dxbzh/jzh hold physical-space fields here)
// J^2 (src/magnetic.f90:596-645) and, for the eta(z) extension, J itself.
// Column walk with 5-plane windows of bx, by for their d/dz (flag 5).  A
// thread owns two neighbouring columns (nx is even), so every load and store
// moves two elements.
template <typename TW>
struct JsqKernel {
    using Args = MagArgs;
    static constexpr int NT = 256;
    static __device__ __forceinline__ void body(const MagArgs& a, uint3 bid, uint3 gdim);
};

template <typename TW>
__device__ __forceinline__ void JsqKernel<TW>::body(const MagArgs& a, uint3 bid, uint3 gdim)
{
    (void)bid;
    (void)gdim;
    using TW2 = TW __attribute__((ext_vector_type(2)));
    const int nxy = a.g.nxy, nz = a.g.nz;
    const long col = 2 * ((long)bid.x * NT + threadIdx.x);
    if (col >= nxy) return;
    const int k0 = bid.y * a.zc;
    const int k1 = (k0 + a.zc < nz) ? k0 + a.zc : nz;
    const TW* by = (const TW*)a.by;
    const TW* bx = (const TW*)a.bx;
    const double h = a.p.ho12;
    auto ld2 = [](const void* p, long q) { return *(const TW2*)((const TW*)p + q); };
    // interior planes [2, nz-2) use the centred stencil from the windows
    const int ki0 = k0 > 2 ? k0 : 2, ki1 = k1 < nz - 2 ? k1 : nz - 2;
    double yw[2][5], xw[2][5];
    if (ki0 < ki1) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long q = col + (long)(ki0 - 2 + j) * nxy;
            const TW2 y2 = ld2(by, q), x2 = ld2(bx, q);
            yw[0][j + 1] = (double)y2.x; yw[1][j + 1] = (double)y2.y;
            xw[0][j + 1] = (double)x2.x; xw[1][j + 1] = (double)x2.y;
        }
    }
    for (int k = k0; k < k1; ++k) {
        const long q = col + (long)k * nxy;
        double wk1[2], wk2[2];
        if (k >= ki0 && k < ki1) {
            const TW2 y2 = ld2(by, q + 2L * nxy), x2 = ld2(bx, q + 2L * nxy);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    yw[c][j] = yw[c][j + 1];
                    xw[c][j] = xw[c][j + 1];
                }
            }
            yw[0][4] = (double)y2.x; yw[1][4] = (double)y2.y;
            xw[0][4] = (double)x2.x; xw[1][4] = (double)x2.y;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                wk1[c] = h * (yw[c][0] - 8.0 * yw[c][1] + 8.0 * yw[c][3] - yw[c][4]);
                wk2[c] = h * (xw[c][0] - 8.0 * xw[c][1] + 8.0 * xw[c][3] - xw[c][4]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                wk1[c] = dz1<double>(k, nz, 5, h, [&](int kk) { return (double)by[col + c + (long)kk * nxy]; });
                wk2[c] = dz1<double>(k, nz, 5, h, [&](int kk) { return (double)bx[col + c + (long)kk * nxy]; });
            }
        }
        if (a.bym) {
            // mixed: bx, by hold no means; d/dz of the means (flag 5) is added back
            const double m1 = dz1<double>(k, nz, 5, h, [&](int kk) { return a.bym[kk]; });
            const double m2 = dz1<double>(k, nz, 5, h, [&](int kk) { return a.bxm[kk]; });
            wk1[0] += m1; wk1[1] += m1;
            wk2[0] += m2; wk2[1] += m2;
        }
        const TW2 ex = ld2(a.dybzh, q), ey = ld2(a.dxbzh, q), ez = ld2(a.jzh, q);
        const double exv[2] = {(double)ex.x, (double)ex.y}, eyv[2] = {(double)ey.x, (double)ey.y},
                     ezv[2] = {(double)ez.x, (double)ez.y};
        TW2 js, jx, jy, jz;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const TW v = (TW)(ezv[c] * ezv[c] + ((exv[c] - wk1[c]) * (exv[c] - wk1[c])) +
                              ((wk2[c] - eyv[c]) * (wk2[c] - eyv[c])));
            const TW vx = (TW)(exv[c] - wk1[c]), vy = (TW)(wk2[c] - eyv[c]), vz = (TW)ezv[c];
            if (c == 0) { js.x = v; jx.x = vx; jy.x = vy; jz.x = vz; }
            else { js.y = v; jx.y = vx; jy.y = vy; jz.y = vz; }
        }
        *(TW2*)((TW*)a.jsq + q) = js;
        if (a.jx) {
            *(TW2*)((TW*)a.jx + q) = jx;
            *(TW2*)((TW*)a.jy + q) = jy;
            *(TW2*)((TW*)a.jz + q) = jz;
        }
    }
}

// mag_init spectral part, src/magnetic.f90:63-233 (truncation folded into reads)
template <typename TW>
struct MagInitSpecKernel {
    using Args = MagArgs;
    static constexpr int NT = 256;
    static __device__ __forceinline__ void body(const MagArgs& a, uint3 bid, uint3 gdim);
};

template <typename TW>
__device__ __forceinline__ void MagInitSpecKernel<TW>::body(const MagArgs& a, uint3 bid, uint3 gdim)
{
    (void)bid;
    (void)gdim;
    const long q = xcd_tile(bid.x, gdim.x) * NT + threadIdx.x;
    if (q >= a.g.n3) return;
    const int nx = a.g.nx, ny = a.g.ny, nxy = a.g.nxy;
    const int k = (int)(q / nxy);
    const long base = (long)k * nxy;
    const int rem = (int)(q - base);
    const int jq = rem / nx, ip = rem - jq * nx;
    const int icx = a.icutx, icy = a.icuty;
    const TW* ex = (const TW*)a.exh;
    const TW* ey = (const TW*)a.eyh;
    const TW* ez = (const TW*)a.ezh;
    auto T = [&](const TW* f, int p2, int q2) -> double {
        return in_band(p2, q2, nx, ny, icx, icy) ? 0.0 : (double)f[base + (long)q2 * nx + p2];
    };
    const double k2 = a.akx2[ip] + a.aky2[jq];
    const bool dc = (ip == 0 && jq == 0);
    // pol = -Bz-hat / (akx2+aky2), 123-136
    const double pol = dc ? 0.0 : -1.0 * T(ez, ip, jq) / k2;
    // wk3 = dBy/dx - dBx/dy (spectral), 141-175
    double dxby = 0.0, dybx = 0.0;
    if (!(ip == 0 || ip == nx / 2)) {
        const double v = T(ey, nx - ip, jq) * a.akx1[ip];
        dxby = (ip < nx / 2) ? -v : v;
    }
    if (!(jq == 0 || jq == ny / 2)) {
        const double v = T(ex, ip, ny - jq) * a.aky1[jq];
        dybx = (jq < ny / 2) ? -v : v;
    }
    const double jz = dxby - dybx;
    const double tor = dc ? 0.0 : -1.0 * jz / k2;
    a.pol[q] = pol;
    a.tor[q] = tor;
    // 199-233: wk3 = -tor*(akx2+aky2), dBz/dy, dBz/dx from the truncated Bz-hat
    ((TW*)a.jzh)[q] = (TW)(-1.0 * tor * k2);
    double dyb = 0.0, dxb = 0.0;
    if (!(jq == 0 || jq == ny / 2)) {
        const double v = T(ez, ip, ny - jq) * a.aky1[jq];
        dyb = (jq < ny / 2) ? -v : v;
    }
    if (!(ip == 0 || ip == nx / 2)) {
        const double v = T(ez, nx - ip, jq) * a.akx1[ip];
        dxb = (ip < nx / 2) ? -v : v;
    }
    ((TW*)a.dybzh)[q] = (TW)dyb;
    ((TW*)a.dxbzh)[q] = (TW)dxb;
    // 111-117: mean fields from the DC of each plane
    if (ip == 0 && jq == 0) {
        a.bxbar[k] = (double)ex[base] * a.p.xn;
        a.bybar[k] = (double)ey[base] * a.p.xn;
    }
}

__global__ void bzbar_kernel(MagArgs a, int wf32, double* out)
{
    const double v = wf32 ? (double)((const float*)a.ezh)[0] : ((const double*)a.ezh)[0];
    out[0] = v * a.p.xn;
}

#define MHD_LAUNCH_W(K, grid, block, shm, s, arg)                                          \
    do {                                                                                   \
        const int e_ = (arg).wf32 ? mhd_launch<K<float>>(grid, block, shm, s, arg)         \
                                  : mhd_launch<K<double>>(grid, block, shm, s, arg);       \
        if (e_) return -22;                                                                \
    } while (0)

int launch_mag_spec1(const MagArgs& a, hipStream_t s)
{
    const long quads = (long)(a.g.nx / 2 + 1) * (a.g.ny / 2 + 1) * a.g.nz;
    MHD_LAUNCH_W(MagSpec1Kernel, dim3(nblk(quads)), dim3(NT), 0, s, a);
    return 0;
}
int launch_mean_emf(const MagArgs& a, hipStream_t s)
{
    MHD_LAUNCH_W(MeanEmfKernel, dim3(1), dim3(256), 4 * a.g.nz * sizeof(double), s, a);
    return 0;
}
int launch_mag_rhs(const MagArgs& a, hipStream_t s)
{
    if (a.zc <= 0) return -10;
    if (a.g.nxy % 2) return -10;   // two columns per thread (nx is even)
    MHD_LAUNCH_W(MagRhsKernel, dim3(nblk(a.g.nxy / 2), (unsigned)((a.g.nz + a.zc - 1) / a.zc)), dim3(NT), 0, s, a);
    return 0;
}
int launch_bandj_spec(const MagArgs& a, hipStream_t s)
{
    const long quads = (long)(a.g.nx / 2 + 1) * (a.g.ny / 2 + 1) * a.g.nz;
    MHD_LAUNCH_W(BandjSpecKernel, dim3(nblk(quads)), dim3(NT), 0, s, a);
    return 0;
}
int launch_jsq(const MagArgs& a, hipStream_t s)
{
    if (a.zc <= 0) return -10;
    if (a.g.nxy % 2) return -10;   // two columns per thread (nx is even)
    MHD_LAUNCH_W(JsqKernel, dim3(nblk(a.g.nxy / 2), (unsigned)((a.g.nz + a.zc - 1) / a.zc)), dim3(NT), 0, s, a);
    return 0;
}
int launch_mag_init_spec(const MagArgs& a, hipStream_t s, double* means_out)
{
    MHD_LAUNCH_W(MagInitSpecKernel, dim3(nblk(a.g.n3)), dim3(NT), 0, s, a);
    hipLaunchKernelGGL(bzbar_kernel, dim3(1), dim3(1), 0, s, a, a.wf32, means_out);
    return hipGetLastError() == hipSuccess ? 0 : -22;
}

// ---------------------------------------------------------------------------
// AB3 update, src/main.F90:413-477, plus magdealias (src/magnetic.f90:666-717)
// ---------------------------------------------------------------------------
template <typename TH>
struct AB3Kernel {
    using Args = AB3Args;
    static constexpr int NT = 256;
    static __device__ __forceinline__ void body(const AB3Args& a, uint3 bid, uint3 gdim);
};

template <typename TH>
__device__ __forceinline__ void AB3Kernel<TH>::body(const AB3Args& a, uint3 bid, uint3 gdim)
{
    (void)bid;
    (void)gdim;
    const long q = (long)bid.x * NT + threadIdx.x;
    if (q < a.g.n3) {
        const double c0 = a.c0, c1 = a.c1, c2 = a.c2;
        // every entry defined; field f is updated when f0 <= f < nfields
        double xn[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int f = 0; f < 7; ++f) {
            if (f >= a.f0 && f < a.nfields) {
                const double* x = a.var[f];
                xn[f] = x[q] + c0 * (double)((const TH*)a.hist[f][2])[q] +
                        c1 * (double)((const TH*)a.hist[f][1])[q] + c2 * (double)((const TH*)a.hist[f][0])[q];
            }
        }
        // src/main.F90:417-425: density floor, its history zeroed
        const bool hydro = a.f0 == 0;
        if (hydro && xn[3] < 0.05) {
            xn[3] = 0.05;
            ((TH*)a.hist[3][0])[q] = (TH)0;
            ((TH*)a.hist[3][1])[q] = (TH)0;
            ((TH*)a.hist[3][2])[q] = (TH)0;
        }
        // magdealias: pol and tor vanish on the 2/3 band
        bool band = false;
        if (a.mag) {
            const int nx = a.g.nx, ny = a.g.ny;
            const int rem = (int)(q % a.g.nxy);
            const int jq = rem / nx, ip = rem - jq * nx;
            band = in_band(ip, jq, nx, ny, a.icutx, a.icuty);
        }
#pragma unroll
        for (int f = 0; f < 7; ++f) {
            if (f >= a.f0 && f < a.nfields) a.var[f][q] = (f >= 5 && band) ? 0.0 : xn[f];
        }
        if (a.shadow[0] && hydro) {
            const int k = (int)(q / a.g.nxy);
            ((float*)a.shadow[0])[q] = (float)xn[0];
            ((float*)a.shadow[1])[q] = (float)xn[1];
            ((float*)a.shadow[2])[q] = (float)xn[2];
            ((float*)a.shadow[3])[q] = (float)(xn[3] - a.bg_r[k]);
            ((float*)a.shadow[4])[q] = (float)(xn[4] - a.bg_t[k]);
        }
    }
    if (q < a.g.nz && a.mag) {
        for (int f = 0; f < 2; ++f) {
            double* x = a.mvar[f];
            x[q] = x[q] + a.c0 * a.mhist[f][2][q] + a.c1 * a.mhist[f][1][q] + a.c2 * a.mhist[f][0][q];
        }
    }
}

int launch_ab3(const AB3Args& a, hipStream_t s)
{
    // with no field to update (f0 = 7) only the nz means are left
    long n = a.f0 < a.nfields ? (a.g.n3 > a.g.nz ? a.g.n3 : a.g.nz) : a.g.nz;
    return a.h32 ? mhd_launch<AB3Kernel<float>>(dim3(nblk(n)), dim3(NT), 0, s, a)
                 : mhd_launch<AB3Kernel<double>>(dim3(nblk(n)), dim3(NT), 0, s, a);
}

// ---------------------------------------------------------------------------
// boundary conditions: impermeable stress-free walls, src/dbydz.F90:3814-3900
// (u, v, w); fixed T (4119, 4183) or the radiative top / fixed-flux bottom
// options; vertical field (4660-4750) or perfect conductors (4314-4400) for
// pol, tor and the means.  elf = 1/11.
// ---------------------------------------------------------------------------
struct BoundaryKernel {
    using Args = BCArgs;
    static constexpr int NT = 256;
    static __device__ __forceinline__ void body(const BCArgs& a, uint3 bid, uint3 gdim);
};

__device__ __forceinline__ void BoundaryKernel::body(const BCArgs& a, uint3 bid, uint3 gdim)
{
    (void)bid;
    (void)gdim;
    const long p = (long)bid.x * NT + threadIdx.x;
    const long m = a.g.nxy;
    const double elf = 1.0 / 11.0;
    if (p < m && a.hydro) {
        const long L = (long)(a.g.nz - 1) * m;
        a.w[p] = 0.0;
        a.u[p] = elf * (18.0 * a.u[m + p] - 9.0 * a.u[2 * m + p] + 2.0 * a.u[3 * m + p]);
        a.v[p] = elf * (18.0 * a.v[m + p] - 9.0 * a.v[2 * m + p] + 2.0 * a.v[3 * m + p]);
        if (a.radbc) {
            // 4037-4057: one-sided dT/dz = theta T^4, ten Newton steps from the old value
            const double tdz = a.theta * a.dz;
            const double rest = -18.0 * a.t[m + p] + 9.0 * a.t[2 * m + p] - 2.0 * a.t[3 * m + p];
            double t1 = a.t[p];
            for (int it = 0; it < 10; ++it) {
                const double ft1 = 6.0 * tdz * t1 * t1 * t1 * t1 + 11.0 * t1 + rest;
                const double dft1 = 24.0 * tdz * t1 * t1 * t1 + 11.0;
                t1 = t1 - (ft1 / dft1);
            }
            a.t[p] = t1;
        } else {
            a.t[p] = 1.0;                                  // 4119
        }
        a.w[L + p] = 0.0;
        a.u[L + p] = elf * (18.0 * a.u[L - m + p] - 9.0 * a.u[L - 2 * m + p] + 2.0 * a.u[L - 3 * m + p]);
        a.v[L + p] = elf * (18.0 * a.v[L - m + p] - 9.0 * a.v[L - 2 * m + p] + 2.0 * a.v[L - 3 * m + p]);
        if (a.constflux)                                   // 4133-4143: dT/dz = theta
            a.t[L + p] = elf * (18.0 * a.t[L - m + p] - 9.0 * a.t[L - 2 * m + p] + 2.0 * a.t[L - 3 * m + p] +
                                6.0 * a.theta * a.dz);
        else
            a.t[L + p] = 1.0 + a.theta;                    // 4183
        if (a.su) {
            const long rows[2] = {p, L + p};
            const int ks[2] = {0, a.g.nz - 1};
            for (int i = 0; i < 2; ++i) {
                a.su[rows[i]] = (float)a.u[rows[i]];
                a.sv[rows[i]] = (float)a.v[rows[i]];
                a.sw[rows[i]] = (float)a.w[rows[i]];
                a.st[rows[i]] = (float)(a.t[rows[i]] - a.t0[ks[i]]);
            }
        }
    }
    if (p < m && a.mag) {
        const long L = (long)(a.g.nz - 1) * m;
        if (a.perfect) {
            // 4314-4400: bz = 0 (pol), zero d/dz of bx, by (tor)
            a.pol[p] = 0.0;
            a.tor[p] = elf * (18.0 * a.tor[m + p] - 9.0 * a.tor[2 * m + p] + 2.0 * a.tor[3 * m + p]);
            a.pol[L + p] = 0.0;
            a.tor[L + p] =
                elf * (18.0 * a.tor[L - m + p] - 9.0 * a.tor[L - 2 * m + p] + 2.0 * a.tor[L - 3 * m + p]);
        } else {
            // 4660-4750: vertical field
            a.tor[p] = 0.0;
            a.pol[p] = elf * (18.0 * a.pol[m + p] - 9.0 * a.pol[2 * m + p] + 2.0 * a.pol[3 * m + p]);
            a.tor[L + p] = 0.0;
            a.pol[L + p] =
                elf * (18.0 * a.pol[L - m + p] - 9.0 * a.pol[L - 2 * m + p] + 2.0 * a.pol[L - 3 * m + p]);
        }
    }
    if (p == 0 && a.mag) {
        const int nz = a.g.nz;
        if (a.perfect) {
            a.bxbar[0] = elf * (18.0 * a.bxbar[1] - 9.0 * a.bxbar[2] + 2.0 * a.bxbar[3]);
            a.bybar[0] = elf * (18.0 * a.bybar[1] - 9.0 * a.bybar[2] + 2.0 * a.bybar[3]);
            a.bxbar[nz - 1] = elf * (18.0 * a.bxbar[nz - 2] - 9.0 * a.bxbar[nz - 3] + 2.0 * a.bxbar[nz - 4]);
            a.bybar[nz - 1] = elf * (18.0 * a.bybar[nz - 2] - 9.0 * a.bybar[nz - 3] + 2.0 * a.bybar[nz - 4]);
        } else {
            a.bxbar[0] = 0.0;
            a.bybar[0] = 0.0;
            a.bxbar[nz - 1] = 0.0;
            a.bybar[nz - 1] = 0.0;
        }
    }
}

int launch_boundary(const BCArgs& a, hipStream_t s)
{
    return mhd_launch<BoundaryKernel>(dim3(nblk(a.g.nxy)), dim3(NT), 0, s, a);
}

__global__ void __launch_bounds__(NT) magdealias_kernel(Geom g, double* pol, double* tor, int icx, int icy)
{
    const long q = (long)blockIdx.x * NT + threadIdx.x;
    if (q >= g.n3) return;
    const int rem = (int)(q % g.nxy);
    const int jq = rem / g.nx, ip = rem - jq * g.nx;
    if (in_band(ip, jq, g.nx, g.ny, icx, icy)) {
        pol[q] = 0.0;
        tor[q] = 0.0;
    }
}

int launch_magdealias(Geom g, double* pol, double* tor, int icutx, int icuty, hipStream_t s)
{
    hipLaunchKernelGGL(magdealias_kernel, dim3(nblk(g.n3)), dim3(NT), 0, s, g, pol, tor, icutx, icuty);
    return hipGetLastError() == hipSuccess ? 0 : -22;
}

// ---------------------------------------------------------------------------
// reductions: one workgroup per plane writes partial[k][0..7]; one finishing
// workgroup combines the planes in k order like the reference's loops
// ---------------------------------------------------------------------------
// Workgroup reduction of NV values (kind: 0 sum, 1 min, 2 max).  Within each
// 64-wide wavefront a butterfly of cross-lane moves (__shfl_xor, no LDS)
// leaves every lane with the wave's result; the NT / 64 wave results then
// pass through LDS and thread 0 combines them in wave order.  The order of
// combination is fixed, so results do not depend on timing.  The result is
// valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], const int (&kind)[NV], double* sh)
{
    static_assert(NT % 64 == 0, "whole wavefronts");
    constexpr int NW = NT / 64;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const double o = __shfl_xor(v[i], off, 64);
            v[i] = kind[i] == 0 ? v[i] + o : kind[i] == 1 ? fmin(v[i], o) : fmax(v[i], o);
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) sh[i * NW + wave] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double x = sh[i * NW];
            for (int w = 1; w < NW; ++w) {
                const double b = sh[i * NW + w];
                x = kind[i] == 0 ? x + b : kind[i] == 1 ? fmin(x, b) : fmax(x, b);
            }
            v[i] = x;
        }
    }
}

template <typename TW>
__device__ __forceinline__ double rdw(const void* p, long q) { return (double)((const TW*)p)[q]; }

template <typename TW>
__global__ void __launch_bounds__(NT) reduce_plane_kernel(RedArgs a)
{
    __shared__ double sh[5 * (NT / 64)];
    // workgroup = (plane k, chunk ch): the points [lo, hi) of the plane
    const int k = blockIdx.x / RED_CHUNKS, ch = blockIdx.x - k * RED_CHUNKS;
    const long base = (long)k * a.g.nxy;
    const int m = a.g.nxy;
    const int lo = (int)((long)ch * m / RED_CHUNKS), hi = (int)((long)(ch + 1) * m / RED_CHUNKS);
    const int tid = threadIdx.x;
    double* part = a.partial + (long)blockIdx.x * 8;
    if (a.op == RED_GLOBALAVG) {
        // src/utils.F90:55-82: rav, urav, vrav, keav, bsqav
        double v[5] = {0, 0, 0, 0, 0};
        for (int i = lo + tid; i < hi; i += NT) {
            const long q = base + i;
            const double uu = a.u[q], vv = a.v[q], rr = a.r[q], ww = a.w[q];
            double bx = rdw<TW>(a.bx, q), by = rdw<TW>(a.by, q), bz = rdw<TW>(a.bz, q);
            if (a.bxm) {
                bx += a.bxm[k];
                by += a.bym[k];
                bz += a.bzm;
            }
            v[0] += rr;
            v[1] += rr * uu;
            v[2] += rr * vv;
            v[3] += 0.5 * rr * (uu * uu + vv * vv + ww * ww);
            v[4] += bx * bx + by * by + bz * bz;
        }
        const int kind[5] = {0, 0, 0, 0, 0};
        block_reduce<5>(v, kind, sh);
        if (tid == 0)
            for (int i = 0; i < 5; ++i) part[i] = v[i];
    } else if (a.op == RED_MINMAX || a.op == RED_DIVB) {
        double v[2] = {1e300, -1e300};
        for (int i = lo + tid; i < hi; i += NT) {
            const long q = base + i;
            double x;
            if (a.op == RED_MINMAX)
                x = (a.f_f32 ? (double)((const float*)a.f0)[q] : ((const double*)a.f0)[q]) +
                    (a.f0m ? a.f0m[k] : 0.0) + a.f0s;
            else
                x = (a.f_f32 ? (double)((const float*)a.f0)[q] : ((const double*)a.f0)[q]) +
                    (a.f_f32 ? (double)((const float*)a.f1)[q] : ((const double*)a.f1)[q]) +
                    (a.f_f32 ? (double)((const float*)a.f2)[q] : ((const double*)a.f2)[q]);
            v[0] = fmin(v[0], x);
            v[1] = fmax(v[1], x);
        }
        const int kind[2] = {1, 2};
        block_reduce<2>(v, kind, sh);
        if (tid == 0) {
            part[0] = v[0];
            part[1] = v[1];
        }
    } else if (a.op == RED_CALCDT) {
        // src/utils.F90:168-186
        double v[3] = {1e300, -1.0e-50, -1.0e-50};
        const double f = a.p.f, g1 = a.p.g1;
        for (int i = lo + tid; i < hi; i += NT) {
            const long q = base + i;
            const double u0 = a.u[q], v0 = a.v[q], w0 = a.w[q], r0 = a.r[q], t0 = a.t[q];
            double bz0 = rdw<TW>(a.bz, q), bx0 = rdw<TW>(a.bx, q), by0 = rdw<TW>(a.by, q);
            if (a.bxm) {
                bx0 += a.bxm[k];
                by0 += a.bym[k];
                bz0 += a.bzm;
            }
            const double bsq0 = bx0 * bx0 + by0 * by0 + bz0 * bz0;
            const double xx = sqrt(u0 * u0 + v0 * v0) + g1 * sqrt(t0) + sqrt((bsq0 * f) / r0);
            const double xz = sqrt(w0 * w0) + g1 * sqrt(t0) + sqrt((bsq0 * f) / r0);
            v[0] = fmin(v[0], r0);
            v[1] = fmax(v[1], xx);
            v[2] = fmax(v[2], xz);
        }
        const int kind[3] = {1, 2, 2};
        block_reduce<3>(v, kind, sh);
        if (tid == 0)
            for (int i = 0; i < 3; ++i) part[i] = v[i];
    } else if (a.op == RED_NUSSELT) {
        // src/dbydz.F90:4878-4886, bottom plane only (RED_CHUNKS workgroups)
        double v[1] = {0.0};
        const long L = (long)(a.g.nz - 1) * m;
        const double c = 1.0 / (6.0 * a.p.dz);
        for (int i = lo + tid; i < hi; i += NT)
            v[0] += c * (11.0 * a.t[L + i] - 18.0 * a.t[L - m + i] + 9.0 * a.t[L - 2 * m + i] - 2.0 * a.t[L - 3 * m + i]);
        const int kind[1] = {0};
        block_reduce<1>(v, kind, sh);
        if (tid == 0) part[0] = v[0];
    }
}

__global__ void __launch_bounds__(NT) reduce_finish_kernel(RedArgs a)
{
    const int nz = a.g.nz;
    double* R = a.result;
    // per plane: its chunks combined in order (sums for the averages and
    // Nusselt, min / max for the others), one (plane, value) per thread, into
    // P[k * 8 + i]: in LDS when the planes fit, else past the chunk partials
    constexpr int LDS_VALS = 4096;
    __shared__ double sP[LDS_VALS];
    const int nk = a.op == RED_NUSSELT ? 1 : nz;
    double* P = nk * 8 <= LDS_VALS ? sP : a.partial + (long)nz * RED_CHUNKS * 8;
    const int nv = a.op == RED_GLOBALAVG ? 5 : (a.op == RED_CALCDT ? 3 : (a.op == RED_NUSSELT ? 1 : 2));
    const bool sum = a.op == RED_GLOBALAVG || a.op == RED_NUSSELT;
    for (int idx = threadIdx.x; idx < nk * nv; idx += blockDim.x) {
        const int k = idx / nv, i = idx - k * nv;
        const double* c = a.partial + (long)k * RED_CHUNKS * 8 + i;
        double y[RED_CHUNKS];
#pragma unroll
        for (int ch = 0; ch < RED_CHUNKS; ++ch) y[ch] = c[ch * 8];
        double x = y[0];
#pragma unroll
        for (int ch = 1; ch < RED_CHUNKS; ++ch) x = sum ? x + y[ch] : (i == 0 ? fmin(x, y[ch]) : fmax(x, y[ch]));
        P[k * 8 + i] = x;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (a.op == RED_GLOBALAVG) {
        // src/utils.F90:83-127
        double urbar = 0, vrbar = 0, rbar = 0, me = 0, ke = 0;
        for (int k = 0; k < nz; ++k) {
            const double* v = P + k * 8;
            if (k == 0 || k == nz - 1) {
                urbar += 0.5 * v[1];
                vrbar += 0.5 * v[2];
                rbar += 0.5 * v[0];
                me += 0.25 * v[4];
                ke += 0.5 * v[3];
            } else {
                urbar += v[1];
                vrbar += v[2];
                rbar += v[0];
                me += 0.5 * v[4];
                ke += v[3];
            }
        }
        R[0] = urbar / rbar;
        R[1] = vrbar / rbar;
        R[2] = ke * a.p.xzn;
        R[3] = me * a.p.xzn;
    } else if (a.op == RED_MINMAX) {
        double mn = P[0], mx = P[1];
        for (int k = 1; k < nz; ++k) {
            mn = fmin(mn, P[k * 8]);
            mx = fmax(mx, P[k * 8 + 1]);
        }
        R[0] = mn;
        R[1] = mx;
    } else if (a.op == RED_DIVB) {
        // src/utils.F90:270-288: divbmax starts at -1e30, divbmin at 1.0
        double mx = -1.0e30, mn = 1.0;
        for (int k = 0; k < nz; ++k) {
            mx = fmax(mx, P[k * 8 + 1]);
            mn = fmin(mn, P[k * 8]);
        }
        R[0] = mx;
        R[1] = mn;
    } else if (a.op == RED_CALCDT) {
        // src/utils.F90:205-242
        double denmn = P[0], vx = P[1], vz = P[2];
        for (int k = 1; k < nz; ++k) {
            denmn = fmin(denmn, P[k * 8]);
            vx = fmax(vx, P[k * 8 + 1]);
            vz = fmax(vz, P[k * 8 + 2]);
        }
        const double dtcflx = 0.16 * a.p.dd / vx;
        const double dtcflz = 0.3 * a.p.dz / vz;
        const double dtviscx = a.p.dtvisc0x * denmn, dtviscz = a.p.dtvisc0z * denmn;
        const double dtdiffx = a.p.dtdiff0x * denmn, dtdiffz = a.p.dtdiff0z * denmn;
        const double mn = fmin(fmin(fmin(dtcflx, dtcflz), fmin(dtviscx, dtviscz)), fmin(dtdiffx, dtdiffz));
        R[0] = a.p.sf * mn;
        R[1] = denmn;
        R[2] = vx;
        R[3] = vz;
    } else if (a.op == RED_NUSSELT) {
        const double tzbar = P[0] * a.p.xn;
        R[0] = (tzbar - a.p.nusstmp) / (a.p.theta - a.p.nusstmp);
    }
}

int launch_reduce(const RedArgs& a, hipStream_t s)
{
    const unsigned g = (a.op == RED_NUSSELT) ? (unsigned)RED_CHUNKS : (unsigned)(a.g.nz * RED_CHUNKS);
    if (a.wf32) hipLaunchKernelGGL(reduce_plane_kernel<float>, dim3(g), dim3(NT), 0, s, a);
    else hipLaunchKernelGGL(reduce_plane_kernel<double>, dim3(g), dim3(NT), 0, s, a);
    hipLaunchKernelGGL(reduce_finish_kernel, dim3(1), dim3(NT), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -22;
}

__global__ void com_shift_kernel(Geom g, double* u, double* v, const double* res)
{
    const long q = (long)blockIdx.x * NT + threadIdx.x;
    if (q >= g.n3) return;
    u[q] = u[q] - res[0];
    v[q] = v[q] - res[1];
}

int launch_com_shift(Geom g, double* u, double* v, const double* res, hipStream_t s)
{
    hipLaunchKernelGGL(com_shift_kernel, dim3(nblk(g.n3)), dim3(NT), 0, s, g, u, v, res);
    return hipGetLastError() == hipSuccess ? 0 : -22;
}

// ---------------------------------------------------------------------------
// mixed precision: FP32 fluctuation shadows of the FP64 state
// ---------------------------------------------------------------------------
// FP32 fluctuation shadows of the FP64 state (mixed precision)
struct ShadowArgs {
    Geom g;
    const double *u, *v, *w, *t, *r, *t0, *r0;
    float *su, *sv, *sw, *st, *sr;
};
struct ShadowKernel {
    using Args = ShadowArgs;
    static constexpr int NT = 256;
    static __device__ __forceinline__ void body(const ShadowArgs& a, uint3 bid, uint3 gdim)
    {
        (void)gdim;
        const long q = (long)bid.x * NT + threadIdx.x;
        if (q >= a.g.n3) return;
        const int k = (int)(q / a.g.nxy);
        a.su[q] = (float)a.u[q];
        a.sv[q] = (float)a.v[q];
        a.sw[q] = (float)a.w[q];
        a.st[q] = (float)(a.t[q] - a.t0[k]);
        a.sr[q] = (float)(a.r[q] - a.r0[k]);
    }
};

int launch_shadow(Geom g, const double* u, const double* v, const double* w, const double* t, const double* r,
                  float* su, float* sv, float* sw, float* st, float* sr, const double* t0, const double* r0,
                  hipStream_t s)
{
    ShadowArgs a;
    a.g = g;
    a.u = u; a.v = v; a.w = w; a.t = t; a.r = r; a.t0 = t0; a.r0 = r0;
    a.su = su; a.sv = sv; a.sw = sw; a.st = st; a.sr = sr;
    return mhd_launch<ShadowKernel>(dim3(nblk(g.n3)), dim3(NT), 0, s, a);
}

__global__ void __launch_bounds__(NT) add_mean_kernel(const float* src, double* dst, const double* mean, double sc,
                                                      Geom g)
{
    const long q = (long)blockIdx.x * NT + threadIdx.x;
    if (q >= g.n3) return;
    const int k = (int)(q / g.nxy);
    dst[q] = (double)src[q] + (mean ? mean[k] : 0.0) + sc;
}

int launch_add_mean(const float* src, double* dst, const double* mean, double sc, Geom g, hipStream_t s)
{
    hipLaunchKernelGGL(add_mean_kernel, dim3(nblk(g.n3)), dim3(NT), 0, s, src, dst, mean, sc, g);
    return hipGetLastError() == hipSuccess ? 0 : -22;
}

// ---------------------------------------------------------------------------
__global__ void copy_cvt_kernel(const void* src, int sf, void* dst, int df, long n)
{
    const long q = (long)blockIdx.x * NT + threadIdx.x;
    if (q >= n) return;
    const double v = sf ? (double)((const float*)src)[q] : ((const double*)src)[q];
    if (df) ((float*)dst)[q] = (float)v;
    else ((double*)dst)[q] = v;
}

int launch_copy_cvt(const void* src, int src_f32, void* dst, int dst_f32, long n, hipStream_t s)
{
    hipLaunchKernelGGL(copy_cvt_kernel, dim3(nblk(n)), dim3(NT), 0, s, src, src_f32, dst, dst_f32, n);
    return hipGetLastError() == hipSuccess ? 0 : -22;
}

__global__ void fill_kernel(void* dst, int f32, double v, long n)
{
    const long q = (long)blockIdx.x * NT + threadIdx.x;
    if (q >= n) return;
    if (f32) ((float*)dst)[q] = (float)v;
    else ((double*)dst)[q] = v;
}

int launch_fill(void* dst, int f32, double v, long n, hipStream_t s)
{
    hipLaunchKernelGGL(fill_kernel, dim3(nblk(n)), dim3(NT), 0, s, dst, f32, v, n);
    return hipGetLastError() == hipSuccess ? 0 : -22;
}

}  // namespace mhd
