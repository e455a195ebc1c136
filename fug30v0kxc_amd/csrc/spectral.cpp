// spectral.cpp -- dispatch of the horizontal transform kernels by length.
#include <type_traits>

#include "fft_reg.hpp"
#include "spectral.hpp"

namespace mhd {

int launch_spectral_g0(int n, const SpecLaunch& L, int f32, hipStream_t s);
int launch_spectral_g1(int n, const SpecLaunch& L, int f32, hipStream_t s);
int launch_spectral_g2(int n, const SpecLaunch& L, int f32, hipStream_t s);
int launch_spectral_g3(int n, const SpecLaunch& L, int f32, hipStream_t s);
int launch_spectral_g4(int n, const SpecLaunch& L, int f32, hipStream_t s);
int launch_spectral_g5(int n, const SpecLaunch& L, int f32, hipStream_t s);

#define MHD_ALL_SIZES(X) X(8) X(12) X(16) X(20) X(24) X(30) X(32) X(40) X(48) X(60) X(64) X(80) X(96) X(120) X(128) X(160) X(192) X(240) X(256) X(320) X(384) X(480) X(512) X(640) X(768) X(1024)

bool spectral_len_fixed(int n)
{
    switch (n) {
#define X(n_) case n_: return true;
        MHD_ALL_SIZES(X)
#undef X
    }
    return false;
}

bool spectral_len_supported(int n)
{
    if (spectral_len_fixed(n)) return true;
    int r[SPEC_GEN_MAX_PASSES], np = 0;
    return generic_plan(n, r, &np);
}

int spectral_variants(int n)
{
    switch (n) {
#define X(n_)                                                                  \
    case n_:                                                                   \
        return std::is_void<AltPlan<n_>::type>::value ? 1 : std::is_void<AltPlan2<n_>::type>::value ? 2 : 3;
        MHD_ALL_SIZES(X)
#undef X
    }
    return 1;
}

static int launch_group(const SpecLaunch& L, int compute_f32, hipStream_t s)
{
    const int n = (L.dir == 0) ? L.g.nx : L.g.ny;
    if (!spectral_len_fixed(n)) return launch_spectral_generic(L, compute_f32, s);
    if (n <= 60) return launch_spectral_g0(n, L, compute_f32, s);
    if (n <= 128) return launch_spectral_g1(n, L, compute_f32, s);
    if (n <= 256) return launch_spectral_g2(n, L, compute_f32, s);
    if (n <= 480) return launch_spectral_g3(n, L, compute_f32, s);
    if (n <= 640) return launch_spectral_g4(n, L, compute_f32, s);
    return launch_spectral_g5(n, L, compute_f32, s);
}

// Three LD_CROSS jobs whose factors are the components of two vectors a, b
// with job q = a_{q+1} b_{q+2} - a_{q+2} b_{q+1}: the EMF u x B.
static bool is_cross_product(const SpecLaunch& L)
{
    if (L.njobs != 3) return false;
    const SpecJob* J = L.job;
    for (int q = 0; q < 3; ++q)
        if (J[q].load_op != LD_CROSS || J[q].in_f32 != J[0].in_f32) return false;
    // a_{q+1} = J[q].in[0], b_{q+2} = J[q].in[1]
    const void *a[3], *b[3];
    for (int q = 0; q < 3; ++q) {
        a[(q + 1) % 3] = J[q].in[0];
        b[(q + 2) % 3] = J[q].in[1];
    }
    for (int q = 0; q < 3; ++q)
        if (J[q].in[2] != a[(q + 2) % 3] || J[q].in[3] != b[(q + 1) % 3]) return false;
    return true;
}

// Among first-derivative jobs, three that differentiate the Maxwell-stress
// loads of one direction: MAXW(b0, b1, b2), PROD(b1, b0) or PROD(b0, b1), and
// PROD(b2, b0), all of one element type.  Fills G with those three (in that
// order, job 0 the first product) and R with the others; false if absent.
static bool split_bstress(const SpecLaunch& L, SpecLaunch& G, SpecLaunch& R)
{
    if (L.mode != SM_DERIV1 || L.njobs < 3) return false;
    for (int m = 0; m < L.njobs; ++m) {
        const SpecJob& M = L.job[m];
        if (M.load_op != LD_MAXW) continue;
        const void *b0 = M.in[0], *b1 = M.in[1], *b2 = M.in[2];
        int p1 = -1, p2 = -1, pat = 0;
        for (int q = 0; q < L.njobs; ++q) {
            const SpecJob& J = L.job[q];
            if (q == m || J.load_op != LD_PROD || J.in_f32 != M.in_f32) continue;
            if (p1 < 0 && J.in[0] == b1 && J.in[1] == b0) { p1 = q; pat = 0; }
            else if (p1 < 0 && J.in[0] == b0 && J.in[1] == b1) { p1 = q; pat = 1; }
            else if (p2 < 0 && J.in[0] == b2 && J.in[1] == b0) p2 = q;
        }
        if (p1 < 0 || p2 < 0) continue;
        G = L;
        G.njobs = 3;
        G.job[0] = L.job[p1];
        G.job[1] = L.job[p2];
        G.job[2] = M;
        G.bstress = 1;
        G.bpat = pat;
        R = L;
        R.njobs = 0;
        for (int q = 0; q < L.njobs; ++q)
            if (q != m && q != p1 && q != p2) R.job[R.njobs++] = L.job[q];
        return true;
    }
    return false;
}

// The load operation is a compile-time parameter of the kernels, so jobs are
// launched in groups of equal load operation -- except, on the
// register-blocked lengths: the three first derivatives of one direction's
// Maxwell-stress loads, which share one load of the field components
// (LD_BSTRESS); other first-derivative jobs of the plain, product and
// Maxwell loads, which share one launch of the run-time-dispatch kernel
// (LD_ANY: one larger launch streams at a higher rate than three small ones);
// and the three components of a cross product along x, which share their six
// factor loads (LD_CROSS3).
int launch_spectral(const SpecLaunch& L, int compute_f32, hipStream_t s)
{
    if (L.njobs <= 0) return 0;
    if (L.mode == SM_R2HC && L.dir == 0 && spectral_len_fixed(L.g.nx) && is_cross_product(L)) {
        SpecLaunch G = L;
        G.cross3 = 1;
        return launch_group(G, compute_f32, s);
    }
    // FP64 columns too, at half column groups (see launch_plan): the triple and
    // the remaining plain jobs took 274 us at 256^2 x 160 against 282 for
    // per-job loads of all six
    if (spectral_len_fixed(L.dir == 0 ? L.g.nx : L.g.ny)) {
        SpecLaunch G, R;
        if (split_bstress(L, G, R)) {
            const int e = launch_group(G, compute_f32, s);
            if (e || R.njobs == 0) return e;
            return launch_spectral(R, compute_f32, s);
        }
    }
    if (L.mode == SM_DERIV1 && L.njobs > 1 && spectral_len_fixed(L.dir == 0 ? L.g.nx : L.g.ny)) {
        bool mergeable = true, mixed_ops = false;
        for (int i = 0; i < L.njobs; ++i) {
            const int op = L.job[i].load_op;
            mergeable = mergeable && (op == LD_PLAIN || op == LD_PROD || op == LD_MAXW);
            mixed_ops = mixed_ops || op != L.job[0].load_op;
        }
        if (mergeable && mixed_ops) {
            SpecLaunch G = L;
            G.any_load = 1;
            return launch_group(G, compute_f32, s);
        }
    }
    bool done[SPEC_MAX_JOBS] = {};
    for (int i = 0; i < L.njobs; ++i) {
        if (done[i]) continue;
        SpecLaunch G = L;
        G.njobs = 0;
        for (int j = i; j < L.njobs; ++j)
            if (!done[j] && L.job[j].load_op == L.job[i].load_op) {
                G.job[G.njobs++] = L.job[j];
                done[j] = true;
            }
        const int e = launch_group(G, compute_f32, s);
        if (e) return e;
    }
    return 0;
}

}  // namespace mhd
