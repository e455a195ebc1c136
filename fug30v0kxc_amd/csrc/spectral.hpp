// spectral.hpp -- host-side interface of the horizontal (x/y) transform
// kernels: spectral derivatives (src/fft.F90 d1bydx/d1bydy/d2bydx/d2bydy),
// 2/3 dealiasing (dealias), and the real<->halfcomplex 2-D transform passes
// (forwardfft/inversefft) in FFTW-2 halfcomplex order.
#pragma once

#include "common.hpp"

namespace mhd {

enum SpecMode { SM_DERIV1 = 0, SM_DERIV2 = 1, SM_R2HC = 2, SM_HC2R = 3 };
// MUL_D2C: mc[0] * d2 + mc[1] * d1(d1(.)); the two differ at the Nyquist
// mode only, which d2 keeps (akx2) and d1 zeroes.
enum MulKind { MUL_D1 = 0, MUL_D2 = 1, MUL_MASK = 2, MUL_D2C = 3 };
// LD_BG: FP64 input minus a per-plane background (the mixed path's
// fluctuation loads); the others read inputs of the compute type.
// Products take optional per-plane means m_i of their factors (the mixed
// path stores B without its means): the load is the product of (in_i + m_i)
// minus its per-plane constant, which no horizontal derivative sees.
//   LD_PROD  scale*(in0+m0)*(in1+m1)
//   LD_CROSS (in0+m0)*(in1+m1) - (in2+m2)*(in3+m3)
//   LD_MAXW  scale*((in0+m0)^2 - (in1+m1)^2 - (in2+m2)^2)
enum LoadOp { LD_PLAIN = 0, LD_PROD = 1, LD_CROSS = 2, LD_BG = 3, LD_MAXW = 4,
              // kernel template value only: each job's own load_op (PLAIN, PROD
              // or MAXW) chosen at run time, so first-derivative jobs of
              // different load operations share one launch
              LD_ANY = 5,
              // kernel template value only: three LD_CROSS jobs that are the
              // components of one cross product a x b, job q loading
              // {a_{q+1}, b_{q+2}, a_{q+2}, b_{q+1}} (indices mod 3); the six
              // factor lines of a line group are loaded once for all three
              LD_CROSS3 = 6,
              // kernel template value only: first derivatives of the three
              // Maxwell-stress loads of one direction, PROD(b1, b0) or PROD(b0, b1),
              // PROD(b2, b0) and MAXW(b0, b1, b2) (SpecLaunch::bpat), from one
              // load of the three field components per line group
              LD_BSTRESS = 7 };

constexpr int SPEC_MAX_JOBS = 8;
// lengths without a register-blocked plan: run-time radix plans up to this length
constexpr int SPEC_GEN_MAX_N = 2048;
constexpr int SPEC_GEN_MAX_PASSES = 12;

// One transform job: one input field (or a fused product of fields) and up
// to two outputs.  Element types: *_f32 = 1 for float arrays, 0 for double.
struct SpecJob {
    const void* in[4];
    int in_f32;
    int load_op;      // LoadOp
    double scale;
    const double* mean[4]; // optional per-plane means of in[0..3] (nz entries)
    // LD_CROSS only: the load also subtracts jcoef[k] * jin (the eta(z) term
    // -(eta - eta_0) J of the EMF), when jin is set
    const void* jin;
    const double* jcoef;
    double mc[2];     // MUL_D2C coefficients
    const double* bg; // optional per-plane value subtracted on load (nz entries)
    void* out[2];     // SM_DERIV2: out[1] may be null (the job's first derivative only)
    int out_f32;
    int mul[2];       // MulKind for out[0], out[1] (SM_DERIV*)
    const double* addk; // optional per-plane value added on store (nz entries)
    double adds;        // scalar added on store
};

struct SpecLaunch {
    Geom g;
    int dir;          // 0: along x (rows), 1: along y (columns)
    int mode;         // SpecMode
    int njobs;
    SpecJob job[SPEC_MAX_JOBS];
    double cx;        // 2*pi/xmax (dir x) or 2*pi/ymax (dir y): mode m has wavenumber m*cx
    const void* tw;   // W_N^t table in the compute type
    int icut;         // dealias cut-off (n/2)/3
    int ncu;          // compute units: the persistent grid is occupancy x ncu workgroups
    int grid_mult;    // ... times this (>= 1; 0 = 1), a tunable launch shape
    int variant;      // radix plan: 0 the default, 1 and 2 the length's alternatives (if any)
    int max_groups;   // optional further cap on the grid; 0 = none
    int group_major;  // set by the launcher: items ordered line group first (see spec_kernel)
    int any_load;     // set by the launcher: jobs of mixed load operations in one LD_ANY launch
    int cross3;       // set by the launcher: the three jobs form one cross product (LD_CROSS3)
    int bstress;      // set by the launcher: the three jobs are Maxwell-stress loads (LD_BSTRESS)
    int bpat;         // ... job 0 is PROD(b1, b0) (0) or PROD(b0, b1) (1)
    // set by the launcher for lengths without a register-blocked plan: the
    // run-time radices and the complex lines per workgroup
    int gen_radix[SPEC_GEN_MAX_PASSES];
    int gen_npass;
    int gen_nlb;
};

// Returns 0 or a negative code (-20: unsupported length).
int launch_spectral(const SpecLaunch& L, int compute_f32, hipStream_t s);

// every even length from 2 to SPEC_GEN_MAX_N: register-blocked kernels for
// the lengths of fft_reg.hpp, run-time radix kernels for the others
bool spectral_len_supported(int n);
// the length has a register-blocked plan
bool spectral_len_fixed(int n);
// run-time radix plan of an even length (false: odd or too long)
bool generic_plan(int n, int radix[SPEC_GEN_MAX_PASSES], int* npass);
int launch_spectral_generic(const SpecLaunch& L, int compute_f32, hipStream_t s);
// radix plans of the length: SpecLaunch::variant runs from 0 to this - 1
int spectral_variants(int n);

}  // namespace mhd
