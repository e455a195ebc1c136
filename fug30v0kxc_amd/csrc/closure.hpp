// closure.hpp -- internal interface of the learned subgrid closure (closure.hip)
// used by the time step (mhd_api.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"

struct mhd_closure;

namespace mhd {

constexpr int CL_NF = 8;      // input fields r t u v w bx by bz
constexpr int CL_NOUT = 12;   // tau xx yy zz xy xz yz, eps x y z, q x y z

// where the features of one field come from: value = (prec ? float : double)[q]
// + (mean ? mean[k] : 0) + add
struct ClSource {
    const void* p[CL_NF];
    int f32[CL_NF];
    const double* mean[CL_NF];
    double add[CL_NF];
};

// bytes of the feature grid of a box of n cells (closure_fluxes' `grid`)
constexpr size_t closure_grid_bytes(long n) { return (size_t)n * 16; }

// features of the solver's state and the network's 12 fluxes (de-normalised)
// into out[12] (float or double arrays of the box's size) on stream s.  The
// features go to `grid` (closure_grid_bytes(n) of device memory owned by the
// caller, so runs sharing one closure never share it), or, when grid is
// null, to the closure's own buffer (calls on the closure's stream only).
int closure_fluxes(mhd_closure* cl, const ClSource& src, int nx, int ny, int nz, void* const out[CL_NOUT],
                   int out_f32, void* grid, hipStream_t s);

// appends to the closure's data set the coarse fields and subgrid targets of
// a fine state given as sources on the device (the fine run's own arrays):
// nothing crosses to the host but 20 pairs of sums for the normalisation
int closure_add_source(mhd_closure* cl, const ClSource& src, int nx, int ny, int nz, int f, double fmag);

// momentum and T right-hand sides of the closure (newest history level):
// dots[m] -= ((dx[m] + dy[m]) + dz[m]) / r for u v w (m = 0 1 2), and
// -= (dx[3] + dy[3]) + dz[3] for T; derivative arrays float when f32,
// history float when h32
int closure_tend(long n, const void* const dx[4], const void* const dy[4], const void* const dz[4], int f32,
                 const double* r, void* const dots[4], int h32, hipStream_t s);

}  // namespace mhd
