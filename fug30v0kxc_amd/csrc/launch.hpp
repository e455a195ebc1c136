// launch.hpp -- kernel launches that can be issued one box at a time or
// recorded and replayed as one launch over a table of boxes.
//
// A kernel is a struct K with
//   using Args = ...;                     its by-value argument block
//   static constexpr int NT = ...;        workgroup size (launch bound)
//   static __device__ void body(const Args&, uint3 block, uint3 grid);
// single_kernel<K> runs one box (arguments read in place from the kernarg
// segment, block and grid from the hardware); batched_kernel<K> runs many
// boxes in one grid: workgroup w belongs to the box b with start[b] <= w <
// start[b+1], takes that box's argument block from device memory and its
// own block index within the box's grid.  A box's workgroups therefore do
// exactly what they do in a launch of their own.
//
// mhd_launch<K>() issues the single-box launch, or, while a LaunchRecorder
// is active on the calling thread, appends the launch (kernel, arguments,
// grid, block, LDS) to it instead; the ensemble builds its batched launches
// from such records (mhd_api.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "common.hpp"

namespace mhd {

struct BatchTable {
    const unsigned char* args;   // nbox argument blocks of `stride` bytes
    const unsigned* start;       // nbox + 1 cumulative workgroup counts
    const uint3* grid;           // each box's grid
    int nbox, stride;
};

template <typename K>
__global__ void __launch_bounds__(K::NT) single_kernel(typename K::Args av)
{
    (void)av;
    const uint3 b = {blockIdx.x, blockIdx.y, blockIdx.z};
    const uint3 g = {gridDim.x, gridDim.y, gridDim.z};
    K::body(kernarg<typename K::Args>(), b, g);
}

template <typename K>
__global__ void __launch_bounds__(K::NT) batched_kernel(BatchTable t)
{
    const unsigned w = blockIdx.x;
    int lo = 0, hi = t.nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t.start[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    const uint3 g = t.grid[lo];
    unsigned v = w - t.start[lo];
    uint3 b;
    b.x = v % g.x;
    v /= g.x;
    b.y = v % g.y;
    b.z = v / g.y;
    K::body(*reinterpret_cast<const typename K::Args*>(t.args + (size_t)lo * t.stride), b, g);
}

struct LaunchRec;
using BatchFn = int (*)(const BatchTable& t, unsigned total, unsigned nt, size_t lds, hipStream_t s);

// one recorded launch
struct LaunchRec {
    const void* kernel;               // &single_kernel<K>: the kernel's identity
    BatchFn batch;                    // issues batched_kernel<K>
    std::vector<unsigned char> args;  // the argument block
    dim3 grid, block;
    size_t lds;
};

struct LaunchRecorder {
    std::vector<LaunchRec> recs;
};

// the recorder of the calling thread (null: launches are issued)
LaunchRecorder*& launch_recorder();

template <typename K>
int batch_launch(const BatchTable& t, unsigned total, unsigned nt, size_t lds, hipStream_t s)
{
    if (lds > 65536) {
        static int opted = 0;
        if (!opted) {
            if (hipFuncSetAttribute((const void*)batched_kernel<K>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    163840) != hipSuccess)
                return -22;
            opted = 1;
        }
    }
    hipLaunchKernelGGL(batched_kernel<K>, dim3(total), dim3(nt), lds, s, t);
    return hipGetLastError() == hipSuccess ? 0 : -22;
}

template <typename K>
int mhd_launch(dim3 grid, dim3 block, size_t lds, hipStream_t s, const typename K::Args& a)
{
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) return 0;
    if (LaunchRecorder* r = launch_recorder()) {
        LaunchRec rec;
        rec.kernel = (const void*)single_kernel<K>;
        rec.batch = &batch_launch<K>;
        rec.args.resize(sizeof(typename K::Args));
        memcpy(rec.args.data(), &a, sizeof(typename K::Args));
        rec.grid = grid;
        rec.block = block;
        rec.lds = lds;
        r->recs.push_back(std::move(rec));
        return 0;
    }
    hipLaunchKernelGGL(single_kernel<K>, grid, block, lds, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -22;
}

}  // namespace mhd
