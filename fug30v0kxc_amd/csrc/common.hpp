// common.hpp -- shared device/host definitions for the MI355X (gfx950)
// compressible-MHD time-step library.
//
// Layout contract (same as the reference's Fortran arrays): every 3-D field
// f(nx,ny,nz) is stored plane-major, element (i,j,k) at i + nx*(j + ny*k),
// 256-byte aligned.  x is the contiguous direction, so row (x) transforms read
// whole cache lines and column (y) transforms read tiles of adjacent columns.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mhd {

// complex value held in registers / LDS
template <typename T>
struct cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
// multiply by +i (s=+1) or -i (s=-1)
template <typename T, int S>
__device__ __forceinline__ cpx<T> cmul_i(cpx<T> a)
{
    if constexpr (S > 0) return {-a.y, a.x};
    else return {a.y, -a.x};
}

// element loads with on-the-fly conversion to the compute type
template <typename TC>
__device__ __forceinline__ TC ldv(const void* p, size_t i, int f32)
{
    return f32 ? (TC)(((const float*)p)[i]) : (TC)(((const double*)p)[i]);
}
template <typename TC>
__device__ __forceinline__ void stv(void* p, size_t i, int f32, TC v)
{
    if (f32) ((float*)p)[i] = (float)v;
    else ((double*)p)[i] = (double)v;
}

// The by-value launch struct of a kernel, read in place from the kernarg
// segment (constant address space).  Indexing its arrays with a
// workgroup-uniform index then compiles to scalar loads instead of a private
// copy of the whole struct.  T must be the kernel's first argument.
template <typename T>
__device__ __forceinline__ const T& kernarg()
{
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const T*)(__builtin_amdgcn_kernarg_segment_ptr());
#else
    // host pass: device-only function, never emitted or executed
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wnull-dereference"
    return *(const T*)nullptr;
#pragma clang diagnostic pop
#endif
}

// Box geometry passed by value to every kernel.
struct Geom {
    int nx, ny, nz;
    int nxy;        // nx*ny
    long n3;        // nx*ny*nz
};

}  // namespace mhd

#define MHD_HIP_CHECK(expr)                                                    \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) return mhd_hip_error(_e, #expr, __FILE__, __LINE__); \
    } while (0)

int mhd_hip_error(hipError_t e, const char* what, const char* file, int line);
