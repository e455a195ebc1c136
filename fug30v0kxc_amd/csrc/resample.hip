// resample.hip -- spectral resampling (the reference's expand,
// interpolate/fft.F90:832-927) and the block box filter of coarse-grained
// data, on gfx950.
//
// expand doubles nx and ny by zero-padding the spectrum of every plane.  It
// is composed from the horizontal transform kernels: x r2hc (length nx), a
// padding kernel into rows of 2nx, x hc2r (length 2nx), y r2hc (length ny,
// stride 2nx), a padding kernel into columns of 2ny, y hc2r (length 2ny).
// The padding follows the reference: r_0, the (r_m, i_m) of m = 1..n/2-1
// and the coarse Nyquist r_{n/2} as an ordinary cosine coefficient of mode
// n/2; each inverse divides by the coarse length (the padding kernels apply
// a factor 2 against the transform's 1/(2n)).
#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/mhd.h"
#include "physics.hpp"
#include "spectral.hpp"

namespace mhd {
namespace {

constexpr int RNT = 256;

// rows of nx halfcomplex values -> rows of 2nx (x padding)
__global__ void __launch_bounds__(RNT) pad_x_kernel(const double* __restrict__ in, double* __restrict__ out, int nx,
                                                    long rows)
{
    const long q = (long)blockIdx.x * RNT + threadIdx.x;
    const int n2 = 2 * nx;
    if (q >= rows * n2) return;
    const long row = q / n2;
    const int i = (int)(q - row * n2);
    const double* h = in + row * nx;
    double v = 0.0;
    if (i <= nx / 2) v = h[i];                                   // r_0 .. r_{nx/2}
    else if (i > n2 - nx / 2) v = h[nx - (n2 - i)];              // i_m at 2nx - m
    out[q] = 2.0 * v;
}

// columns (stride nxf) of ny halfcomplex values -> columns of 2ny (y padding)
__global__ void __launch_bounds__(RNT) pad_y_kernel(const double* __restrict__ in, double* __restrict__ out, int nxf,
                                                    int ny, int nz)
{
    const long q = (long)blockIdx.x * RNT + threadIdx.x;
    const int n2 = 2 * ny;
    const long plane = (long)nxf * n2;
    if (q >= plane * nz) return;
    const int k = (int)(q / plane);
    const long rem = q - (long)k * plane;
    const int j = (int)(rem / nxf), i = (int)(rem - (long)j * nxf);
    const double* h = in + (long)k * nxf * ny + i;
    double v = 0.0;
    if (j <= ny / 2) v = h[(long)j * nxf];
    else if (j > n2 - ny / 2) v = h[(long)(ny - (n2 - j)) * nxf];
    out[q] = 2.0 * v;
}

// coarse (I, J, K) = mean of the fine f x f x f block; sums over k, j, i
__global__ void __launch_bounds__(RNT) box_kernel(const double* __restrict__ in, double* __restrict__ out, int nx,
                                                  int ny, int f, int cx, int cy, long nc, double inv)
{
    const long q = (long)blockIdx.x * RNT + threadIdx.x;
    if (q >= nc) return;
    const long pc = (long)cx * cy;
    const int K = (int)(q / pc);
    const long rem = q - (long)K * pc;
    const int J = (int)(rem / cx), I = (int)(rem - (long)J * cx);
    double s = 0.0;
    for (int kk = 0; kk < f; ++kk)
        for (int jj = 0; jj < f; ++jj) {
            const double* row = in + ((long)(f * K + kk) * ny + (f * J + jj)) * nx + (long)f * I;
            for (int ii = 0; ii < f; ++ii) s += row[ii];
        }
    out[q] = s * inv;
}

unsigned nb(long n) { return (unsigned)((n + RNT - 1) / RNT); }

// W_n^t = exp(-2 pi i t / n) for t < n, as (cos, sin) pairs, computed on the
// device so that no host staging is needed
__global__ void __launch_bounds__(RNT) twiddle_kernel(double* __restrict__ tw, int n)
{
    const int t = blockIdx.x * RNT + threadIdx.x;
    if (t >= n) return;
    double s, c;
    sincospi(-2.0 * (double)t / (double)n, &s, &c);
    tw[2 * t] = c;
    tw[2 * t + 1] = s;
}

int ncus()
{
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return n;
}

// one real <-> halfcomplex pass over a (nx, ny, nz) FP64 box, from `in` to
// `out` (the same array for an in-place pass)
int hc_pass(const double* in, double* out, int nx, int ny, int nz, int dir, int mode, const double* tw,
            hipStream_t s)
{
    SpecLaunch L;
    memset(&L, 0, sizeof(L));
    L.g.nx = nx; L.g.ny = ny; L.g.nz = nz; L.g.nxy = nx * ny; L.g.n3 = (long)nx * ny * nz;
    L.dir = dir;
    L.mode = mode;
    L.njobs = 1;
    L.tw = tw;
    L.ncu = ncus();
    SpecJob& J = L.job[0];
    J.in[0] = in;
    J.out[0] = out;
    J.load_op = LD_PLAIN;
    J.scale = 1.0;
    const int e = launch_spectral(L, 0, s);
    return e == -20 ? MHD_E_SIZE : (e ? MHD_E_DEVICE : 0);
}

// device doubles expand needs besides its input and output: the x-padded box
// (2 nx ny nz) and the four twiddle tables (lengths nx, 2nx, ny, 2ny)
size_t expand_work_doubles(int nx, int ny, int nz)
{
    return 2 * (size_t)nx * ny * nz + 2 * (size_t)(3 * nx + 3 * ny);
}

int expand_check(int nx, int ny, int nz, const void* in, const void* out)
{
    if (!in || !out || nz < 1 || nx < 2 || ny < 2) return MHD_E_ARG;
    if (nx % 2) return MHD_E_ODD_NX;
    if (ny % 2) return MHD_E_ODD_NY;
    if (!spectral_len_supported(nx) || !spectral_len_supported(2 * nx) || !spectral_len_supported(ny) ||
        !spectral_len_supported(2 * ny))
        return MHD_E_SIZE;
    return 0;
}

// expand of device arrays: in (nx, ny, nz) -> out (2nx, 2ny, nz), with
// `work` holding expand_work_doubles() doubles; `in` is only read.  The x
// forward pass writes the first quarter of `out`, which the y padding later
// overwrites.
int expand_device(int nx, int ny, int nz, const double* in, double* out, double* work, hipStream_t s)
{
    const long n3 = (long)nx * ny * nz;
    double* b = work;                       // 2 n3: rows of 2nx
    double* twx = b + 2 * n3;
    double* twx2 = twx + 2 * nx;
    double* twy = twx2 + 4 * nx;
    double* twy2 = twy + 2 * ny;
    const int lens[4] = {nx, 2 * nx, ny, 2 * ny};
    double* tws[4] = {twx, twx2, twy, twy2};
    for (int q = 0; q < 4; ++q)
        hipLaunchKernelGGL(twiddle_kernel, dim3(nb(lens[q])), dim3(RNT), 0, s, tws[q], lens[q]);
    if (hipGetLastError() != hipSuccess) return MHD_E_DEVICE;
    int e = hc_pass(in, out, nx, ny, nz, 0, SM_R2HC, twx, s);                 // x forward, length nx
    if (e) return e;
    hipLaunchKernelGGL(pad_x_kernel, dim3(nb(2 * n3)), dim3(RNT), 0, s, out, b, nx, (long)ny * nz);
    if ((e = hc_pass(b, b, 2 * nx, ny, nz, 0, SM_HC2R, twx2, s))) return e;   // x inverse, 2nx
    if ((e = hc_pass(b, b, 2 * nx, ny, nz, 1, SM_R2HC, twy, s))) return e;    // y forward, ny
    hipLaunchKernelGGL(pad_y_kernel, dim3(nb(4 * n3)), dim3(RNT), 0, s, b, out, 2 * nx, ny, nz);
    if ((e = hc_pass(out, out, 2 * nx, 2 * ny, nz, 1, SM_HC2R, twy2, s))) return e;   // y inverse, 2ny
    return hipGetLastError() == hipSuccess ? 0 : MHD_E_DEVICE;
}

}  // namespace
}  // namespace mhd

using namespace mhd;

// Work memory of mhd_expand_noalloc: a fixed device array.  Every GPU the
// library runs on holds its own copy, at its own address, until the process
// ends.  Host code never uses the array's name as a device address: it takes
// the address hipGetSymbolAddress returns for the current device.
static constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)MHD_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[SCRATCH_MAX_DEVICES];

// the device address of `scratch` on the current device, looked up once per device
int scratch_on_device(double** ptr)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_MAX_DEVICES) return MHD_E_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MHD_E_DEVICE;
        g_scratch_ptr[dev] = (double*)p;
    }
    *ptr = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" {

int mhd_expand(int nx, int ny, int nz, const double* in, double* out)
{
    {
        const int e = expand_check(nx, ny, nz, in, out);
        if (e) return e;
    }
    const long n3 = (long)nx * ny * nz;
    double *a = nullptr, *d = nullptr, *work = nullptr;
    hipStream_t s = nullptr;
    int rc = MHD_E_ALLOC;
    if (hipMalloc((void**)&a, n3 * sizeof(double)) == hipSuccess &&
        hipMalloc((void**)&d, 4 * n3 * sizeof(double)) == hipSuccess &&
        hipMalloc((void**)&work, expand_work_doubles(nx, ny, nz) * sizeof(double)) == hipSuccess &&
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess) {
        rc = MHD_E_DEVICE;
        do {
            if (hipMemcpyAsync(a, in, n3 * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess) break;
            const int e = expand_device(nx, ny, nz, a, d, work, s);
            if (e) { rc = e; break; }
            if (hipMemcpyAsync(out, d, 4 * n3 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) break;
            if (hipStreamSynchronize(s) != hipSuccess) break;
            rc = MHD_OK;
        } while (0);
    }
    if (s) (void)hipStreamSynchronize(s);
    for (double* p : {a, d, work})
        if (p) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}

int mhd_expand_noalloc(int nx, int ny, int nz, const double* in, double* out, void* stream)
{
    {
        const int e = expand_check(nx, ny, nz, in, out);
        if (e) return e;
    }
    if (expand_work_doubles(nx, ny, nz) > NOALLOC_SCRATCH_DOUBLES) return MHD_E_SCRATCH;
    double* work = nullptr;
    {
        const int e = scratch_on_device(&work);
        if (e) return e;
    }
    return expand_device(nx, ny, nz, in, out, work, (hipStream_t)stream);
}

int mhd_box_filter(int nx, int ny, int nz, int f, const double* in, double* out)
{
    if (!in || !out || f < 1 || nx < 1 || ny < 1 || nz < 1) return MHD_E_ARG;
    if (nx % f || ny % f || nz % f) return MHD_E_ARG;
    const long n3 = (long)nx * ny * nz;
    const int cx = nx / f, cy = ny / f, cz = nz / f;
    const long nc = (long)cx * cy * cz;
    double *a = nullptr, *c = nullptr;
    int rc = MHD_E_ALLOC;
    if (hipMalloc((void**)&a, n3 * sizeof(double)) == hipSuccess &&
        hipMalloc((void**)&c, nc * sizeof(double)) == hipSuccess) {
        rc = MHD_E_DEVICE;
        const double inv = 1.0 / ((double)f * (double)f * (double)f);
        if (hipMemcpy(a, in, n3 * sizeof(double), hipMemcpyHostToDevice) == hipSuccess) {
            hipLaunchKernelGGL(box_kernel, dim3(nb(nc)), dim3(RNT), 0, 0, a, c, nx, ny, f, cx, cy, nc, inv);
            if (hipGetLastError() == hipSuccess &&
                hipMemcpy(out, c, nc * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess)
                rc = MHD_OK;
        }
    }
    if (a) (void)hipFree(a);
    if (c) (void)hipFree(c);
    return rc;
}

}  // extern "C"
