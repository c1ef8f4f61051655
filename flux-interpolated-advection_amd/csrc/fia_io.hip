// fia_io.hip - reductions behind the reference's diagnostic output: total mass and momenta of a
// grid (reference src/main.cpp:518-597, written to the *.dat series) and the momentum the walls
// take out through the pressure at each projection (reference src/calc_pressure_3d.cpp:235-266).
// FP64 sums over FP32 fields in a fixed order (per-block partials, then one block).
#include <mutex>

#include "fia_internal.hpp"
#include "fia_math.hpp"

namespace fia {

constexpr int kIB = 256;
constexpr int kIBlocks = 256;

__global__ void __launch_bounds__(kIB) sum_kernel(const float *a, size_t n, double *part) {
    __shared__ double sh[kIB];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * kIB + threadIdx.x; i < n; i += (size_t)gridDim.x * kIB) s += (double)a[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = kIB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// sums of the pressure over the six boundary layers of cells: x = 0, x = nx-1, y = 0, y = ny-1,
// z = 0, z = nz-1
__global__ void __launch_bounds__(kIB) wall_sum_kernel(const float *p, int nx, int ny, int nz, double *part) {
    __shared__ double sh[6][kIB];
    double s[6] = {0, 0, 0, 0, 0, 0};
    const size_t n = (size_t)nx * ny * nz;
    for (size_t c = (size_t)blockIdx.x * kIB + threadIdx.x; c < n; c += (size_t)gridDim.x * kIB) {
        const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((size_t)ny * nz));
        const double v = p[c];
        if (i == 0) s[0] += v;
        if (i == nx - 1) s[1] += v;
        if (j == 0) s[2] += v;
        if (j == ny - 1) s[3] += v;
        if (k == 0) s[4] += v;
        if (k == nz - 1) s[5] += v;
    }
#pragma unroll
    for (int a = 0; a < 6; ++a) sh[a][threadIdx.x] = s[a];
    __syncthreads();
    for (int off = kIB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
#pragma unroll
            for (int a = 0; a < 6; ++a) sh[a][threadIdx.x] += sh[a][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int a = 0; a < 6; ++a) part[a * kIBlocks + blockIdx.x] = sh[a][0];
}

// out[k] = sum of the kIBlocks partials of set k, for k < sets
__global__ void finish_sums_kernel(const double *part, int sets, double *out) {
    const int k = threadIdx.x;
    if (k >= sets) return;
    double s = 0.0;
    for (int b = 0; b < kIBlocks; ++b) s += part[k * kIBlocks + b];
    out[k] = s;
}

size_t io_partial_words() { return 6 * (size_t)kIBlocks + 8; }

// Fixed device scratch of the allocation-free totals (fia_grid_totals_noalloc): the partial sums
// and the four results of one call, io_partial_words() doubles (12,352 bytes), held by every GPU
// the library runs on until the process ends.  In host code `scratch` only stands for the symbol:
// kernels and copies use the device address io_scratch() returns for the current device.
static __device__ __align__(256) double scratch[6 * kIBlocks + 8];

constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mutex;

// Looks up the scratch array's address on `device` (which must be current) once and keeps it.
hipError_t io_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = (double *)ptr;
    return hipSuccess;
}

double *io_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

hipError_t launch_field_sums(const float *const fields[4], const size_t sizes[4], double *part, double *out,
                             hipStream_t s) {
    for (int k = 0; k < 4; ++k)   // a missing field sums to 0 (its blocks write zero partials)
        hipLaunchKernelGGL(sum_kernel, dim3(kIBlocks), dim3(kIB), 0, s, fields[k], fields[k] ? sizes[k] : (size_t)0,
                           part + k * kIBlocks);
    hipLaunchKernelGGL(finish_sums_kernel, dim3(1), dim3(64), 0, s, part, 4, out);
    return hipGetLastError();
}

hipError_t launch_wall_sums(const float *p, int nx, int ny, int nz, double *part, double *out, hipStream_t s) {
    hipLaunchKernelGGL(wall_sum_kernel, dim3(kIBlocks), dim3(kIB), 0, s, p, nx, ny, nz, part);
    hipLaunchKernelGGL(finish_sums_kernel, dim3(1), dim3(64), 0, s, part, 6, out);
    return hipGetLastError();
}

}  // namespace fia
