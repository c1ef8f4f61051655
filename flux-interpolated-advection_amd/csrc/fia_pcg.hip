// fia_pcg.hip - pressure projection on gfx950: Jacobi-preconditioned CG on the reference's
// 7-point Neumann Laplacian (reference src/calc_pressure_3d.cpp:18-180), FP64 vectors, FP64
// dot products reduced in a fixed order (deterministic), and the velocity update
// u -= grad p / h (reference src/calc_pressure_3d.cpp:196-234).
//
// System: L p = b with L = -A (A as assembled by the reference: diag -6 + #walls, +1 per
// neighbour), b = -div(u) h, the last cell pinned to p = 0.  Obstacle cells (optional mask) are
// removed from the system and their faces act as walls.
//
// Scalars live in device memory (scal[]): the iteration runs without host round trips; once the
// residual test passes, every update becomes a no-op, so the host only polls now and then.
#include "fia_internal.hpp"
#include "fia_math.hpp"

namespace fia {

constexpr int kPB = 256;

// scal layout
enum { S_RZ = 0, S_PQ = 1, S_RZNEW = 2, S_RR = 3, S_BB = 4, S_TOL2 = 5, S_DONE = 6, S_ITERS = 7 };

__device__ __forceinline__ bool fluid(const PcgGrid &g, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) return false;
    return !(g.solid && g.solid[idx3(i, j, k, g.ny, g.nz)]);
}

__global__ void pcg_setup_kernel(PcgGrid g, const float *u, const float *v, const float *w, double h, double *b,
                                 double *diag) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)g.nx * g.ny * g.nz;
    if (c >= n) return;
    const int k = (int)(c % g.nz), j = (int)((c / g.nz) % g.ny), i = (int)(c / ((long)g.ny * g.nz));
    if (!fluid(g, i, j, k) || c == g.pin) { b[c] = 0.0; diag[c] = 0.0; return; }
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const double div = ((double)u[idx3(i + 1, j, k, ny, nz)] - (double)u[idx3(i, j, k, ny, nz)])
                     + ((double)v[idx3(i, j + 1, k, ny + 1, nz)] - (double)v[idx3(i, j, k, ny + 1, nz)])
                     + ((double)w[idx3(i, j, k + 1, ny, nz + 1)] - (double)w[idx3(i, j, k, ny, nz + 1)]);
    (void)nx;
    b[c] = -div * h;
    diag[c] = (double)(fluid(g, i - 1, j, k) + fluid(g, i + 1, j, k) + fluid(g, i, j - 1, k) +
                       fluid(g, i, j + 1, k) + fluid(g, i, j, k - 1) + fluid(g, i, j, k + 1));
}

// y = L x on free cells (zero elsewhere)
__global__ void pcg_apply_kernel(PcgGrid g, const double *diag, const double *x, double *y) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)g.nx * g.ny * g.nz;
    if (c >= n) return;
    const double d = diag[c];
    if (d == 0.0) { y[c] = 0.0; return; }
    const int k = (int)(c % g.nz), j = (int)((c / g.nz) % g.ny), i = (int)(c / ((long)g.ny * g.nz));
    double s = d * x[c];
    // neighbours outside the system (walls, obstacles, the pinned cell) hold x = 0
    if (i > 0) s -= x[c - (long)g.ny * g.nz];
    if (i + 1 < g.nx) s -= x[c + (long)g.ny * g.nz];
    if (j > 0) s -= x[c - g.nz];
    if (j + 1 < g.ny) s -= x[c + g.nz];
    if (k > 0) s -= x[c - 1];
    if (k + 1 < g.nz) s -= x[c + 1];
    y[c] = s;
}

// r = b - L x
__global__ void pcg_residual_kernel(PcgGrid g, const double *diag, const double *b, const double *x, double *r) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)g.nx * g.ny * g.nz;
    if (c >= n) return;
    const double d = diag[c];
    if (d == 0.0) { r[c] = 0.0; return; }
    const int k = (int)(c % g.nz), j = (int)((c / g.nz) % g.ny), i = (int)(c / ((long)g.ny * g.nz));
    double s = d * x[c];
    if (i > 0) s -= x[c - (long)g.ny * g.nz];
    if (i + 1 < g.nx) s -= x[c + (long)g.ny * g.nz];
    if (j > 0) s -= x[c - g.nz];
    if (j + 1 < g.ny) s -= x[c + g.nz];
    if (k > 0) s -= x[c - 1];
    if (k + 1 < g.nz) s -= x[c + 1];
    r[c] = b[c] - s;
}

// fixed-order two-level dot product: partials[block], then one block sums the partials
__global__ void __launch_bounds__(kPB) dot_kernel(const double *a, const double *b, size_t n, double *partials) {
    __shared__ double sh[kPB];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += a[i] * b[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = kPB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kPB) finish_dot_kernel(const double *partials, int nparts, double *out) {
    __shared__ double sh[kPB];
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kPB) s += partials[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = kPB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

__global__ void pcg_update_xr_kernel(double *x, double *r, const double *p, const double *q, const double *scal,
                                     size_t n) {
    if (scal[S_DONE] != 0.0) return;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double alpha = scal[S_RZ] / scal[S_PQ];
    x[i] += alpha * p[i];
    r[i] -= alpha * q[i];
}

__global__ void pcg_precond_kernel(const double *r, const double *diag, double *z, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = diag[i];
    z[i] = d > 0.0 ? r[i] / d : 0.0;
}

__global__ void pcg_update_p_kernel(double *p, const double *z, const double *scal, size_t n) {
    if (scal[S_DONE] != 0.0) return;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double beta = scal[S_RZNEW] / scal[S_RZ];
    p[i] = z[i] + beta * p[i];
}

// stage 0: after the initial residual: rz = rznew, done = rr <= tol2 * bb
// stage 1: after an iteration's dots: iters++, done test, then beta uses rz (old) in update_p,
//          so rz is advanced by stage 2 after update_p
__global__ void pcg_scalars_kernel(double *scal, int stage) {
    if (stage == 0) {
        scal[S_RZ] = scal[S_RZNEW];
        scal[S_DONE] = scal[S_RR] <= scal[S_TOL2] * scal[S_BB] ? 1.0 : 0.0;
        scal[S_ITERS] = 0.0;
    } else if (stage == 1) {
        if (scal[S_DONE] != 0.0) return;
        scal[S_ITERS] += 1.0;
    } else {
        if (scal[S_DONE] != 0.0) return;
        scal[S_RZ] = scal[S_RZNEW];
        if (scal[S_RR] <= scal[S_TOL2] * scal[S_BB]) scal[S_DONE] = 1.0;
    }
}

__global__ void project_kernel(PcgGrid g, const double *p, double h, float *u, float *v, float *w,
                               float *pressure_out) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)g.nx * g.ny * g.nz;
    if (c >= n) return;
    const int k = (int)(c % g.nz), j = (int)((c / g.nz) % g.ny), i = (int)(c / ((long)g.ny * g.nz));
    const int ny = g.ny, nz = g.nz;
    if (pressure_out) pressure_out[c] = (float)p[c];
    if (!fluid(g, i, j, k)) return;
    // the face on the low side of this cell in each direction (interior faces only)
    if (i > 0 && fluid(g, i - 1, j, k))
        u[idx3(i, j, k, ny, nz)] = (float)((double)u[idx3(i, j, k, ny, nz)] - (p[c] - p[c - (long)ny * nz]) / h);
    if (j > 0 && fluid(g, i, j - 1, k))
        v[idx3(i, j, k, ny + 1, nz)] = (float)((double)v[idx3(i, j, k, ny + 1, nz)] - (p[c] - p[c - nz]) / h);
    if (k > 0 && fluid(g, i, j, k - 1))
        w[idx3(i, j, k, ny, nz + 1)] = (float)((double)w[idx3(i, j, k, ny, nz + 1)] - (p[c] - p[c - 1]) / h);
}

__global__ void zero_solid_faces_kernel(PcgGrid g, float *u, float *v, float *w) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)g.nx * g.ny * g.nz;
    if (c >= n) return;
    const int k = (int)(c % g.nz), j = (int)((c / g.nz) % g.ny), i = (int)(c / ((long)g.ny * g.nz));
    if (fluid(g, i, j, k)) return;
    const int ny = g.ny, nz = g.nz;
    u[idx3(i, j, k, ny, nz)] = 0.0f; u[idx3(i + 1, j, k, ny, nz)] = 0.0f;
    v[idx3(i, j, k, ny + 1, nz)] = 0.0f; v[idx3(i, j + 1, k, ny + 1, nz)] = 0.0f;
    w[idx3(i, j, k, ny, nz + 1)] = 0.0f; w[idx3(i, j, k + 1, ny, nz + 1)] = 0.0f;
}

__global__ void fill_double_kernel(double *a, double v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void pressure_in_kernel(PcgGrid g, const float *p_in, double *x) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)g.nx * g.ny * g.nz;
    if (c >= n) return;
    const int k = (int)(c % g.nz), j = (int)((c / g.nz) % g.ny), i = (int)(c / ((long)g.ny * g.nz));
    x[c] = (fluid(g, i, j, k) && c != g.pin) ? (double)p_in[c] : 0.0;
}

static inline unsigned blocks(size_t n) { return (unsigned)((n + kPB - 1) / kPB); }

hipError_t launch_pcg_setup(const PcgGrid &g, const float *u, const float *v, const float *w, double h, double *b,
                            double *diag, hipStream_t s) {
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(pcg_setup_kernel, dim3(blocks(n)), dim3(kPB), 0, s, g, u, v, w, h, b, diag);
    return hipGetLastError();
}
hipError_t launch_pcg_apply(const PcgGrid &g, const double *diag, const double *x, double *y, hipStream_t s) {
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(pcg_apply_kernel, dim3(blocks(n)), dim3(kPB), 0, s, g, diag, x, y);
    return hipGetLastError();
}
hipError_t launch_pcg_residual(const PcgGrid &g, const double *diag, const double *b, const double *x, double *r,
                               hipStream_t s) {
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(pcg_residual_kernel, dim3(blocks(n)), dim3(kPB), 0, s, g, diag, b, x, r);
    return hipGetLastError();
}
hipError_t launch_dot(const double *a, const double *b, size_t n, double *partials, int nparts, hipStream_t s) {
    hipLaunchKernelGGL(dot_kernel, dim3(nparts), dim3(kPB), 0, s, a, b, n, partials);
    return hipGetLastError();
}
hipError_t launch_finish_dot(const double *partials, int nparts, double *out, hipStream_t s) {
    hipLaunchKernelGGL(finish_dot_kernel, dim3(1), dim3(kPB), 0, s, partials, nparts, out);
    return hipGetLastError();
}
hipError_t launch_pcg_update_xr(double *x, double *r, const double *p, const double *q, const double *scal, size_t n,
                                hipStream_t s) {
    hipLaunchKernelGGL(pcg_update_xr_kernel, dim3(blocks(n)), dim3(kPB), 0, s, x, r, p, q, scal, n);
    return hipGetLastError();
}
hipError_t launch_pcg_precond(const double *r, const double *diag, double *z, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(pcg_precond_kernel, dim3(blocks(n)), dim3(kPB), 0, s, r, diag, z, n);
    return hipGetLastError();
}
hipError_t launch_pcg_update_p(double *p, const double *z, const double *scal, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(pcg_update_p_kernel, dim3(blocks(n)), dim3(kPB), 0, s, p, z, scal, n);
    return hipGetLastError();
}
hipError_t launch_pcg_scalars(double *scal, int stage, hipStream_t s) {
    hipLaunchKernelGGL(pcg_scalars_kernel, dim3(1), dim3(1), 0, s, scal, stage);
    return hipGetLastError();
}
hipError_t launch_project(const PcgGrid &g, const double *p, double h, float *u, float *v, float *w,
                          float *pressure_out, hipStream_t s) {
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(project_kernel, dim3(blocks(n)), dim3(kPB), 0, s, g, p, h, u, v, w, pressure_out);
    return hipGetLastError();
}
hipError_t launch_zero_solid_faces(const PcgGrid &g, float *u, float *v, float *w, hipStream_t s) {
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(zero_solid_faces_kernel, dim3(blocks(n)), dim3(kPB), 0, s, g, u, v, w);
    return hipGetLastError();
}
hipError_t launch_fill_double(double *a, double v, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_double_kernel, dim3(blocks(n)), dim3(kPB), 0, s, a, v, n);
    return hipGetLastError();
}
hipError_t launch_pressure_in(const PcgGrid &g, const float *p_in, double *x, hipStream_t s) {
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(pressure_in_kernel, dim3(blocks(n)), dim3(kPB), 0, s, g, p_in, x);
    return hipGetLastError();
}

}  // namespace fia
