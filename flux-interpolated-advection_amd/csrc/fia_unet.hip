// fia_unet.hip - kernels and host sequencing of the learned pressure warm start (the U-Net of
// fia_unet.hpp).  No counterpart in the reference.
//
// 3x3x3 convolutions are implicit GEMMs on the matrix cores: rows = voxels, columns = output
// channels, K = 27 taps x input channels.  A block owns a 4 x 4 x 16 voxel tile; its input plus a
// one-voxel halo is staged in LDS 32 channels at a time and read by all 27 taps, one
// v_mfma_f32_16x16x32_bf16 per 16 voxels x 16 output channels x 32 input channels.  A wave owns
// 64 voxels (4 row blocks) and all output channels, so every A fragment feeds all column blocks.
// The data gradient is the same kernel on the flipped, transposed weights.  The weight gradient stages 64-voxel tiles of the pre-activation gradient
// and of the tap-shifted input in LDS and accumulates per tap and voxel chunk; chunks are then
// summed in a fixed order (no atomics: training is deterministic).
//
// MFMA operand layouts (64 lanes): 16x16x16 BF16: lane l holds A[l % 16][4 (l / 16) + i] and
// B[4 (l / 16) + i][l % 16] for i < 4; 16x16x32 BF16: A[l % 16][8 (l / 16) + i] and
// B[8 (l / 16) + i][l % 16] for i < 8; both: D[4 (l / 16) + i][l % 16] for i < 4.
#include <math.h>
#include <string.h>

#include <vector>

#include "fia_unet.hpp"

namespace fia {

const UNetConv kUNetTable[kUNetConvs] = {
    {16, 16, 0, 2},    // e0a: rhs, mask (+14 zero channels)
    {16, 16, 0, 16},   // e0b -> s0
    {16, 32, 1, 16},   // e1a (pooled s0)
    {32, 32, 1, 32},   // e1b -> s1
    {32, 48, 2, 32},   // e2a (pooled s1)
    {48, 48, 2, 48},   // e2b -> s2
    {48, 64, 3, 48},   // ba (pooled s2)
    {64, 64, 3, 64},   // bb
    {112, 48, 2, 112}, // d2a: up(bb) 64 + s2 48
    {48, 48, 2, 48},   // d2b
    {80, 32, 1, 80},   // d1a: up(d2b) 48 + s1 32
    {32, 32, 1, 32},   // d1b
    {48, 16, 0, 48},   // d0a: up(d1b) 32 + s0 16
    {16, 16, 0, 16},   // d0b
};

size_t unet_weight_offset(int conv) {
    size_t o = 0;
    for (int j = 0; j < conv; ++j) o += (size_t)27 * kUNetTable[j].cin * kUNetTable[j].cout + kUNetTable[j].cout;
    return o;
}
size_t unet_bias_offset(int conv) {
    return unet_weight_offset(conv) + (size_t)27 * kUNetTable[conv].cin * kUNetTable[conv].cout;
}
size_t unet_out_offset() { return unet_weight_offset(kUNetConvs); }
size_t unet_param_count() { return unet_out_offset() + 17; }

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((unsigned)h << 16); }
// round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ long vox(int x, int y, int z, int ny, int nz) { return ((long)x * ny + y) * nz + z; }

// channels [c, c + 4) of a convolution's input at an in-range voxel
__device__ __forceinline__ bf16x4 load4(const ConvArgs &a, int x, int y, int z, int c) {
    const uint16_t *p;
    if (c < a.cA)
        p = a.upA ? a.srcA + vox(x >> 1, y >> 1, z >> 1, a.ny >> 1, a.nz >> 1) * a.cA + c
                  : a.srcA + vox(x, y, z, a.ny, a.nz) * a.cA + c;
    else
        p = a.srcB + vox(x, y, z, a.ny, a.nz) * a.cB + (c - a.cA);
    return *reinterpret_cast<const bf16x4 *>(p);
}

constexpr int kCB = 256;   // threads per convolution block (4 waves x 64 voxels)

// Block tile of the convolution: 4 (x) x 4 (y) x 16 (z) voxels; wave w owns x-plane w of the tile,
// its 4 row blocks are the tile's 4 y-rows of 16 z-voxels.  The input of the tile plus a one-voxel
// halo (6 x 6 x 18 voxels) is staged in LDS for 32 input channels at a time and read by all 27
// taps.  Full 32-channel slices use v_mfma_f32_16x16x32_bf16 (lane l: A[l % 16][8 (l / 16) + i],
// B[8 (l / 16) + i][l % 16], i < 8); a 16-channel tail uses the 16x16x16 form.
constexpr int kTX = 4, kTY = 4, kTZ = 16;
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2, kHalo = kHX * kHY * kHZ;
constexpr int kSlice = 32, kSliceStride = kSlice + 8;   // LDS row: 32 channels + padding (bank spread)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short i16x8 __attribute__((ext_vector_type(8)));

// eight channels [c, c + 8) of a convolution's input at an in-range voxel (c % 8 == 0: an 8-group
// never straddles the two sources, whose channel counts are multiples of 16)
__device__ __forceinline__ i16x8 load8(const ConvArgs &a, int x, int y, int z, int c) {
    const uint16_t *p;
    if (c < a.cA)
        p = a.upA ? a.srcA + vox(x >> 1, y >> 1, z >> 1, a.ny >> 1, a.nz >> 1) * a.cA + c
                  : a.srcA + vox(x, y, z, a.ny, a.nz) * a.cA + c;
    else
        p = a.srcB + vox(x, y, z, a.ny, a.nz) * a.cB + (c - a.cA);
    return *reinterpret_cast<const i16x8 *>(p);
}

// Stages input channels [c0, c0 + 8 groups) of the block's halo tile in LDS.
__device__ __forceinline__ void stage_slice(const ConvArgs &a, uint16_t *tile, int x0, int y0, int z0, int c0,
                                            int groups) {
    for (int i = threadIdx.x; i < kHalo * groups; i += kCB) {
        const int g = i % groups, v = i / groups;
        const int hz = v % kHZ, hy = (v / kHZ) % kHY, hx = v / (kHZ * kHY);
        const int x = x0 - 1 + hx, y = y0 - 1 + hy, z = z0 - 1 + hz;
        i16x8 val = i16x8{0, 0, 0, 0, 0, 0, 0, 0};
        if (x >= 0 && x < a.nx && y >= 0 && y < a.ny && z >= 0 && z < a.nz) val = load8(a, x, y, z, c0 + 8 * g);
        *reinterpret_cast<i16x8 *>(&tile[v * kSliceStride + 8 * g]) = val;
    }
}

// All 27 taps over one staged slice: K = 32 channels (FULL, 16x16x32) or 16 (16x16x16).
template <int NB, bool FULL>
__device__ __forceinline__ void slice_taps(const ConvArgs &a, const uint16_t *tile, int c0, int cin, int wave, int row,
                                           int grp, f32x4 (&acc)[4][NB]) {
    for (int t = 0; t < 27; ++t) {
        const int dx = t / 9, dy = (t / 3) % 3, dz = t % 3;   // halo offsets 0..2
        const uint16_t *wt = a.w + (size_t)t * a.cout * cin + c0;
        if (FULL) {
            bf16x8 A[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int v = ((wave + dx) * kHY + (m + dy)) * kHZ + (row + dz);
                A[m] = *reinterpret_cast<const bf16x8 *>(&tile[v * kSliceStride + 8 * grp]);
            }
#pragma unroll
            for (int n = 0; n < NB; ++n) {
                const bf16x8 B = *reinterpret_cast<const bf16x8 *>(wt + (size_t)(n * 16 + row) * cin + 8 * grp);
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, acc[m][n], 0, 0, 0);
            }
        } else {
            bf16x4 A[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int v = ((wave + dx) * kHY + (m + dy)) * kHZ + (row + dz);
                A[m] = *reinterpret_cast<const bf16x4 *>(&tile[v * kSliceStride + 4 * grp]);
            }
#pragma unroll
            for (int n = 0; n < NB; ++n) {
                const bf16x4 B = *reinterpret_cast<const bf16x4 *>(wt + (size_t)(n * 16 + row) * cin + 4 * grp);
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[m], B, acc[m][n], 0, 0, 0);
            }
        }
    }
}

template <int NB, int MODE>
__global__ void __launch_bounds__(kCB) conv3_kernel(ConvArgs a) {
    __shared__ uint16_t tile[kHalo * kSliceStride];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, grp = lane >> 4;
    const int tx = (a.nx + kTX - 1) / kTX, ty = (a.ny + kTY - 1) / kTY;
    const int bx = blockIdx.x % tx, by = (blockIdx.x / tx) % ty, bz = blockIdx.x / (tx * ty);
    const int x0 = bx * kTX, y0 = by * kTY, z0 = bz * kTZ;
    const int cin = a.cA + a.cB;
    f32x4 acc[4][NB];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int full = cin / kSlice * kSlice;
    for (int c0 = 0; c0 < full; c0 += kSlice) {
        __syncthreads();   // the previous slice's reads are done
        stage_slice(a, tile, x0, y0, z0, c0, kSlice / 8);
        __syncthreads();
        slice_taps<NB, true>(a, tile, c0, cin, wave, row, grp, acc);
    }
    if (full < cin) {   // a 16-channel tail
        __syncthreads();
        stage_slice(a, tile, x0, y0, z0, full, 2);
        __syncthreads();
        slice_taps<NB, false>(a, tile, full, cin, wave, row, grp, acc);
    }
    // D[i][j] in lane l: j = l % 16 (output channel), i = 4 (l / 16) + r (z-voxel of the row block)
    const int x = x0 + wave;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int y = y0 + m;
#pragma unroll
        for (int n = 0; n < NB; ++n) {
            const int co = n * 16 + row;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int z = z0 + 4 * grp + r;
                if (x >= a.nx || y >= a.ny || z >= a.nz) continue;
                const long v = vox(x, y, z, a.ny, a.nz);
                float val = acc[m][n][r];
                if (MODE == 0) {
                    val += a.bias[co];
                    if (a.relu) val = fmaxf(val, 0.f);
                    static_cast<uint16_t *>(a.out)[v * a.cout + co] = f2bf(val);
                } else if (MODE == 1) {
                    static_cast<float *>(a.out)[v * a.cout + co] = val;
                } else {
                    static_cast<float *>(a.out)[v * a.cout + co] += val;
                }
            }
        }
    }
}

template <int MODE>
static hipError_t conv3_nb(const ConvArgs &a, hipStream_t s) {
    const long tiles = (long)((a.nx + kTX - 1) / kTX) * ((a.ny + kTY - 1) / kTY) * ((a.nz + kTZ - 1) / kTZ);
    const dim3 grid((unsigned)tiles);
    switch (a.cout / 16) {
    case 1: hipLaunchKernelGGL((conv3_kernel<1, MODE>), grid, dim3(kCB), 0, s, a); break;
    case 2: hipLaunchKernelGGL((conv3_kernel<2, MODE>), grid, dim3(kCB), 0, s, a); break;
    case 3: hipLaunchKernelGGL((conv3_kernel<3, MODE>), grid, dim3(kCB), 0, s, a); break;
    case 4: hipLaunchKernelGGL((conv3_kernel<4, MODE>), grid, dim3(kCB), 0, s, a); break;
    case 5: hipLaunchKernelGGL((conv3_kernel<5, MODE>), grid, dim3(kCB), 0, s, a); break;
    case 7: hipLaunchKernelGGL((conv3_kernel<7, MODE>), grid, dim3(kCB), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_conv3(const ConvArgs &a, int mode, hipStream_t s) {
    if (a.cout % 16 || (a.cA + a.cB) % 16 || a.cA % 16 || a.nx <= 0 || a.ny <= 0 || a.nz <= 0)
        return hipErrorInvalidValue;
    if (a.upA && (a.nx % 2 || a.ny % 2 || a.nz % 2)) return hipErrorInvalidValue;
    if (mode == 0) return conv3_nb<0>(a, s);
    if (mode == 1) return conv3_nb<1>(a, s);
    return conv3_nb<2>(a, s);
}

// ---- weight gradient ---------------------------------------------------------------------------

constexpr int kWgVox = 64;        // voxels per LDS tile
constexpr int kWgChunk = 4096;    // voxels per block
constexpr int kWgMaxTiles = 6;    // 16x16 output tiles per wave (cout/16 * cin/16 <= 24)
constexpr int kWgMaxCout = 64, kWgMaxCin = 112;

size_t conv3_wgrad_chunks(size_t n) { return (n + kWgChunk - 1) / kWgChunk; }

__global__ void __launch_bounds__(256) conv3_wgrad_kernel(ConvArgs a, const uint16_t *dpre, float *part) {
    __shared__ uint16_t sD[kWgVox][kWgMaxCout + 4];
    __shared__ uint16_t sI[kWgVox][kWgMaxCin + 4];
    const int t = blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kq = (lane >> 4) * 4;
    const int cin = a.cA + a.cB, cout = a.cout;
    const int nbi = cin / 16, ntiles = (cout / 16) * nbi;
    const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
    const long nvox = (long)a.nx * a.ny * a.nz;
    const long vb = (long)blockIdx.x * kWgChunk, ve = min(nvox, vb + kWgChunk);
    f32x4 acc[kWgMaxTiles];
#pragma unroll
    for (int s = 0; s < kWgMaxTiles; ++s) acc[s] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long v0 = vb; v0 < ve; v0 += kWgVox) {
        const int q4o = cout / 4, q4i = cin / 4;
        for (int i = threadIdx.x; i < kWgVox * q4o; i += 256) {
            const int r = i / q4o, c4 = (i % q4o) * 4;
            const long v = v0 + r;
            const bf16x4 val = v < ve ? *reinterpret_cast<const bf16x4 *>(dpre + v * cout + c4) : bf16x4{0, 0, 0, 0};
#pragma unroll
            for (int e = 0; e < 4; ++e) sD[r][c4 + e] = (uint16_t)val[e];
        }
        for (int i = threadIdx.x; i < kWgVox * q4i; i += 256) {
            const int r = i / q4i, c4 = (i % q4i) * 4;
            const long v = v0 + r;
            bf16x4 val = bf16x4{0, 0, 0, 0};
            if (v < ve) {
                const int z = (int)(v % a.nz) + dz, y = (int)((v / a.nz) % a.ny) + dy, x = (int)(v / ((long)a.ny * a.nz)) + dx;
                if (x >= 0 && x < a.nx && y >= 0 && y < a.ny && z >= 0 && z < a.nz) val = load4(a, x, y, z, c4);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) sI[r][c4 + e] = (uint16_t)val[e];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kWgMaxTiles; ++s) {
            const int tile = wave + 4 * s;
            if (tile >= ntiles) continue;
            const int ob = tile / nbi, ib = tile % nbi;
#pragma unroll
            for (int ks = 0; ks < kWgVox / 16; ++ks) {
                bf16x4 A, B;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    A[e] = (short)sD[ks * 16 + kq + e][ob * 16 + row];
                    B[e] = (short)sI[ks * 16 + kq + e][ib * 16 + row];
                }
                acc[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, B, acc[s], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < kWgMaxTiles; ++s) {
        const int tile = wave + 4 * s;
        if (tile >= ntiles) continue;
        const int ob = tile / nbi, ib = tile % nbi;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int co = ob * 16 + kq + r, ci = ib * 16 + row;
            part[(((size_t)blockIdx.x * 27 + t) * cout + co) * cin + ci] = acc[s][r];
        }
    }
}

// grad_w[i] += sum over chunks of part[chunk][i] (chunk order)
__global__ void reduce_chunks_kernel(const float *part, size_t chunks, size_t per, float *grad_w) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per) return;
    float s = 0.f;
    for (size_t c = 0; c < chunks; ++c) s += part[c * per + i];
    grad_w[i] += s;
}

hipError_t launch_conv3_wgrad(const ConvArgs &a, const uint16_t *dpre, float *part, float *grad_w, hipStream_t s) {
    const int cin = a.cA + a.cB;
    if (a.cout > kWgMaxCout || cin > kWgMaxCin || (a.cout / 16) * (cin / 16) > 4 * kWgMaxTiles)
        return hipErrorInvalidValue;
    const long nvox = (long)a.nx * a.ny * a.nz;
    const size_t chunks = conv3_wgrad_chunks((size_t)nvox);
    hipLaunchKernelGGL(conv3_wgrad_kernel, dim3((unsigned)chunks, 27), dim3(256), 0, s, a, dpre, part);
    const size_t per = (size_t)27 * a.cout * cin;
    hipLaunchKernelGGL(reduce_chunks_kernel, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, s, part, chunks, per,
                       grad_w);
    return hipGetLastError();
}

// ---- element-wise pieces ---------------------------------------------------------------------

constexpr int kEB = 256;
static inline unsigned eblocks(size_t n) { return (unsigned)((n + kEB - 1) / kEB); }

constexpr int kRedBlocks = 256;
// red_ slots: [0, 4 kRedBlocks) partials, then the scale, its inverse, the loss, |b|^2, the loss
// gradient coefficient, the residual norm ratio |L y - b|^2 / |b|^2
enum { kRedScale = 4 * kRedBlocks, kRedInvScale, kRedLoss, kRedLossDen, kRedLossCoef, kRedResidual, kRedSlots };

// sum of rhs^2 and count over the fluid cells, per block
__global__ void __launch_bounds__(kEB) rhs_sumsq_kernel(const double *rhs, const uint8_t *solid, size_t n, double *part) {
    __shared__ double sh[2][kEB];
    double s = 0.0, c = 0.0;
    for (size_t i = (size_t)blockIdx.x * kEB + threadIdx.x; i < n; i += (size_t)gridDim.x * kEB) {
        if (solid && solid[i]) continue;
        s += rhs[i] * rhs[i];
        c += 1.0;
    }
    sh[0][threadIdx.x] = s; sh[1][threadIdx.x] = c;
    __syncthreads();
    for (int off = kEB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) { sh[0][threadIdx.x] += sh[0][threadIdx.x + off]; sh[1][threadIdx.x] += sh[1][threadIdx.x + off]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[blockIdx.x] = sh[0][0]; part[kRedBlocks + blockIdx.x] = sh[1][0]; }
}

__global__ void scale_finish_kernel(double *red) {
    if (threadIdx.x != 0) return;
    double s = 0.0, c = 0.0;
    for (int b = 0; b < kRedBlocks; ++b) { s += red[b]; c += red[kRedBlocks + b]; }
    double sc = c > 0.0 ? sqrt(s / c) : 0.0;
    if (!(sc > 0.0)) sc = 1.0;
    red[kRedScale] = sc;
    red[kRedInvScale] = 1.0 / sc;
}

// input channels: rhs / scale, obstacle mask, 14 zeros
__global__ void unet_input_kernel(const double *rhs, const uint8_t *solid, const double *red, uint16_t *in0, size_t n) {
    const size_t i = (size_t)blockIdx.x * kEB + threadIdx.x;
    if (i >= n) return;
    const bool sol = solid && solid[i];
    const float b = sol ? 0.f : (float)(rhs[i] * red[kRedInvScale]);
    uint16_t *o = in0 + i * 16;
    o[0] = f2bf(b);
    o[1] = f2bf(sol ? 1.f : 0.f);
#pragma unroll
    for (int c = 2; c < 16; ++c) o[c] = 0;
}

// 2x2x2 average (fine resolution nx, ny, nz)
__global__ void pool_kernel(const uint16_t *in, uint16_t *out, int nx, int ny, int nz, int C) {
    const size_t i = (size_t)blockIdx.x * kEB + threadIdx.x;
    const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
    const size_t n = (size_t)cx * cy * cz * C;
    if (i >= n) return;
    const int c = (int)(i % C);
    const size_t v = i / C;
    const int z = (int)(v % cz), y = (int)((v / cz) % cy), x = (int)(v / ((size_t)cz * cy));
    float s = 0.f;
#pragma unroll
    for (int d = 0; d < 8; ++d)
        s += bf2f(in[vox(2 * x + (d >> 2), 2 * y + ((d >> 1) & 1), 2 * z + (d & 1), ny, nz) * C + c]);
    out[i] = f2bf(0.125f * s);
}

// gfine[v][c] += gcoarse[v / 2][c] / 8
__global__ void pool_bwd_add_kernel(const float *gc, float *gf, int nx, int ny, int nz, int C) {
    const size_t i = (size_t)blockIdx.x * kEB + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz * C;
    if (i >= n) return;
    const int c = (int)(i % C);
    const size_t v = i / C;
    const int z = (int)(v % nz), y = (int)((v / nz) % ny), x = (int)(v / ((size_t)nz * ny));
    gf[i] += 0.125f * gc[vox(x >> 1, y >> 1, z >> 1, ny / 2, nz / 2) * C + c];
}

// decoder input gradient [fine][cA + cB] -> upsampled part summed into gcoarse [coarse][cA]
__global__ void up_bwd_kernel(const float *gcat, float *gc, int nx, int ny, int nz, int cA, int cB) {
    const size_t i = (size_t)blockIdx.x * kEB + threadIdx.x;
    const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
    const size_t n = (size_t)cx * cy * cz * cA;
    if (i >= n) return;
    const int c = (int)(i % cA);
    const size_t v = i / cA;
    const int z = (int)(v % cz), y = (int)((v / cz) % cy), x = (int)(v / ((size_t)cz * cy));
    float s = 0.f;
#pragma unroll
    for (int d = 0; d < 8; ++d)
        s += gcat[vox(2 * x + (d >> 2), 2 * y + ((d >> 1) & 1), 2 * z + (d & 1), ny, nz) * (cA + cB) + c];
    gc[i] = s;
}

// ... and the skip part copied into gskip [fine][cB]
__global__ void skip_bwd_kernel(const float *gcat, float *gs, size_t nvox, int cA, int cB) {
    const size_t i = (size_t)blockIdx.x * kEB + threadIdx.x;
    if (i >= nvox * cB) return;
    const size_t v = i / cB;
    const int c = (int)(i % cB);
    gs[i] = gcat[v * (cA + cB) + cA + c];
}

// BF16 gradient of a pre-activation: g where the ReLU output is positive
__global__ void relu_bwd_kernel(const float *g, const uint16_t *act, uint16_t *dpre, size_t n) {
    const size_t i = (size_t)blockIdx.x * kEB + threadIdx.x;
    if (i >= n) return;
    dpre[i] = (act[i] & 0x7fffu) != 0 && !(act[i] & 0x8000u) ? f2bf(g[i]) : (uint16_t)0;
}

// bias gradient partials: part[chunk][c] = sum over the chunk's voxels of dpre[v][c]
__global__ void bias_part_kernel(const uint16_t *dpre, size_t nvox, int C, float *part) {
    const int c = threadIdx.x % C, lanes = blockDim.x / C, sub = threadIdx.x / C;
    __shared__ float sh[256];
    float s = 0.f;
    const size_t vb = (size_t)blockIdx.x * kWgChunk, ve = min(nvox, vb + (size_t)kWgChunk);
    if (sub < lanes)
        for (size_t v = vb + sub; v < ve; v += lanes) s += bf2f(dpre[v * C + c]);
    sh[threadIdx.x] = s;
    __syncthreads();
    if ((int)threadIdx.x < C) {
        float t = 0.f;
        for (int k = 0; k < lanes; ++k) t += sh[k * C + threadIdx.x];
        part[(size_t)blockIdx.x * C + threadIdx.x] = t;
    }
}

// 1x1 output layer: y = w . act + b
__global__ void out_fwd_kernel(const uint16_t *act, const float *wb, float *y, size_t n) {
    const size_t i = (size_t)blockIdx.x * kEB + threadIdx.x;
    if (i >= n) return;
    float s = wb[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) s += wb[c] * bf2f(act[i * 16 + c]);
    y[i] = s;
}

// gact = dy w; per-block partials of dy * act (16) and dy (1)
__global__ void __launch_bounds__(kEB) out_bwd_kernel(const float *dy, const uint16_t *act, const float *wb, float *gact,
                                                     size_t n, float *part) {
    __shared__ float sh[17][kEB];
    float acc[17];
#pragma unroll
    for (int c = 0; c < 17; ++c) acc[c] = 0.f;
    for (size_t i = (size_t)blockIdx.x * kEB + threadIdx.x; i < n; i += (size_t)gridDim.x * kEB) {
        const float d = dy[i];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            gact[i * 16 + c] = d * wb[c];
            acc[c] += d * bf2f(act[i * 16 + c]);
        }
        acc[16] += d;
    }
#pragma unroll
    for (int c = 0; c < 17; ++c) sh[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int off = kEB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
#pragma unroll
            for (int c = 0; c < 17; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + off];
        __syncthreads();
    }
    if ((int)threadIdx.x < 17) part[(size_t)blockIdx.x * 17 + threadIdx.x] = sh[threadIdx.x][0];
}

// Residual of the network output on the fluid cells: res = L y - rhs / scale (L: the Neumann
// Laplacian of the fluid cells, flag bits 0-5), and per-block sums of res^2, b^2, y.res, b.y
// (b = rhs / scale).
__global__ void __launch_bounds__(kEB) loss_res_kernel(const uint8_t *flags, const float *y, const double *rhs,
                                                      const double *red, int nx, int ny, int nz, float *res, double *part) {
    __shared__ double sh[4][kEB];
    const size_t n = (size_t)nx * ny * nz;
    const long sx = (long)ny * nz, sy = nz;
    const double inv = red[kRedInvScale];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t c = (size_t)blockIdx.x * kEB + threadIdx.x; c < n; c += (size_t)gridDim.x * kEB) {
        const unsigned f = flags[c];
        float r = 0.f;
        if (f & 63u) {
            double s = (double)__builtin_popcount(f & 63u) * y[c];
            if (f & 1u) s -= y[c - sx];
            if (f & 2u) s -= y[c + sx];
            if (f & 4u) s -= y[c - sy];
            if (f & 8u) s -= y[c + sy];
            if (f & 16u) s -= y[c - 1];
            if (f & 32u) s -= y[c + 1];
            const double b = rhs[c] * inv;
            r = (float)(s - b);
            a[0] += (double)r * r;
            a[1] += b * b;
            a[2] += (double)y[c] * r;
            a[3] += b * y[c];
        }
        res[c] = r;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int off = kEB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
#pragma unroll
            for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) part[k * kRedBlocks + blockIdx.x] = sh[k][0];
}

// mode 0 (residual norm): loss = |r|^2 / |b|^2, dy = (2 / |b|^2) L r
// mode 1 (energy): loss = (y.L y / 2 - b.y) / |b|^2 = (y.r - b.y) / (2 |b|^2), dy = r / |b|^2
__global__ void loss_finish_kernel(double *red, int mode) {
    if (threadIdx.x != 0) return;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < 4; ++k)
        for (int b = 0; b < kRedBlocks; ++b) a[k] += red[k * kRedBlocks + b];
    const double den = a[1] > 0.0 ? a[1] : 1.0;
    red[kRedResidual] = a[0] / den;
    red[kRedLossDen] = a[1];
    if (mode == 0) {
        red[kRedLoss] = a[0] / den;
        red[kRedLossCoef] = 2.0 / den;
    } else {
        red[kRedLoss] = 0.5 * (a[2] - a[3]) / den;
        red[kRedLossCoef] = 1.0 / den;
    }
}

// dy = coef L res (mode 0) or coef res (mode 1)
__global__ void loss_grad_kernel(const uint8_t *flags, const float *res, const double *red, int nx, int ny, int nz,
                                 int mode, float *dy) {
    const size_t c = (size_t)blockIdx.x * kEB + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (c >= n) return;
    const long sx = (long)ny * nz, sy = nz;
    const unsigned f = flags[c];
    float g = 0.f;
    if (f & 63u) {
        double s = res[c];
        if (mode == 0) {
            s *= (double)__builtin_popcount(f & 63u);
            if (f & 1u) s -= res[c - sx];
            if (f & 2u) s -= res[c + sx];
            if (f & 4u) s -= res[c - sy];
            if (f & 8u) s -= res[c + sy];
            if (f & 16u) s -= res[c - 1];
            if (f & 32u) s -= res[c + 1];
        }
        g = (float)(red[kRedLossCoef] * s);
    }
    dy[c] = g;
}

__global__ void adam_kernel(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2,
                            float eps, float bc1, float bc2, float gs) {
    const size_t i = (size_t)blockIdx.x * kEB + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] * gs;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// BF16 weights of conv at offset off: forward [t][co][ci] and flipped-transposed [26 - t][ci][co]
__global__ void pack_kernel(const float *p, uint16_t *wf, uint16_t *wt, size_t off, int cin, int cout) {
    const size_t i = (size_t)blockIdx.x * kEB + threadIdx.x;
    const size_t n = (size_t)27 * cin * cout;
    if (i >= n) return;
    const int ci = (int)(i % cin), co = (int)((i / cin) % cout), t = (int)(i / ((size_t)cin * cout));
    const uint16_t h = f2bf(p[off + i]);
    wf[off + i] = h;
    wt[off + ((size_t)(26 - t) * cin + ci) * cout + co] = h;
}

__global__ void fluid_flags_kernel(const uint8_t *solid, int nx, int ny, int nz, uint8_t *flags) {
    const size_t c = (size_t)blockIdx.x * kEB + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (c >= n) return;
    const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((size_t)ny * nz));
    auto fl = [&](int a, int b, int d) {
        return a >= 0 && b >= 0 && d >= 0 && a < nx && b < ny && d < nz && !(solid && solid[vox(a, b, d, ny, nz)]);
    };
    if (!fl(i, j, k)) { flags[c] = 0; return; }
    flags[c] = (uint8_t)((unsigned)fl(i - 1, j, k) | (unsigned)fl(i + 1, j, k) << 1 | (unsigned)fl(i, j - 1, k) << 2 |
                         (unsigned)fl(i, j + 1, k) << 3 | (unsigned)fl(i, j, k - 1) << 4 | (unsigned)fl(i, j, k + 1) << 5);
}

hipError_t launch_fluid_flags(const uint8_t *solid, int nx, int ny, int nz, uint8_t *flags, hipStream_t s) {
    const size_t n = (size_t)nx * ny * nz;
    hipLaunchKernelGGL(fluid_flags_kernel, dim3(eblocks(n)), dim3(kEB), 0, s, solid, nx, ny, nz, flags);
    return hipGetLastError();
}

__global__ void guess_kernel(const float *y, long pin, const uint8_t *pflags, const double *red, double *x, size_t n) {
    const size_t c = (size_t)blockIdx.x * kEB + threadIdx.x;
    if (c >= n) return;
    x[c] = (pflags[c] & kPcgUnknown) ? red[kRedScale] * ((double)y[c] - (double)y[pin]) : 0.0;
}

__global__ void fill_f32_kernel(float *a, float v, size_t n) {
    const size_t i = (size_t)blockIdx.x * kEB + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---- host side --------------------------------------------------------------------------------

static size_t level_vox(int nx, int ny, int nz, int level) {
    return (size_t)(nx >> level) * (ny >> level) * (nz >> level);
}

UNet::~UNet() {
    param_.release(); grad_.release(); m_.release(); v_.release(); wf_.release(); wt_.release(); in0_.release();
    for (int i = 0; i < kUNetConvs; ++i) { act_[i].release(); gact_[i].release(); }
    for (int i = 0; i < 3; ++i) { pool_[i].release(); gpool_[i].release(); }
    gcat_.release(); dpre_.release(); wpart_.release(); red_.release();
    if (host_red_) (void)hipHostFree(host_red_);
}

size_t UNet::device_bytes() const {
    size_t b = param_.bytes() + grad_.bytes() + m_.bytes() + v_.bytes() + wf_.bytes() + wt_.bytes() + in0_.bytes() +
               gcat_.bytes() + dpre_.bytes() + wpart_.bytes() + red_.bytes();
    for (int i = 0; i < kUNetConvs; ++i) b += act_[i].bytes() + gact_[i].bytes();
    for (int i = 0; i < 3; ++i) b += pool_[i].bytes() + gpool_[i].bytes();
    return b;
}

hipError_t UNet::alloc_params() {
    const size_t n = unet_param_count();
    hipError_t e;
    if ((e = param_.ensure(n)) != hipSuccess) return e;
    if ((e = grad_.ensure(n)) != hipSuccess) return e;
    if ((e = m_.ensure(n)) != hipSuccess) return e;
    if ((e = v_.ensure(n)) != hipSuccess) return e;
    if ((e = wf_.ensure(n)) != hipSuccess) return e;
    if ((e = wt_.ensure(n)) != hipSuccess) return e;
    if ((e = red_.ensure(kRedSlots)) != hipSuccess) return e;
    if (!host_red_ && (e = hipHostMalloc((void **)&host_red_, kRedSlots * sizeof(double))) != hipSuccess) return e;
    return hipSuccess;
}

// splitmix64 + Box-Muller: the same numbers on every host
static double unit_uniform(unsigned long long &st) {
    st += 0x9E3779B97F4A7C15ULL;
    unsigned long long z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

hipError_t UNet::init(unsigned long long seed, hipStream_t s) {
    std::vector<float> h(unet_param_count(), 0.f);
    unsigned long long st = seed;
    auto normal = [&]() {
        const double u1 = unit_uniform(st), u2 = unit_uniform(st);
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    };
    for (int i = 0; i < kUNetConvs; ++i) {
        const UNetConv &L = kUNetTable[i];
        const double sd = sqrt(2.0 / (27.0 * L.cin_real));
        const size_t off = unet_weight_offset(i);
        for (int t = 0; t < 27; ++t)
            for (int co = 0; co < L.cout; ++co)
                for (int ci = 0; ci < L.cin; ++ci)
                    h[off + ((size_t)t * L.cout + co) * L.cin + ci] = ci < L.cin_real ? (float)(sd * normal()) : 0.f;
    }
    const size_t o = unet_out_offset();
    for (int c = 0; c < 16; ++c) h[o + c] = (float)(0.25 * normal());
    return set_params(h.data(), s);
}

hipError_t UNet::set_params(const float *host, hipStream_t s) {
    hipError_t e;
    if ((e = alloc_params()) != hipSuccess) return e;
    const size_t n = unet_param_count();
    if ((e = hipMemcpyAsync(param_.p, host, n * sizeof(float), hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(fill_f32_kernel, dim3(eblocks(n)), dim3(kEB), 0, s, m_.p, 0.f, n);
    hipLaunchKernelGGL(fill_f32_kernel, dim3(eblocks(n)), dim3(kEB), 0, s, v_.p, 0.f, n);
    hipLaunchKernelGGL(fill_f32_kernel, dim3(eblocks(n)), dim3(kEB), 0, s, grad_.p, 0.f, n);
    adam_t_ = 0;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return pack(s);
}

hipError_t UNet::get_params(float *host, hipStream_t s) {
    if (!param_.p) return hipErrorInvalidValue;
    hipError_t e = hipMemcpyAsync(host, param_.p, unet_param_count() * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

hipError_t UNet::pack(hipStream_t s) {
    for (int i = 0; i < kUNetConvs; ++i) {
        const UNetConv &L = kUNetTable[i];
        const size_t n = (size_t)27 * L.cin * L.cout;
        hipLaunchKernelGGL(pack_kernel, dim3(eblocks(n)), dim3(kEB), 0, s, param_.p, wf_.p, wt_.p, unet_weight_offset(i),
                           L.cin, L.cout);
    }
    return hipGetLastError();
}

hipError_t UNet::zero_grad(hipStream_t s) {
    const size_t n = unet_param_count();
    hipLaunchKernelGGL(fill_f32_kernel, dim3(eblocks(n)), dim3(kEB), 0, s, grad_.p, 0.f, n);
    return hipGetLastError();
}

hipError_t UNet::ensure_grid(int nx, int ny, int nz) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx % 8 || ny % 8 || nz % 8) return hipErrorInvalidValue;
    nx_ = nx; ny_ = ny; nz_ = nz;
    for (int l = 0; l < 4; ++l) vox_[l] = level_vox(nx, ny, nz, l);
    hipError_t e;
    if ((e = in0_.ensure(vox_[0] * 16)) != hipSuccess) return e;
    size_t gcat = 0, dpre = 0, part = 0;
    for (int i = 0; i < kUNetConvs; ++i) {
        const UNetConv &L = kUNetTable[i];
        const size_t nv = vox_[L.level];
        if ((e = act_[i].ensure(nv * L.cout)) != hipSuccess) return e;
        if ((e = gact_[i].ensure(nv * L.cout)) != hipSuccess) return e;
        gcat = std::max(gcat, nv * L.cin);
        dpre = std::max(dpre, nv * L.cout);
        part = std::max(part, conv3_wgrad_chunks(nv) * 27 * L.cin * L.cout);
        part = std::max(part, conv3_wgrad_chunks(nv) * L.cout);
    }
    part = std::max(part, (size_t)kRedBlocks * 17);
    for (int p = 0; p < 3; ++p) {
        const int C = kUNetTable[2 * p + 1].cout;
        if ((e = pool_[p].ensure(vox_[p + 1] * C)) != hipSuccess) return e;
        if ((e = gpool_[p].ensure(vox_[p + 1] * C)) != hipSuccess) return e;
    }
    if ((e = gcat_.ensure(gcat)) != hipSuccess) return e;
    if ((e = dpre_.ensure(dpre)) != hipSuccess) return e;
    if ((e = wpart_.ensure(part)) != hipSuccess) return e;
    return hipSuccess;
}

// the input of conv i in the forward pass (and in its weight gradient)
ConvArgs UNet::conv_in(int i) const {
    const UNetConv &L = kUNetTable[i];
    ConvArgs a;
    memset(&a, 0, sizeof(a));
    a.nx = nx_ >> L.level; a.ny = ny_ >> L.level; a.nz = nz_ >> L.level;
    a.cout = L.cout;
    a.w = wf_.p + unet_weight_offset(i);
    a.bias = param_.p + unet_bias_offset(i);
    a.relu = 1;
    a.out = act_[i].p;
    switch (i) {
    case 0: a.srcA = in0_.p; a.cA = 16; break;
    case 2: a.srcA = pool_[0].p; a.cA = 16; break;
    case 4: a.srcA = pool_[1].p; a.cA = 32; break;
    case 6: a.srcA = pool_[2].p; a.cA = 48; break;
    case 8: a.srcA = act_[7].p; a.cA = 64; a.upA = 1; a.srcB = act_[5].p; a.cB = 48; break;
    case 10: a.srcA = act_[9].p; a.cA = 48; a.upA = 1; a.srcB = act_[3].p; a.cB = 32; break;
    case 12: a.srcA = act_[11].p; a.cA = 32; a.upA = 1; a.srcB = act_[1].p; a.cB = 16; break;
    default: a.srcA = act_[i - 1].p; a.cA = kUNetTable[i - 1].cout; break;
    }
    return a;
}

hipError_t UNet::forward(const double *rhs, const uint8_t *solid, float *y, double *scale, hipStream_t s) {
    if (!nx_ || !param_.p) return hipErrorInvalidValue;
    const size_t n = vox_[0];
    hipLaunchKernelGGL(rhs_sumsq_kernel, dim3(kRedBlocks), dim3(kEB), 0, s, rhs, solid, n, red_.p);
    hipLaunchKernelGGL(scale_finish_kernel, dim3(1), dim3(64), 0, s, red_.p);
    hipLaunchKernelGGL(unet_input_kernel, dim3(eblocks(n)), dim3(kEB), 0, s, rhs, solid, red_.p, in0_.p, n);
    hipError_t e;
    for (int i = 0; i < kUNetConvs; ++i) {
        if ((e = launch_conv3(conv_in(i), 0, s)) != hipSuccess) return e;
        if (i == 1 || i == 3 || i == 5) {
            const int p = i / 2, lv = kUNetTable[i].level, C = kUNetTable[i].cout;
            const size_t m = vox_[lv + 1] * C;
            hipLaunchKernelGGL(pool_kernel, dim3(eblocks(m)), dim3(kEB), 0, s, act_[i].p, pool_[p].p, nx_ >> lv,
                               ny_ >> lv, nz_ >> lv, C);
        }
    }
    hipLaunchKernelGGL(out_fwd_kernel, dim3(eblocks(n)), dim3(kEB), 0, s, act_[13].p, param_.p + unet_out_offset(), y, n);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (scale) {
        if ((e = hipMemcpyAsync(host_red_ + kRedScale, red_.p + kRedScale, sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess)
            return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
        *scale = host_red_[kRedScale];
    }
    return hipSuccess;
}

hipError_t UNet::loss(int mode, const uint8_t *flags, const double *rhs, const float *y, float *dy, double *loss,
                      double *residual, hipStream_t s) {
    if (mode != 0 && mode != 1) return hipErrorInvalidValue;
    const size_t n = vox_[0];
    // the residual goes into the full-resolution gradient buffer of d0a, free at this point
    float *res = gact_[12].p;
    hipLaunchKernelGGL(loss_res_kernel, dim3(kRedBlocks), dim3(kEB), 0, s, flags, y, rhs, red_.p, nx_, ny_, nz_, res, red_.p);
    hipLaunchKernelGGL(loss_finish_kernel, dim3(1), dim3(64), 0, s, red_.p, mode);
    if (dy)
        hipLaunchKernelGGL(loss_grad_kernel, dim3(eblocks(n)), dim3(kEB), 0, s, flags, res, red_.p, nx_, ny_, nz_, mode,
                           dy);
    hipError_t e;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (loss || residual) {
        if ((e = hipMemcpyAsync(host_red_ + kRedLoss, red_.p + kRedLoss, (kRedResidual - kRedLoss + 1) * sizeof(double),
                                hipMemcpyDeviceToHost, s)) != hipSuccess)
            return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
        if (loss) *loss = host_red_[kRedLoss];
        if (residual) *residual = host_red_[kRedResidual];
    }
    return hipSuccess;
}

// out-layer parameter gradients: grad[out + c] += sum of the partials
__global__ void out_grad_finish_kernel(const float *part, int blocks, float *g) {
    const int c = threadIdx.x;
    if (c >= 17) return;
    float s = 0.f;
    for (int b = 0; b < blocks; ++b) s += part[(size_t)b * 17 + c];
    g[c] += s;
}

hipError_t UNet::backward(const float *dy, hipStream_t s) {
    if (!nx_) return hipErrorInvalidValue;
    const size_t n = vox_[0];
    hipError_t e;
    // output layer
    hipLaunchKernelGGL(out_bwd_kernel, dim3(kRedBlocks), dim3(kEB), 0, s, dy, act_[13].p, param_.p + unet_out_offset(),
                       gact_[13].p, n, wpart_.p);
    hipLaunchKernelGGL(out_grad_finish_kernel, dim3(1), dim3(64), 0, s, wpart_.p, kRedBlocks, grad_.p + unet_out_offset());
    for (int i = kUNetConvs - 1; i >= 0; --i) {
        const UNetConv &L = kUNetTable[i];
        const size_t nv = vox_[L.level];
        const int lx = nx_ >> L.level, ly = ny_ >> L.level, lz = nz_ >> L.level;
        hipLaunchKernelGGL(relu_bwd_kernel, dim3(eblocks(nv * L.cout)), dim3(kEB), 0, s, gact_[i].p, act_[i].p, dpre_.p,
                           nv * L.cout);
        const ConvArgs in = conv_in(i);
        if ((e = launch_conv3_wgrad(in, dpre_.p, wpart_.p, grad_.p + unet_weight_offset(i), s)) != hipSuccess) return e;
        const size_t chunks = conv3_wgrad_chunks(nv);
        const int threads = std::min(256, (256 / L.cout) * L.cout);
        hipLaunchKernelGGL(bias_part_kernel, dim3((unsigned)chunks), dim3(threads), 0, s, dpre_.p, nv, L.cout, wpart_.p);
        hipLaunchKernelGGL(reduce_chunks_kernel, dim3(1), dim3(256), 0, s, wpart_.p, chunks, (size_t)L.cout,
                           grad_.p + unet_bias_offset(i));
        if (i == 0) break;
        // data gradient: the same convolution on the flipped, transposed weights
        ConvArgs d;
        memset(&d, 0, sizeof(d));
        d.nx = lx; d.ny = ly; d.nz = lz;
        d.srcA = dpre_.p; d.cA = L.cout;
        d.w = wt_.p + unet_weight_offset(i);
        d.cout = L.cin;
        const bool cat = (i == 8 || i == 10 || i == 12), pooled = (i == 2 || i == 4 || i == 6);
        d.out = cat ? (void *)gcat_.p : pooled ? (void *)gpool_[i / 2 - 1].p : (void *)gact_[i - 1].p;
        if ((e = launch_conv3(d, 1, s)) != hipSuccess) return e;
        if (cat) {
            const int up = i - 1, skip = (i == 8) ? 5 : (i == 10) ? 3 : 1;
            const size_t nc = vox_[L.level + 1] * in.cA;
            hipLaunchKernelGGL(up_bwd_kernel, dim3(eblocks(nc)), dim3(kEB), 0, s, gcat_.p, gact_[up].p, lx, ly, lz, in.cA,
                               in.cB);
            hipLaunchKernelGGL(skip_bwd_kernel, dim3(eblocks(nv * in.cB)), dim3(kEB), 0, s, gcat_.p, gact_[skip].p, nv,
                               in.cA, in.cB);
        } else if (pooled) {
            const int skip = i - 1, fl = L.level - 1, C = kUNetTable[skip].cout;
            const size_t nf = vox_[fl] * C;
            hipLaunchKernelGGL(pool_bwd_add_kernel, dim3(eblocks(nf)), dim3(kEB), 0, s, gpool_[i / 2 - 1].p, gact_[skip].p,
                               nx_ >> fl, ny_ >> fl, nz_ >> fl, C);
        }
    }
    return hipGetLastError();
}

hipError_t UNet::guess(const float *y, long pin, const uint8_t *pcg_flags, double *x, hipStream_t s) {
    if (!nx_ || pin < 0 || (size_t)pin >= vox_[0]) return hipErrorInvalidValue;
    hipLaunchKernelGGL(guess_kernel, dim3(eblocks(vox_[0])), dim3(kEB), 0, s, y, pin, pcg_flags, red_.p, x, vox_[0]);
    return hipGetLastError();
}

hipError_t UNet::adam_step(float lr, float beta1, float beta2, float eps, float grad_scale, hipStream_t s) {
    ++adam_t_;
    const float bc1 = 1.f - powf(beta1, (float)adam_t_), bc2 = 1.f - powf(beta2, (float)adam_t_);
    const size_t n = unet_param_count();
    hipLaunchKernelGGL(adam_kernel, dim3(eblocks(n)), dim3(kEB), 0, s, param_.p, grad_.p, m_.p, v_.p, n, lr, beta1, beta2,
                       eps, bc1, bc2, grad_scale);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return pack(s);
}

}  // namespace fia
