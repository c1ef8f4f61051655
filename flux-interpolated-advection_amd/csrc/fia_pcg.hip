// fia_pcg.hip - pressure projection on gfx950: Jacobi-preconditioned CG on the reference's
// 7-point Neumann Laplacian (reference src/calc_pressure_3d.cpp:18-180), FP64 vectors, FP64
// dot products reduced in a fixed order (deterministic), and the velocity update
// u -= grad p / h (reference src/calc_pressure_3d.cpp:196-234).
//
// System: L p = b with L = -A (A as assembled by the reference: diag -6 + #walls, +1 per
// neighbour), b = -div(u) h, the last cell pinned to p = 0.  Obstacle cells (optional mask) are
// removed from the system and their faces act as walls.
//
// One CG iteration is two kernels:
//   direction: beta from the previous residual, p = z + beta p_old (z = r / diag, not stored),
//              q = L p (2.5-D tiled stencil sweep), partial p.q;
//   update:    alpha = rz / p.q, x += alpha p, r -= alpha q, partial r.z and r.r.
// The update has two forms, chosen per solve from the grid's tiling (pcg_direction_tiling):
//   streaming: the direction kernel stores q and the update reads it, grid-stride over the cells;
//   recompute: the update sweeps the direction kernel's tiles and recomputes q = L p from p.
// p alternates between two buffers (a thread reads its neighbours' old p while others write the
// new one).  x and r are FP64; the search direction p is stored in FP32: the direction kernel
// rounds p = z + beta p_old before its stencil, so q = L p and p.q are exactly those of the stored
// p, x += alpha p and r -= alpha q use that same p, and the recurrence keeps r = b - L x.  The
// matrix is not stored: one flag byte per cell holds its fluid neighbours.  Per cell and iteration
// the direction kernel reads 8 (r) + 4 (p old) + 1 (flags) and writes 4 (p), plus 8 (q) when
// streaming; the update reads 4 (p) + 8 x 2 (x, r) + 1, plus 8 (q) when streaming, and writes 16
// (x, r): 70 bytes streaming, 54 recomputing.  Recomputing pays where the tiles sweep
// long x chunks (large grids); on small grids, whose chunks are a few planes, the halo planes
// cost more than q and the streaming form is faster.  Under Jacobi on grids of at least 4 M
// cells one kernel per iteration replaces the two (pcg_fused_kernel, below: 35 bytes per cell
// and iteration), and a geometric multigrid V-cycle can replace the Jacobi preconditioner.
//
// Reductions without fences: every block writes its partial sums with plain stores, and every
// block of the NEXT kernel sums all partials itself, in block order, before its own work (the
// kernel boundary makes the partials visible; the sum is the same in every block and does not
// depend on scheduling).  The CG scalars are double-buffered by iteration parity, so no block
// reads a scalar that another block of the same kernel writes.  Once converged every later kernel
// of the solve returns at once, so the host only polls now and then (after a 1-block check).
#include <atomic>

#include "fia_internal.hpp"
#include "fia_math.hpp"

namespace fia {

constexpr int kPB = 256;
constexpr int kTileZ = 64, kTileY = 16, kRows = 4;   // column tile of the direction kernel, rows per thread
// The update recomputes q rather than streaming it on grids of at least kRecomputeCells cells
// whose tiles sweep at least kRecomputeChunk planes.  Measured per iteration, streaming vs
// recomputing: 64^3 20 vs 24 us (chunks of 2) and 31 us (chunks forced to 4), 96^3 27 vs 33 us,
// 128^3 43 vs 50 us (chunks of 2) and 42 us (chunks of 4); 256^3 296 vs 254 us (chunks of 16).
constexpr int kRecomputeChunk = 4;
constexpr size_t kRecomputeCells = (size_t)1 << 22;

__device__ __forceinline__ bool fluid(const PcgGrid &g, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) return false;
    return !(g.solid && g.solid[idx3(i, j, k, g.ny, g.nz)]);
}

// 1 / (number of fluid neighbours), 0 for a cell outside the system
__device__ __forceinline__ double inv_diag(unsigned f) {
    if (!(f & kPcgUnknown)) return 0.0;
    const int d = __builtin_popcount(f & 63u);
    return d == 6 ? 1.0 / 6.0 : d == 5 ? 0.2 : d == 4 ? 0.25 : d == 3 ? 1.0 / 3.0 : d == 2 ? 0.5 : d == 1 ? 1.0 : 0.0;
}

__global__ void pcg_setup_kernel(PcgGrid g, const float *u, const float *v, const float *w, double h, double *b,
                                 uint8_t *flags) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)g.nx * g.ny * g.nz;
    if (c >= n) return;
    const int k = (int)(c % g.nz), j = (int)((c / g.nz) % g.ny), i = (int)(c / ((long)g.ny * g.nz));
    if (!fluid(g, i, j, k)) { b[c] = 0.0; flags[c] = 0; return; }
    unsigned f = (unsigned)fluid(g, i - 1, j, k) | (unsigned)fluid(g, i + 1, j, k) << 1 |
                 (unsigned)fluid(g, i, j - 1, k) << 2 | (unsigned)fluid(g, i, j + 1, k) << 3 |
                 (unsigned)fluid(g, i, j, k - 1) << 4 | (unsigned)fluid(g, i, j, k + 1) << 5;
    if (c == g.pin || (f & 63u) == 0u) { b[c] = 0.0; flags[c] = (uint8_t)f; return; }
    const int ny = g.ny, nz = g.nz;
    const double div = ((double)u[idx3(i + 1, j, k, ny, nz)] - (double)u[idx3(i, j, k, ny, nz)])
                     + ((double)v[idx3(i, j + 1, k, ny + 1, nz)] - (double)v[idx3(i, j, k, ny + 1, nz)])
                     + ((double)w[idx3(i, j, k + 1, ny, nz + 1)] - (double)w[idx3(i, j, k, ny, nz + 1)]);
    b[c] = -div * h;
    flags[c] = (uint8_t)(f | kPcgUnknown);
}

// Sums NV values over the block in a fixed order: a shuffle tree in each wave, then the waves'
// sums in order.  Every thread gets the same totals, and every block sums in the same order.
template <int NV>
__device__ __forceinline__ void block_total(double (&v)[NV]) {
    constexpr int kWaves = kPB / 64;
    __shared__ double sh[NV][kWaves];
#pragma unroll
    for (int a = 0; a < NV; ++a)
        for (int off = 32; off > 0; off >>= 1) v[a] += __shfl_down(v[a], off, 64);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < NV; ++a) sh[a][wave] = v[a];
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < NV; ++a) {
        double t = sh[a][0];
        for (int w = 1; w < kWaves; ++w) t += sh[a][w];
        v[a] = t;
    }
    __syncthreads();
}

// Writes block blk's sums of v into part[a * nb + blk].
template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double *part, int nb, int blk) {
    block_total<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < NV; ++a) part[(size_t)a * nb + blk] = v[a];
    }
}

// Totals of partials written by an earlier kernel, identical in every thread of every block.
// Every block runs this before its own work, so up to kPcgMaxBlocks partials its loads are all
// issued at once: each thread takes the partials threadIdx.x + j kPB in order of j from clamped
// indices, and one past nb counts as 0 through a multiply by 0 (a select next to the load, or a
// loop to nb, became one round trip per partial).  The sums are those of the loop that larger
// counts take: x * 1 = x, and s + (+-0) = s for the finite partials.
template <int NV>
__device__ __forceinline__ void total_of(const double *part, int nb, double (&tot)[NV]) {
    if (nb <= kPcgMaxBlocks) {
        constexpr int kPer = kPcgMaxBlocks / kPB;
        double v[kPer][NV];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int i = min((int)threadIdx.x + j * kPB, nb - 1);
#pragma unroll
            for (int a = 0; a < NV; ++a) v[j][a] = part[(size_t)a * nb + i];
        }
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < kPer; ++j) s += v[j][a] * (double)((int)threadIdx.x + j * kPB < nb);
            tot[a] = s;
        }
    } else {
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            double s = 0.0;
            for (int i = threadIdx.x; i < nb; i += kPB) s += part[(size_t)a * nb + i];
            tot[a] = s;
        }
    }
    block_total<NV>(tot);
}

// The converged flag, read by one thread and shared, so that every thread of the block takes the
// same branch even when block 0 of the same kernel is setting the flag.
__device__ __forceinline__ bool solve_done(const double *scal) {
    __shared__ int flag;
    if (threadIdx.x == 0) flag = scal[kPcgDone] != 0.0;
    __syncthreads();
    return flag != 0;
}
// the same with a flag value each thread has loaded: thread 0's decides
__device__ __forceinline__ bool block_flag(double v) {
    __shared__ int flag;
    if (threadIdx.x == 0) flag = v != 0.0;
    __syncthreads();
    return flag != 0;
}

// The residual state after k iterations: rz, rr (and bb) from the partials of the start kernel
// (k = 0) or of update kernel k - 1.
__device__ __forceinline__ void residual_state(const PcgState &st, int k, double &rz, double &rr, double &bb) {
    if (k == 0) {
        double tot[3];
        total_of<3>(st.part_start, st.nblocks, tot);
        rz = tot[0]; rr = tot[1]; bb = tot[2];
    } else if (st.fused) {   // the sums of sweep k - 1 (its residual is iteration k - 1's)
        double tot[5];
        total_of<5>(st.part_fused + (size_t)((k - 1) & 1) * 5 * st.dir_blocks, st.dir_blocks, tot);
        rz = tot[0]; rr = tot[1]; bb = st.scal[kPcgBB];
    } else {
        double tot[2];
        total_of<2>(st.part_update, st.upd_blocks, tot);
        rz = tot[0]; rr = tot[1]; bb = st.scal[kPcgBB];
    }
    if (st.z) {   // multigrid: r.z of the last V-cycle
        double t[1];
        total_of<1>(st.part_mg, st.dir_blocks, t);
        rz = t[0];
    }
}

// Cells of block blk of a solve.  Workgroups are dispatched to the 8 XCDs round-robin, so block b
// runs on XCD b % 8 (for a solve launched alone, blk is the block index); each XCD gets one contiguous eighth of the cells (x-slabs), which its blocks sweep
// together, so the y and x neighbours a block reads were fetched into the same XCD's L2 by its
// sibling blocks.  nblocks is a multiple of 8.
struct CellRange {
    size_t first, end, step;
};
__device__ __forceinline__ CellRange block_cells(size_t n, int nblocks, int blk) {
    const int xcd = blk % 8, slot = blk / 8, per = nblocks / 8;
    const size_t chunk = (n + 7) / 8;
    const size_t lo = min(n, (size_t)xcd * chunk), hi = min(n, lo + chunk);
    return {lo + (size_t)slot * kPB + threadIdx.x, hi, (size_t)per * kPB};
}

// The 7-point stencil sum diag * v(c) - sum of v over the fluid neighbours, for a cell with
// flag byte f.  Every load is issued unconditionally (a missing neighbour reads the cell itself
// and is then dropped by a select), so the seven loads of a cell are in flight together.
template <class F>
__device__ __forceinline__ double stencil(unsigned f, size_t c, long sx, long sy, F val) {
    const double v0 = val(c);
    const double v1 = val((f & 1u) ? c - sx : c), v2 = val((f & 2u) ? c + sx : c);
    const double v3 = val((f & 4u) ? c - sy : c), v4 = val((f & 8u) ? c + sy : c);
    const double v5 = val((f & 16u) ? c - 1 : c), v6 = val((f & 32u) ? c + 1 : c);
    double s = (double)__builtin_popcount(f & 63u) * v0;
    s -= (f & 1u) ? v1 : 0.0;
    s -= (f & 2u) ? v2 : 0.0;
    s -= (f & 4u) ? v3 : 0.0;
    s -= (f & 8u) ? v4 : 0.0;
    s -= (f & 16u) ? v5 : 0.0;
    s -= (f & 32u) ? v6 : 0.0;
    return s;
}

// r = b - L x, p_old = 0, and partial rz (Jacobi), rr, bb; block 0 resets the tolerance and the flags
__device__ __forceinline__ void pcg_start_body(const PcgState &st, int blk, double tol2) {
    const size_t n = (size_t)st.g.nx * st.g.ny * st.g.nz;
    const long sx = (long)st.g.ny * st.g.nz, sy = st.g.nz;
    const double *x = st.x;
    double acc[3] = {0.0, 0.0, 0.0};
    const CellRange cr = block_cells(n, st.nblocks, blk);
    for (size_t c = cr.first; c < cr.end; c += cr.step) {
        const unsigned f = st.flags[c];
        const double bc = st.b[c];
        const double s = stencil(f, c, sx, sy, [&](size_t i) { return x[i]; });
        const double rv = (f & kPcgUnknown) ? bc - s : 0.0;
        st.r[c] = rv;
        st.p[0][c] = 0.0f;
        acc[0] += rv * (rv * inv_diag(f));
        acc[1] += rv * rv;
        acc[2] += bc * bc;
    }
    write_partials<3>(acc, st.part_start, st.nblocks, blk);
    if (blk == 0 && threadIdx.x == 0) {
        double *sc = st.scal;
        sc[kPcgTol2] = tol2;
        sc[kPcgDone] = 0.0;
        sc[kPcgIters] = 0.0;
        sc[kPcgNext] = 0.0;
    }
}

// Convergence test after k iterations (one block): the state the host polls.
__device__ __forceinline__ void pcg_check_body(const PcgState &st, int k) {
    if (solve_done(st.scal)) return;
    double rz, rr, bb;
    residual_state(st, k, rz, rr, bb);
    if (threadIdx.x == 0) {
        double *sc = st.scal;
        sc[kPcgRR] = rr;
        if (k == 0) sc[kPcgBB] = bb;
        sc[kPcgIters] = (double)(st.fused && k > 0 ? k - 1 : k);
        if (rr <= sc[kPcgTol2] * bb) sc[kPcgDone] = 1.0;
    }
}

// Both kernels of an iteration sweep the grid with one 2.5-D stencil: a block owns a tile of
// kTileZ x kTileY cell columns (each thread kRows of them, kTileY / kRows rows apart) and marches
// along x over its chunk of planes.  The value v of every cell (and of a one-cell halo around the
// tile) is computed once per plane: the x neighbours stay in registers, the y and z neighbours are
// read from the plane in LDS.  The raw data of the plane after next is loaded one step ahead.
// For each cell of the tile the operation receives v and q = L v (0 on cells that are not
// unknowns: the pinned cell has fluid neighbours).
//
// Iteration k, direction kernel: converged test on the residual of iteration k - 1,
// beta = rz_k / rz_(k-1), v = p = z + beta p_old (outside the system r = p_old = 0, so p = 0
// there), written out, and partial p.q.
// Update kernel: v = p read back, q recomputed with the same arithmetic (the same bits as the
// direction kernel's q, 16 bytes per cell cheaper than storing and reloading it),
// x += alpha p, r -= alpha q (alpha = rz / p.q), partial rz, rr.
// Raw records hold the loaded values as stored (FP32 p included): converting at the load would
// make the sweep wait for the plane it prefetches two steps ahead.
// Every load of a sweep is issued unconditionally, from an address clamped into the grid, and a
// select zeroes what lies outside (also for the planes past the chunk and for threads without a
// halo cell, whose values are not used): a load under a branch that all lanes may skip leaves the
// compiler unable to count the loads issued after an earlier one, and it then waits for all of
// them (vmcnt(0)), which undoes the prefetch.
__device__ __forceinline__ bool in_grid(const PcgGrid &g, int x, int y, int z) {
    return x >= 0 && x < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz;
}
__device__ __forceinline__ size_t clamped_cell(const PcgGrid &g, int x, int y, int z) {
    return idx3(min(max(x, 0), g.nx - 1), min(max(y, 0), g.ny - 1), min(max(z, 0), g.nz - 1), g.ny, g.nz);
}
struct DirRaw {
    unsigned f;
    double r;
    float p;
};
struct DirOp {
    using Raw = DirRaw;
    const PcgState &st;
    const float *__restrict__ pold;
    float *__restrict__ pnew;
    double beta;
    // Jacobi: r (z = r / diag on the fly); multigrid: the V-cycle's z
    __device__ __forceinline__ Raw load(int x, int y, int z, bool) const {
        const size_t c = clamped_cell(st.g, x, y, z);
        const double *rz = st.z ? st.z : st.r;
        const unsigned f = st.flags[c];
        const double r = rz[c];
        const float p = pold[c];
        const bool in = in_grid(st.g, x, y, z);
        return Raw{in ? f : 0u, in ? r : 0.0, in ? p : 0.0f};
    }
    // the direction rounded to FP32 here, before the stencil: q = L p and p.q are those of the
    // stored p, which the update reads back
    __device__ __forceinline__ double value(const Raw &a) const {
        return (double)(float)((st.z ? a.r : a.r * inv_diag(a.f)) + beta * (double)a.p);
    }
    __device__ __forceinline__ void cell(size_t c, const Raw &, double v, double q, double *acc) const {
        pnew[c] = (float)v;
        if (st.q) st.q[c] = q;
        acc[0] += v * q;
    }
};
struct UpdRaw {
    unsigned f;
    float p;
    double x, r;
};
struct UpdOp {
    using Raw = UpdRaw;
    const PcgState &st;
    const float *__restrict__ p;
    double alpha;
    // x and r only for the tile's own cells (own), not for the halo
    __device__ __forceinline__ Raw load(int x, int y, int z, bool own) const {
        const size_t c = clamped_cell(st.g, x, y, z);
        const bool in = in_grid(st.g, x, y, z);
        const unsigned f = st.flags[c];
        const float pv = p[c];
        Raw a{in ? f : 0u, in ? pv : 0.0f, 0.0, 0.0};
        if (own) {   // a constant at every call
            const double xv = st.x[c], rv = st.r[c];
            a.x = in ? xv : 0.0;
            a.r = in ? rv : 0.0;
        }
        return a;
    }
    __device__ __forceinline__ double value(const Raw &a) const { return (double)a.p; }
    __device__ __forceinline__ void cell(size_t c, const Raw &a, double v, double q, double *acc) const {
        const double rv = a.r - alpha * q;
        st.x[c] = a.x + alpha * v;
        st.r[c] = rv;
        acc[0] += rv * (rv * inv_diag(a.f));
        acc[1] += rv * rv;
    }
};

// the halo cell a thread loads (2 kTileZ y-halo cells, then 2 kTileY z-halo cells), or false
__device__ __forceinline__ bool halo_cell(int t, int y0, int z0, int &y, int &z, int &ly, int &lz) {
    if (t < kTileZ) { y = y0 - 1; z = z0 + t; ly = 0; lz = t + 1; return true; }
    if (t < 2 * kTileZ) { y = y0 + kTileY; z = z0 + t - kTileZ; ly = kTileY + 1; lz = t - kTileZ + 1; return true; }
    if (t < 2 * kTileZ + kTileY) { y = y0 + t - 2 * kTileZ; z = z0 - 1; ly = y - y0 + 1; lz = 0; return true; }
    if (t < 2 * kTileZ + 2 * kTileY) { y = y0 + t - 2 * kTileZ - kTileY; z = z0 + kTileZ; ly = y - y0 + 1; lz = kTileZ + 1; return true; }
    return false;
}

// An Op with kSelectReads = true keeps the plane reads of tile_sweep under selects of the value
// (fewer registers, more branches); the others read a zero slot for a missing neighbour.
template <class Op> constexpr bool select_reads(long) { return false; }
template <class Op> constexpr auto select_reads(int) -> decltype((void)Op::kSelectReads, bool()) { return Op::kSelectReads; }

// Tile blk of a solve's sweep (dir_blocks tiles; blocks are dispatched to the 8 XCDs round-robin,
// so consecutive logical tiles go to one XCD, which sweeps whole y-z slabs of neighbouring tiles).
template <class Op>
__device__ __forceinline__ void tile_sweep(const PcgState &st, int blk, const Op &op, double *acc) {
    static_assert(kTileZ * kTileY == kPB * kRows && 2 * (kTileZ + kTileY) <= kPB, "tile shape");
    using Raw = typename Op::Raw;
    __shared__ double plane[2][kTileY + 2][kTileZ + 2];
    // +0.0, read in place of a neighbour that is not fluid: the plane reads take their address from
    // the flag bit instead of sitting under a select of the value, which the compiler turns into a
    // branch with its own wait per read
    __shared__ double zero_slot;
    if (threadIdx.x == 0) zero_slot = 0.0;
    const int per = st.dir_blocks / 8;
    const int L = (blk % 8) * per + blk / 8;
    constexpr int kRowStep = kTileY / kRows;
    const int ty = threadIdx.x / kTileZ, tz = threadIdx.x % kTileZ;
    if (L >= st.dir_tiles_y * st.dir_tiles_z * st.dir_chunks) return;
    const int y0 = (L % st.dir_tiles_y) * kTileY;
    const int z0 = ((L / st.dir_tiles_y) % st.dir_tiles_z) * kTileZ;
    const int xa = (L / (st.dir_tiles_y * st.dir_tiles_z)) * st.dir_chunk_x;
    const int xb = min(st.g.nx, xa + st.dir_chunk_x);
    const int z = z0 + tz;
    int hy = 0, hz = 0, hly = 0, hlz = 0;
    const bool has_halo = halo_cell(threadIdx.x, y0, z0, hy, hz, hly, hlz);
    const long sx = (long)st.g.ny * st.g.nz;
    // prologue: plane xa - 1 (own columns), plane xa (own + halo into LDS), raw plane xa + 1
    double dm[kRows], dc[kRows];
    Raw oc[kRows], on[kRows];
    bool active[kRows];
    size_t c[kRows];
#pragma unroll
    for (int w = 0; w < kRows; ++w) {
        const int y = y0 + ty + w * kRowStep;
        active[w] = y < st.g.ny && z < st.g.nz;
        dm[w] = op.value(op.load(xa - 1, y, z, false));
        oc[w] = op.load(xa, y, z, true);
        on[w] = op.load(xa + 1, y, z, true);
        dc[w] = op.value(oc[w]);
        plane[0][ty + w * kRowStep + 1][tz + 1] = dc[w];
        c[w] = idx3(xa, min(y, st.g.ny - 1), min(z, st.g.nz - 1), st.g.ny, st.g.nz);
    }
    // a thread without a halo cell loads (hy, hz) = (0, 0) and does not use it
    const Raw hc = op.load(xa, hy, hz, false);
    Raw hn = op.load(xa + 1, hy, hz, false);
    if (has_halo) plane[0][hly][hlz] = op.value(hc);
    int cur = 0;
    __syncthreads();
    for (int x = xa; x < xb; ++x) {
        // plane x + 2: used while it is still this chunk's or its +x neighbour plane (x + 2 <= xb)
        Raw on2[kRows];
#pragma unroll
        for (int w = 0; w < kRows; ++w) on2[w] = op.load(x + 2, y0 + ty + w * kRowStep, z, true);
        const Raw hn2 = op.load(x + 2, hy, hz, false);
        double dn[kRows];
#pragma unroll
        for (int w = 0; w < kRows; ++w) {
            dn[w] = op.value(on[w]);
            plane[cur ^ 1][ty + w * kRowStep + 1][tz + 1] = dn[w];
        }
        if (has_halo) plane[cur ^ 1][hly][hlz] = op.value(hn);
#pragma unroll
        for (int w = 0; w < kRows; ++w) {
            if (!active[w]) continue;
            const int ly = ty + w * kRowStep + 1;
            const unsigned f = oc[w].f;
            double s = (double)__builtin_popcount(f & 63u) * dc[w];
            s -= (f & 1u) ? dm[w] : 0.0;
            s -= (f & 2u) ? dn[w] : 0.0;
            if constexpr (select_reads<Op>(0)) {
                s -= (f & 4u) ? plane[cur][ly - 1][tz + 1] : 0.0;
                s -= (f & 8u) ? plane[cur][ly + 1][tz + 1] : 0.0;
                s -= (f & 16u) ? plane[cur][ly][tz] : 0.0;
                s -= (f & 32u) ? plane[cur][ly][tz + 2] : 0.0;
            } else {
                const double n4 = *((f & 4u) ? &plane[cur][ly - 1][tz + 1] : &zero_slot);
                const double n8 = *((f & 8u) ? &plane[cur][ly + 1][tz + 1] : &zero_slot);
                const double n16 = *((f & 16u) ? &plane[cur][ly][tz] : &zero_slot);
                const double n32 = *((f & 32u) ? &plane[cur][ly][tz + 2] : &zero_slot);
                s -= n4;
                s -= n8;
                s -= n16;
                s -= n32;
            }
            op.cell(c[w], oc[w], dc[w], (f & kPcgUnknown) ? s : 0.0, acc);
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < kRows; ++w) {
            dm[w] = dc[w]; dc[w] = dn[w]; oc[w] = on[w]; on[w] = on2[w]; c[w] += sx;
        }
        hn = hn2;
        cur ^= 1;
    }
}

__device__ __forceinline__ void pcg_direction_body(const PcgState &st, int blk, int k) {
    const float *__restrict__ pold = st.p[k & 1];
    float *__restrict__ pnew = st.p[(k + 1) & 1];
    // the converged flag read with the partials (one round trip), tested before any write
    const double flag = st.scal[kPcgDone];
    double rz, rr, bb;
    residual_state(st, k, rz, rr, bb);
    if (block_flag(flag)) return;
    const bool done = rr <= st.scal[kPcgTol2] * bb;
    const double beta = k == 0 ? 0.0 : rz / st.scal[kPcgRZ + ((k - 1) & 1)];
    if (blk == 0 && threadIdx.x == 0) {
        double *sc = st.scal;
        sc[kPcgRZ + (k & 1)] = rz;
        sc[kPcgRR] = rr;
        if (k == 0) sc[kPcgBB] = bb;   // later kernels read it
        sc[kPcgIters] = sc[kPcgNext];  // k, counted on the device
        sc[kPcgNext] += 1.0;
        if (done) sc[kPcgDone] = 1.0;
    }
    if (done) return;
    double acc[1] = {0.0};
    tile_sweep(st, blk, DirOp{st, pold, pnew, beta}, acc);
    write_partials<1>(acc, st.part_dir, st.dir_blocks, blk);
}

// cells per thread and loop step of the streaming update (independent cells, so their loads overlap)
constexpr int kUpdCells = 4;

// Streaming update: x += alpha p, r -= alpha q with the stored q, grid-stride over nblocks blocks.
// Outside the system p = q = r = 0, so the update leaves those cells as they are.
__device__ __forceinline__ void update_stream(const PcgState &st, int blk, int k, double alpha, double *acc) {
    const float *__restrict__ pnew = st.p[(k + 1) & 1];
    const size_t n = (size_t)st.g.nx * st.g.ny * st.g.nz;
    const uint8_t *__restrict__ fl = st.flags;
    double *__restrict__ x = st.x;
    double *__restrict__ r = st.r;
    const double *__restrict__ q = st.q;
    const CellRange cr = block_cells(n, st.nblocks, blk);
    const size_t stride = cr.step;
    for (size_t c0 = cr.first; c0 < cr.end; c0 += kUpdCells * stride) {
        double xv[kUpdCells], rv[kUpdCells], id[kUpdCells];
#pragma unroll
        for (int u = 0; u < kUpdCells; ++u) {
            const size_t cu = c0 + u * stride;
            const size_t c = cu < cr.end ? cu : c0;
            xv[u] = x[c] + alpha * (double)pnew[c];
            rv[u] = r[c] - alpha * q[c];
            id[u] = inv_diag(fl[c]);
        }
#pragma unroll
        for (int u = 0; u < kUpdCells; ++u) {
            const size_t c = c0 + u * stride;
            if (c < cr.end) {
                x[c] = xv[u];
                r[c] = rv[u];
                acc[0] += rv[u] * (rv[u] * id[u]);
                acc[1] += rv[u] * rv[u];
            }
        }
    }
}

// Iteration k's update: alpha from the direction kernel's partials, then the solve's form.
template <bool RECOMPUTE> __device__ __forceinline__ void pcg_update_body(const PcgState &st, int blk, int k) {
    const double flag = st.scal[kPcgDone];   // read with the partials, as in the direction step
    double pq[1];
    total_of<1>(st.part_dir, st.dir_blocks, pq);
    if (block_flag(flag)) return;
    const double alpha = pq[0] != 0.0 ? st.scal[kPcgRZ + (k & 1)] / pq[0] : 0.0;
    double acc[2] = {0.0, 0.0};
    if constexpr (RECOMPUTE) tile_sweep(st, blk, UpdOp{st, st.p[(k + 1) & 1], alpha}, acc);
    else update_stream(st, blk, k, alpha, acc);
    write_partials<2>(acc, st.part_update, st.upd_blocks, blk);
}

// ---- one sweep per iteration (Jacobi) ----------------------------------------------------------
//
// With z = r / diag, the residual dot of the next iteration follows from dots over the current
// vectors: rz_(k+1) = sum (r_k - alpha_k q_k)^2 / d = rz_k - 2 alpha_k (r_k.q_k / d)
// + alpha_k^2 (q_k.q_k / d).  So one sweep can finish iteration k and start iteration k + 1: with
// alpha_k = rz_k / p_k.q_k and beta = rz_(k+1) / rz_k from the previous sweep's sums, per cell
//   q_k = L p_k (recomputed from the stored p_k), r_(k+1) = r_k - alpha_k q_k,
//   x_(k+1) = x_k + alpha_k p_k, p_(k+1) = r_(k+1) / d + beta p_k (rounded to FP32),
//   q_(k+1) = L p_(k+1),
// storing r and p and summing rz, rr, p.q, r.q / d and q.q / d of the new vectors (the true
// rz_(k+1) enters the next alpha; only beta uses the expansion, whose terms come from one sweep, so
// its rounding does not accumulate).  q_(k+1) of a cell needs p_(k+1) of its neighbours, so the
// sweep computes r_(k+1) and p_(k+1) on the tile and a one-cell ring around it, whose q_k needs
// p_k on a two-cell ring; along x, stage A (r, x, p of plane X) runs one plane ahead of stage B
// (q_(k+1) and the sums of plane X - 1).  The ring reads r_k of the neighbouring tiles, which
// their own blocks overwrite during the same sweep, so r alternates between two buffers as p does
// (x is read only by its own cells and stays in place).
//
// x takes two steps at once: even sweeps k >= 2 store x_k = x_(k-2) + (alpha_(k-2) p_(k-2) +
// alpha_(k-1) p_(k-1)), odd sweeps leave x alone.  p_(k-1) is the sweep's p_old; p_(k-2) is still
// in the buffer the sweep writes p_k to, and a block reads it there only at its own cells, each
// before the same thread overwrites it; alpha_(k-2) is stored by the odd sweep k - 1.  x is
// exactly the sum of the stored p times the same alphas, so the recurrence r = b - L x holds as
// before, up to the grouping of two FP64 additions.  Per cell, odd sweeps move 25 bytes (flags 1,
// p 4 + 4, r 8 + 8), even ones 45 (x 8 + 8, p_(k-2) 4 more): 35 on average, against 41 with x
// every sweep and 54 for the direction and update sweeps, and one launch per iteration.  A solve
// that stops after an odd iteration I leaves alpha_(I-1) p_(I-1) unadded; pcg_fused_flush_kernel
// adds it.
//
// Sweep 0 (alpha = beta = 0, p_old = 0) forms p_0 = z_0 and the sums of r_0; sweep j >= 1 runs
// iteration j after testing the residual of iteration j - 1.
constexpr int kRing1 = 2 * (kTileZ + 2) + 2 * kTileY;         // cells of the one-cell ring
constexpr int kRing2 = 2 * (kTileZ + 4) + 2 * (kTileY + 2);   // cells of the second ring
static_assert(kRing1 <= kPB && kRing2 <= kPB, "ring cells: one per thread");
static_assert(kRows == 4, "the fused sweep packs the rows' flag bytes into one word");

// local (y, z) of ring cell t (tile rows 0 .. kTileY - 1, columns 0 .. kTileZ - 1), as selects
__device__ __forceinline__ int2 ring1_cell(int t) {
    const int a = kTileZ + 2, b = 2 * a, c = b + kTileY;
    const int ly = t < a ? -1 : t < b ? kTileY : t < c ? t - b : t - c;
    const int lz = t < a ? t - 1 : t < b ? t - a - 1 : t < c ? -1 : kTileZ;
    return make_int2(ly, lz);
}
__device__ __forceinline__ int2 ring2_cell(int t) {
    const int a = kTileZ + 4, b = 2 * a, c = b + kTileY + 2;
    const int ly = t < a ? -2 : t < b ? kTileY + 1 : t < c ? t - b - 1 : t - c - 1;
    const int lz = t < a ? t - 2 : t < b ? t - a - 2 : t < c ? -2 : kTileZ + 1;
    return make_int2(ly, lz);
}

// flags and residual of a cell, as stored (zero outside the grid)
struct FusedRaw {
    unsigned f;
    double r;
};

// The fused sweep of one tile: sums acc = {rz, rr, p.q, r.q / d, q.q / d} of its own cells; with
// XUPD it also adds a2 p_(k-2) + alpha p_(k-1) to x.  Registers per thread: p_k two planes ahead,
// flags and r one plane ahead, and with XUPD x and p_(k-2) of the next plane, whose x is stored
// at the end of the plane's iteration.  The one-cell ring (a stencil each) goes to the last kRing1 threads and the
// second ring (a load each) to the first kRing2, so the four waves share the extra work.
template <bool XUPD>
__device__ __forceinline__ void fused_sweep(const PcgState &st, int blk, const float *__restrict__ pold,
                                            float *__restrict__ pnew, const double *__restrict__ rk,
                                            double *__restrict__ rw, double alpha, double beta, double a2,
                                            double (&acc)[5]) {
    // p_k of plane X (two-cell ring) and p_(k+1) of planes X and X - 1 (one-cell ring), both by
    // the parity of the plane, so one barrier per plane suffices: stage A of plane X reads P0[X & 1]
    // and writes P1[X & 1], stage B reads P1 of plane X - 1, and a buffer is written again two
    // planes later, after the next plane's barrier, which every thread passes only when done with
    // it.  (R1 is per thread.)
    __shared__ float P0[2][kTileY + 4][kTileZ + 4];
    __shared__ float P1[2][kTileY + 2][kTileZ + 2];
    // with XUPD, r_(k+1) of the own cells waits for stage B here (per thread, by plane parity)
    // rather than in registers, which x and p_(k-2) need
    __shared__ double R1[XUPD ? 2 : 1][kRows][XUPD ? kPB : 1];
    // +0.0f, read in place of a neighbour that is not fluid, and 1 / diag by neighbour count (index
    // 7: not an unknown).  The stencils' LDS reads take their address from the flag bit instead of
    // sitting under a select of the value, which the compiler turns into a branch with its own wait
    // per read; inv_diag's chain of selects likewise becomes a tree of branches.
    __shared__ float zero_slot;
    __shared__ double inv_tab[8];
    const int per = st.dir_blocks / 8;
    const int L = (blk % 8) * per + blk / 8;
    if (L >= st.dir_tiles_y * st.dir_tiles_z * st.dir_chunks) return;
    if (threadIdx.x < 8) inv_tab[threadIdx.x] = inv_diag(threadIdx.x < 7 ? kPcgUnknown | ((1u << threadIdx.x) - 1u) : 0u);
    if (threadIdx.x == 0) zero_slot = 0.0f;
    // read after the first barrier of the plane loop
    auto inv_of = [&](unsigned f) { return inv_tab[(f & kPcgUnknown) ? __builtin_popcount(f & 63u) : 7]; };
    const int nx = st.g.nx, ny = st.g.ny, nz = st.g.nz;
    const int y0 = (L % st.dir_tiles_y) * kTileY;
    const int z0 = ((L / st.dir_tiles_y) % st.dir_tiles_z) * kTileZ;
    const int xa = (L / (st.dir_tiles_y * st.dir_tiles_z)) * st.dir_chunk_x;
    const int xb = min(nx, xa + st.dir_chunk_x);
    constexpr int kRowStep = kTileY / kRows;
    const int ty = threadIdx.x / kTileZ, tz = threadIdx.x % kTileZ;
    const long sx = (long)ny * nz;
    const uint8_t *__restrict__ fl = st.flags;
    double *__restrict__ xw = st.x;
    // Every load of the plane loop is unconditional, from an address clamped into the grid, and a
    // select zeroes what lies outside: a load under a branch that all lanes may skip leaves the
    // compiler unable to count the loads issued after an earlier one, and it then waits for all of
    // them (vmcnt(0)) before the plane's first LDS store, which undoes the one-plane prefetch.
    auto inside = [&](int x, int y, int z) { return x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz; };
    auto clamped = [&](int x, int y, int z) {
        return idx3(min(max(x, 0), nx - 1), min(max(y, 0), ny - 1), min(max(z, 0), nz - 1), ny, nz);
    };
    auto load_p = [&](int x, int y, int z) -> float {
        const float v = pold[clamped(x, y, z)];
        return inside(x, y, z) ? v : 0.0f;
    };
    auto load_raw = [&](int x, int y, int z) -> FusedRaw {
        const size_t c = clamped(x, y, z);
        const unsigned f = fl[c];
        const double r = rk[c];
        const bool in = inside(x, y, z);
        return FusedRaw{in ? f : 0u, in ? r : 0.0};
    };
    // ring cells of this thread
    const int t1 = (int)threadIdx.x - (kPB - kRing1);
    const bool has1 = t1 >= 0, has2 = threadIdx.x < kRing2;
    const int2 g1 = ring1_cell(has1 ? t1 : 0), g2 = ring2_cell(has2 ? (int)threadIdx.x : 0);
    const int r1y = g1.x, r1z = g1.y, r2y = g2.x, r2z = g2.y;
    // the ring cells' (y, z) address parts, clamped into the grid, and whether they lie in it; the
    // plane loop adds the plane part, which is the same for every thread
    auto yz_part = [&](int y, int z) { return (size_t)min(max(y, 0), ny - 1) * nz + min(max(z, 0), nz - 1); };
    auto yz_in = [&](int y, int z) { return y >= 0 && y < ny && z >= 0 && z < nz; };
    const size_t r1yz = yz_part(y0 + r1y, z0 + r1z), r2yz = yz_part(y0 + r2y, z0 + r2z);
    const bool r1in = yz_in(y0 + r1y, z0 + r1z), r2in = yz_in(y0 + r2y, z0 + r2z);
    const int z = z0 + tz;
    bool act[kRows];
    size_t c[kRows];   // own cell of plane max(X, 0), y and z clamped into the grid
    // own cells: p_k of planes X - 1 .. X + 2; flags and r of X and X + 1
    float pm[kRows], pc[kRows], pn[kRows];
    // flags of the kRows rows are packed one byte each (ocf: plane X, fbf: plane X - 1)
    double ocr[kRows];
    unsigned ocf = 0u;
    // own cells for stage B: p_(k+1) of planes X - 2 and X - 1, r_(k+1) and flags of X - 1
    float qm[kRows], qc[kRows];
    double rb[kRows];
    unsigned fbf = 0u;
    // ring cells: p_k of X - 1 .. X + 1 (ring 1), X and X + 1 (ring 2); flags and r of X (ring 1)
    float gm = 0.f, gc = 0.f, gn = 0.f, hc = 0.f, hn = 0.f;
    FusedRaw ac{0u, 0.0};
    const int xs = xa - 1;   // first plane of stage A
#pragma unroll
    for (int w = 0; w < kRows; ++w) {
        const int y = y0 + ty + w * kRowStep;
        act[w] = y < ny && z < nz;
        c[w] = idx3(max(xs, 0), min(y, ny - 1), min(z, nz - 1), ny, nz);
        pm[w] = load_p(xs - 1, y, z);
        pc[w] = load_p(xs, y, z);
        pn[w] = load_p(xs + 1, y, z);
        const FusedRaw o = load_raw(xs, y, z);
        ocr[w] = o.r;
        ocf |= o.f << (8 * w);
        qm[w] = 0.f; qc[w] = 0.f; rb[w] = 0.0;
    }
    if (has1) {
        gm = load_p(xs - 1, y0 + r1y, z0 + r1z);
        gc = load_p(xs, y0 + r1y, z0 + r1z);
        gn = load_p(xs + 1, y0 + r1y, z0 + r1z);
        ac = load_raw(xs, y0 + r1y, z0 + r1z);
    }
    if (has2) {
        hc = load_p(xs, y0 + r2y, z0 + r2z);
        hn = load_p(xs + 1, y0 + r2y, z0 + r2z);
    }
    for (int X = xs; X <= xb; ++X) {
        const bool own_plane = X >= xa && X < xb;
        // loads for the next planes: raw of X + 1 (a stage-A plane while X + 1 <= xb), p_k of X + 2
        // (needed by stage A while X + 2 <= xb + 1); past those, loaded and not used.  X + 1 >= 0,
        // so only the top of the grid clamps the planes; an own row's cell of plane min(X + k,
        // nx - 1) is its cell of plane max(X, 0) plus a plane offset.
        const int x1 = min(X + 1, nx - 1), x2 = min(X + 2, nx - 1);
        const bool in1 = X + 1 < nx, in2 = X + 2 < nx;
        const long o1 = (long)(x1 - max(X, 0)) * sx, o2 = (long)(x2 - max(X, 0)) * sx;
        // As loaded; the selects for cells outside the grid come at the end of the plane: a select
        // next to its load, on a condition uniform across the block (in1, in2), is compiled into a
        // branch around the load with a wait for it.
        unsigned on_f[kRows];
        double on_r[kRows];
        float pn2[kRows];
#pragma unroll
        for (int w = 0; w < kRows; ++w) {
            on_f[w] = fl[(long)c[w] + o1];
            on_r[w] = rk[(long)c[w] + o1];
            pn2[w] = pold[(long)c[w] + o2];
        }
        // ring cells (threads without one load a cell of the first ring and do not use it)
        const size_t a1 = r1yz + (size_t)x1 * sx;
        const unsigned an_f = fl[a1];
        const double an_r = rk[a1];
        const float gn2 = pold[r1yz + (size_t)x2 * sx], hn2 = pold[r2yz + (size_t)x2 * sx];
        // p_k of plane X into LDS
        const int b1 = X & 1;
#pragma unroll
        for (int w = 0; w < kRows; ++w) P0[b1][ty + w * kRowStep + 2][tz + 2] = pc[w];
        if (has1) P0[b1][r1y + 2][r1z + 2] = gc;
        if (has2) P0[b1][r2y + 2][r2z + 2] = hc;
        __syncthreads();
        // stage A on plane X: q_k, r_(k+1), p_(k+1) on the tile and the one-cell ring
        auto stage_a = [&](int ly, int lz, unsigned f, double r, float pmv, float pcv, float pnv, double &r1,
                           float &p1) {
            const float n4 = *((f & 4u) ? &P0[b1][ly + 1][lz + 2] : &zero_slot);
            const float n8 = *((f & 8u) ? &P0[b1][ly + 3][lz + 2] : &zero_slot);
            const float n16 = *((f & 16u) ? &P0[b1][ly + 2][lz + 1] : &zero_slot);
            const float n32 = *((f & 32u) ? &P0[b1][ly + 2][lz + 3] : &zero_slot);
            const double id = inv_of(f);
            double q = (double)__builtin_popcount(f & 63u) * (double)pcv;
            q -= (f & 1u) ? (double)pmv : 0.0;
            q -= (f & 2u) ? (double)pnv : 0.0;
            q -= (double)n4;
            q -= (double)n8;
            q -= (double)n16;
            q -= (double)n32;
            q = (f & kPcgUnknown) ? q : 0.0;
            r1 = r - alpha * q;
            p1 = (float)(r1 * id + beta * (double)pcv);
            P1[b1][ly + 1][lz + 1] = p1;
        };
        double r1o[kRows];
        float p1o[kRows];
#pragma unroll
        for (int w = 0; w < kRows; ++w) {
            const int ly = ty + w * kRowStep;
            stage_a(ly, tz, (ocf >> (8 * w)) & 0xffu, ocr[w], pm[w], pc[w], pn[w], r1o[w], p1o[w]);
            if constexpr (XUPD) R1[b1][w][threadIdx.x] = r1o[w];
        }
        // the stores after all rows' stencils, so that the rows' LDS reads share one block
        if (own_plane) {
#pragma unroll
            for (int w = 0; w < kRows; ++w) {
                if (!act[w]) continue;
                rw[c[w]] = r1o[w];
                pnew[c[w]] = p1o[w];
            }
        }
        if (has1) {
            double r1;
            float p1;
            stage_a(r1y, r1z, ac.f, ac.r, gm, gc, gn, r1, p1);
        }
        // x and p_(k-2) of the own cells of plane X + 1, stored after stage B (the cell's p_(k-2)
        // is read before this thread overwrites it, in stage A of the next plane)
        const bool own_next = X + 1 < xb;
        double xo[kRows];
        float po[kRows];
        // from plane max(X, 0) of c to plane min(X + 1, nx - 1) (past xb: in the grid, not used)
        const long xoff = o1;
        if constexpr (XUPD) {
#pragma unroll
            for (int w = 0; w < kRows; ++w) {
                xo[w] = xw[(long)c[w] + xoff];
                po[w] = pnew[(long)c[w] + xoff];
            }
        }
        // stage B on plane X - 1 (own cells): q_(k+1) and the sums
        if (X > xa) {
            const int b0 = b1 ^ 1;
#pragma unroll
            for (int w = 0; w < kRows; ++w) {
                if (!act[w]) continue;
                const int ly = ty + w * kRowStep + 1, lz = tz + 1;
                const unsigned f = (fbf >> (8 * w)) & 0xffu;
                const float n4 = *((f & 4u) ? &P1[b0][ly - 1][lz] : &zero_slot);
                const float n8 = *((f & 8u) ? &P1[b0][ly + 1][lz] : &zero_slot);
                const float n16 = *((f & 16u) ? &P1[b0][ly][lz - 1] : &zero_slot);
                const float n32 = *((f & 32u) ? &P1[b0][ly][lz + 1] : &zero_slot);
                const double id = inv_of(f);
                double q = (double)__builtin_popcount(f & 63u) * (double)qc[w];
                q -= (f & 1u) ? (double)qm[w] : 0.0;
                q -= (f & 2u) ? (double)p1o[w] : 0.0;
                q -= (double)n4;
                q -= (double)n8;
                q -= (double)n16;
                q -= (double)n32;
                q = (f & kPcgUnknown) ? q : 0.0;
                double r;
                if constexpr (XUPD) r = R1[b0][w][threadIdx.x];
                else r = rb[w];
                const double pv = (double)qc[w];
                acc[0] += r * (r * id);
                acc[1] += r * r;
                acc[2] += pv * q;
                acc[3] += r * (q * id);
                acc[4] += q * (q * id);
            }
        }
        // x of plane X + 1: p_(k-1) there is pn
        if constexpr (XUPD) {
#pragma unroll
            for (int w = 0; w < kRows; ++w)
                if (own_next && act[w]) xw[(long)c[w] + xoff] = xo[w] + (a2 * (double)po[w] + alpha * (double)pn[w]);
        }
        // advance one plane
        fbf = ocf;
        ocf = 0u;
#pragma unroll
        for (int w = 0; w < kRows; ++w) {
            qm[w] = qc[w]; qc[w] = p1o[w];
            if constexpr (!XUPD) rb[w] = r1o[w];
            pm[w] = pc[w]; pc[w] = pn[w]; pn[w] = in2 && act[w] ? pn2[w] : 0.0f;
            ocr[w] = in1 && act[w] ? on_r[w] : 0.0;
            ocf |= (in1 && act[w] ? on_f[w] : 0u) << (8 * w);
            if (X >= 0) c[w] += sx;
        }
        gm = gc; gc = gn; gn = in2 && r1in ? gn2 : 0.0f;
        ac = FusedRaw{in1 && r1in ? an_f : 0u, in1 && r1in ? an_r : 0.0};
        hc = hn; hn = in2 && r2in ? hn2 : 0.0f;
    }
}

// Sweep k of a fused solve (see above): the sums of sweep k - 1 (of the start kernel for k = 0)
// give the convergence test, alpha and beta; then the sweep.
template <bool XUPD> __device__ __forceinline__ void pcg_fused_body(const PcgState &st, int blk, int k) {
    // the converged flag is read before the previous sweep's partials and tested after them, so
    // the two round trips overlap; thread 0's value decides for the block (as in solve_done), and
    // the test comes before block 0 updates the counters
    const double flag = st.scal[kPcgDone];
    double rz, rr, bb, alpha = 0.0, beta = 0.0;
    if (k == 0) {
        double tot[3];
        total_of<3>(st.part_start, st.nblocks, tot);
        rz = tot[0]; rr = tot[1]; bb = tot[2];
    } else {
        double tot[5];
        total_of<5>(st.part_fused + (size_t)((k - 1) & 1) * 5 * st.dir_blocks, st.dir_blocks, tot);
        rz = tot[0]; rr = tot[1]; bb = st.scal[kPcgBB];
        alpha = tot[2] != 0.0 ? rz / tot[2] : 0.0;
        const double rz1 = rz - 2.0 * alpha * tot[3] + alpha * alpha * tot[4];
        beta = rz != 0.0 ? rz1 / rz : 0.0;
    }
    if (block_flag(flag)) return;
    const bool done = rr <= st.scal[kPcgTol2] * bb;
    if (blk == 0 && threadIdx.x == 0) {
        double *sc = st.scal;
        sc[kPcgRR] = rr;
        if (k == 0) sc[kPcgBB] = bb;
        const double next = sc[kPcgNext];   // sweeps run: iterations done are one fewer
        sc[kPcgIters] = next > 0.0 ? next - 1.0 : 0.0;
        sc[kPcgNext] = next + 1.0;
        if (done) sc[kPcgDone] = 1.0;
        else if (k & 1) sc[kPcgAlpha2] = alpha;   // alpha_(k-1), for the x step of sweep k + 1
    }
    if (done) return;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // residuals alternate between r (even sweeps read it) and r_alt; the start kernel wrote r
    fused_sweep<XUPD>(st, blk, st.p[k & 1], st.p[(k + 1) & 1], (k & 1) ? st.r_alt : st.r, (k & 1) ? st.r : st.r_alt,
                      alpha, beta, XUPD ? st.scal[kPcgAlpha2] : 0.0, acc);
    write_partials<5>(acc, st.part_fused + (size_t)(k & 1) * 5 * st.dir_blocks, st.dir_blocks, blk);
}

// Energy-optimal scaling of an initial guess: partial sums of x.b and x.(L x) ...
__global__ void __launch_bounds__(kPB) pcg_energy_kernel(PcgState st) {
    const int blk = blockIdx.x;
    const size_t n = (size_t)st.g.nx * st.g.ny * st.g.nz;
    const long sx = (long)st.g.ny * st.g.nz, sy = st.g.nz;
    const double *x = st.x;
    double acc[2] = {0.0, 0.0};
    const CellRange cr = block_cells(n, st.nblocks, blk);
    for (size_t c = cr.first; c < cr.end; c += cr.step) {
        const unsigned f = st.flags[c];
        const double s = stencil(f, c, sx, sy, [&](size_t i) { return x[i]; });
        if (f & kPcgUnknown) {
            acc[0] += x[c] * st.b[c];
            acc[1] += x[c] * s;
        }
    }
    write_partials<2>(acc, st.part_start, st.nblocks, blk);
}

// ... then x *= gamma = x.b / x.L x, the factor that minimises the energy (the A-norm error) of
// gamma x; 0 when x.L x is not positive.  Block 0 keeps gamma in the scalars.
__global__ void __launch_bounds__(kPB) pcg_scale_kernel(PcgState st) {
    const int blk = blockIdx.x;
    double tot[2];
    total_of<2>(st.part_start, st.nblocks, tot);
    const double gamma = tot[1] > 0.0 ? tot[0] / tot[1] : 0.0;
    if (blk == 0 && threadIdx.x == 0) st.scal[kPcgGamma] = gamma;
    const size_t n = (size_t)st.g.nx * st.g.ny * st.g.nz;
    const CellRange cr = block_cells(n, st.nblocks, blk);
    for (size_t c = cr.first; c < cr.end; c += cr.step) st.x[c] *= gamma;
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

// x = initial guess on the unknowns (0 on obstacles and on the pinned cell)
__global__ void pcg_guess_kernel(PcgState st, const float *p_in) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)st.g.nx * st.g.ny * st.g.nz;
    if (c >= n) return;
    st.x[c] = (p_in && (st.flags[c] & kPcgUnknown)) ? (double)p_in[c] : 0.0;
}

// ---- multigrid preconditioner -----------------------------------------------------------------
// Jacobi damping of the smoother, sweeps on the coarsest level and the scale of the coarse-grid
// correction.  The damped Jacobi iteration matrix I - w D^-1 A of a graph Laplacian has spectral
// radius below 1 for w < 1, and a coarse correction scaled by a < 2 (with a coarse solve that
// under-approximates the inverse, as damped Jacobi from zero does) leaves I - a P Ac^-1 P^T A an
// A-norm contraction, so the V-cycle stays symmetric positive definite.  Piecewise-constant
// aggregation under-corrects smooth errors; the scale makes up for it.  Measured at 256^3 (CG
// iterations to 1e-6): w 0.8 a 1.0: 85; 0.67 / 1.5: 52; 0.8 / 1.5: 42; 0.8 / 1.8: 40; 0.9 / 1.5: 37;
// 0.9 / 1.8: 33.
constexpr float kMgOmega = 0.9f;
constexpr int kMgCoarseSweeps = 48;
constexpr float kMgCorrScale = 1.8f;
constexpr int kMgCoarsestCells = 256;

// A cell of a level: unknown, diagonal and the weights to its six neighbours (-x, +x, -y, +y, -z,
// +z).  Level 0 reads the flag byte (weight 1 to each fluid neighbour, the diagonal counts them);
// coarser levels read {diag, w+x, w+y, w+z} and the - weights from the neighbours' + weights.
struct MgCell {
    bool unknown;
    float d;
    float w[6];
};
__device__ __forceinline__ MgCell mg_cell0(const uint8_t *flags, size_t c) {
    const unsigned f = flags[c];
    MgCell a;
    a.unknown = (f & kPcgUnknown) != 0;
    a.d = (float)__builtin_popcount(f & 63u);
#pragma unroll
    for (int q = 0; q < 6; ++q) a.w[q] = (f >> q) & 1u ? 1.0f : 0.0f;
    return a;
}
__device__ __forceinline__ MgCell mg_cell(const MgLevel &L, int i, int j, int k) {
    const size_t sx = (size_t)L.ny * L.nz, sy = L.nz;
    const size_t c = idx3(i, j, k, L.ny, L.nz);
    const float4 o = L.op[c];
    MgCell a;
    a.unknown = o.x > 0.0f;
    a.d = o.x;
    a.w[0] = i > 0 ? L.op[c - sx].y : 0.0f;
    a.w[1] = o.y;
    a.w[2] = j > 0 ? L.op[c - sy].z : 0.0f;
    a.w[3] = o.z;
    a.w[4] = k > 0 ? L.op[c - 1].w : 0.0f;
    a.w[5] = o.w;
    return a;
}
__device__ __forceinline__ bool mg_unknown(const PcgState &st, const MgLevel &L, int l, size_t c) {
    return l == 0 ? (st.flags[c] & kPcgUnknown) != 0 : L.op[c].x > 0.0f;
}
__device__ __forceinline__ MgCell mg_get(const PcgState &st, const MgLevel &L, int l, int i, int j, int k) {
    return l == 0 ? mg_cell0(st.flags, idx3(i, j, k, L.ny, L.nz)) : mg_cell(L, i, j, k);
}
__device__ __forceinline__ void mg_neighbour(int q, int i, int j, int k, int &ni, int &nj, int &nk) {
    ni = i + (q == 1) - (q == 0);
    nj = j + (q == 3) - (q == 2);
    nk = k + (q == 5) - (q == 4);
}

// Level l + 1's operator, one thread per aggregate: the diagonal sums the children's diagonals
// less twice the weights of the faces between children; a child's face to an unknown of the next
// aggregate up an axis adds to that + weight (the - weights are the neighbours' + weights).
__global__ void __launch_bounds__(kPB) mg_build_kernel(PcgState st, MgLevel F, MgLevel C, int l) {
    const size_t n = (size_t)C.nx * C.ny * C.nz;
    const size_t I = (size_t)blockIdx.x * kPB + threadIdx.x;
    if (I >= n) return;
    const int K = (int)(I % C.nz), J = (int)((I / C.nz) % C.ny), II = (int)(I / ((size_t)C.ny * C.nz));
    float d = 0.0f, wp[3] = {0.0f, 0.0f, 0.0f};
    for (int ci = 0; ci < 2; ++ci)
        for (int cj = 0; cj < 2; ++cj)
            for (int ck = 0; ck < 2; ++ck) {
                const int i = 2 * II + ci, j = 2 * J + cj, k = 2 * K + ck;
                if (i >= F.nx || j >= F.ny || k >= F.nz) continue;
                const MgCell a = mg_get(st, F, l, i, j, k);
                if (!a.unknown) continue;
                d += a.d;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    if (a.w[q] == 0.0f) continue;
                    int ni, nj, nk;
                    mg_neighbour(q, i, j, k, ni, nj, nk);
                    if (ni < 0 || nj < 0 || nk < 0 || ni >= F.nx || nj >= F.ny || nk >= F.nz) continue;
                    if (!mg_unknown(st, F, l, idx3(ni, nj, nk, F.ny, F.nz))) continue;   // Dirichlet: in d
                    if ((ni >> 1) == II && (nj >> 1) == J && (nk >> 1) == K) d -= a.w[q];
                    else if (q & 1) wp[q >> 1] += a.w[q];
                }
            }
    C.op[I] = make_float4(d, wp[0], wp[1], wp[2]);
}

// b and L x of a cell at level l (level 0: b = r, FP64 in memory)
__device__ __forceinline__ float mg_b(const PcgState &st, const MgLevel &L, int l, size_t c) {
    return l == 0 ? (float)st.r[c] : L.b[c];
}

// Down the V-cycle at level l, one thread per cell of a 2 x 2 x 64 tile (32 aggregates of level
// l + 1 along z): every unknown cell gets the first smoothing sweep from zero, x = w b / d, and
// its residual b - L x (the neighbours' x recomputed from their b); each aggregate's right-hand
// side is the sum of its eight children's residuals, added in a fixed order through LDS.
// The down step's residual of cell (i, j, k) of a level l >= 1 (x = w b / d on every unknown),
// with every load issued at once: the cell's operator record and those of its -x, -y, -z
// neighbours (the weights), and each neighbour's record and right-hand side, from indices clamped
// into the grid.  A neighbour that is outside, not an unknown, or joined by a zero weight adds
// (its term) * 0 with a denominator that cannot be 0; the others add what the tested form adds.
__device__ __forceinline__ float mg_down_cell_coarse(const MgLevel &F, int i, int j, int k) {
    const int sx = F.ny * F.nz, sy = F.nz;
    const int c = (i * F.ny + j) * F.nz + k;
    const float4 o = F.op[c];
    const float4 oxm = F.op[i > 0 ? c - sx : c], oym = F.op[j > 0 ? c - sy : c], ozm = F.op[k > 0 ? c - 1 : c];
    const float bc = F.b[c];
    const float w[6] = {oxm.y * (float)(i > 0), o.y, oym.z * (float)(j > 0), o.z, ozm.w * (float)(k > 0), o.w};
    float dn[6], bn[6];
    bool in[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        int ni, nj, nk;
        mg_neighbour(q, i, j, k, ni, nj, nk);
        in[q] = ni >= 0 && nj >= 0 && nk >= 0 && ni < F.nx && nj < F.ny && nk < F.nz;
        const int cn = (clampi(ni, 0, F.nx - 1) * F.ny + clampi(nj, 0, F.ny - 1)) * F.nz + clampi(nk, 0, F.nz - 1);
        dn[q] = F.op[cn].x;
        bn[q] = F.b[cn];
    }
    if (!(o.x > 0.0f)) {
        F.x[c] = 0.0f;
        return 0.0f;
    }
    const float d = o.x;
    const float xc = kMgOmega * bc / d;
    F.x[c] = xc;
    float lx = d * xc;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const bool use = w[q] != 0.0f && in[q] && dn[q] > 0.0f;
        const float dq = dn[q] + (use ? 0.0f : 1.0f);
        lx -= (w[q] * (kMgOmega * bn[q] / dq)) * (float)use;
    }
    return bc - lx;
}

__device__ __forceinline__ void mg_down_body(const PcgState &st, const MgLevel &F, const MgLevel &C, int l, int blk) {
    if (solve_done(st.scal)) return;
    __shared__ float res[kPB];
    const int tz = threadIdx.x % kMgTileZ, ty = (threadIdx.x / kMgTileZ) & 1, tx = threadIdx.x / (2 * kMgTileZ);
    const int tiles_z = (F.nz + kMgTileZ - 1) / kMgTileZ, tiles_y = (F.ny + 1) / 2;
    const int bz = blk % tiles_z, by = (blk / tiles_z) % tiles_y, bx = blk / (tiles_z * tiles_y);
    const int i = 2 * bx + tx, j = 2 * by + ty, k = bz * kMgTileZ + tz;
    float rc = 0.0f;
    if (l >= 1) {
        if (i < F.nx && j < F.ny && k < F.nz) rc = mg_down_cell_coarse(F, i, j, k);
    } else if (i < F.nx && j < F.ny && k < F.nz) {
        const size_t c = idx3(i, j, k, F.ny, F.nz);
        const MgCell a = mg_get(st, F, l, i, j, k);
        if (a.unknown) {
            const float bc = mg_b(st, F, l, c);
            const float xc = kMgOmega * bc / a.d;
            F.x[c] = xc;
            float lx = a.d * xc;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if (a.w[q] == 0.0f) continue;
                int ni, nj, nk;
                mg_neighbour(q, i, j, k, ni, nj, nk);
                if (ni < 0 || nj < 0 || nk < 0 || ni >= F.nx || nj >= F.ny || nk >= F.nz) continue;
                const MgCell b = mg_get(st, F, l, ni, nj, nk);
                if (!b.unknown) continue;
                lx -= a.w[q] * (kMgOmega * mg_b(st, F, l, idx3(ni, nj, nk, F.ny, F.nz)) / b.d);
            }
            rc = bc - lx;
        } else {
            F.x[c] = 0.0f;
        }
    }
    res[threadIdx.x] = rc;
    __syncthreads();
    if (threadIdx.x < kMgTileZ / 2) {
        const int az = threadIdx.x;
        const int I = bx, J = by, K = bz * (kMgTileZ / 2) + az;
        if (I < C.nx && J < C.ny && K < C.nz) {
            float sum = 0.0f;
#pragma unroll
            for (int cx = 0; cx < 2; ++cx)
#pragma unroll
                for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                    for (int cz = 0; cz < 2; ++cz) sum += res[cx * 2 * kMgTileZ + cy * kMgTileZ + 2 * az + cz];
            const size_t ci = idx3(I, J, K, C.ny, C.nz);
            C.b[ci] = C.op[ci].x > 0.0f ? sum : 0.0f;
        }
    }
}
__global__ void __launch_bounds__(kPB) mg_down_kernel(PcgState st, MgLevel F, MgLevel C, int l) {
    mg_down_body(st, F, C, l, blockIdx.x);
}

// The coarsest level (at most kMgCoarsestCells cells, one per thread): kMgCoarseSweeps damped
// Jacobi sweeps from zero, in LDS.
__device__ __forceinline__ void mg_coarsest_body(const PcgState &st, const MgLevel &L) {
    if (solve_done(st.scal)) return;
    __shared__ float xs[2][kMgCoarsestCells];
    const int n = L.nx * L.ny * L.nz;
    const int c = threadIdx.x;
    int i = 0, j = 0, k = 0;
    MgCell a{false, 0.0f, {0, 0, 0, 0, 0, 0}};
    float bc = 0.0f;
    if (c < n) {
        k = c % L.nz; j = (c / L.nz) % L.ny; i = c / (L.ny * L.nz);
        a = mg_cell(L, i, j, k);
        bc = L.b[c];
    }
    xs[0][c] = 0.0f;
    __syncthreads();
    int cur = 0;
    for (int s = 0; s < kMgCoarseSweeps; ++s) {
        float v = 0.0f;
        if (c < n && a.unknown) {
            float lx = a.d * xs[cur][c];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if (a.w[q] == 0.0f) continue;
                int ni, nj, nk;
                mg_neighbour(q, i, j, k, ni, nj, nk);
                lx -= a.w[q] * xs[cur][(ni * L.ny + nj) * L.nz + nk];
            }
            v = xs[cur][c] + kMgOmega * (bc - lx) / a.d;
        }
        xs[cur ^ 1][c] = v;
        __syncthreads();
        cur ^= 1;
    }
    if (c < n) L.x[c] = xs[cur][c];
}
__global__ void __launch_bounds__(kPB) mg_coarsest_kernel(PcgState st, MgLevel L) { mg_coarsest_body(st, L); }

// Up the V-cycle at level l: x' = x + the aggregate's correction, then one smoothing sweep
// x'' = x' + w (b - L x') / d (the neighbours' x' recomputed the same way).
__device__ __forceinline__ float mg_up_cell(const PcgState &st, const MgLevel &F, const float *corr, int cny, int cnz,
                                            int l, size_t c, float bc) {
    const int k = (int)(c % F.nz), j = (int)((c / F.nz) % F.ny), i = (int)(c / ((size_t)F.ny * F.nz));
    const MgCell a = mg_get(st, F, l, i, j, k);
    if (!a.unknown) return 0.0f;
    const float xc = F.x[c] + kMgCorrScale * corr[idx3(i >> 1, j >> 1, k >> 1, cny, cnz)];
    float lx = a.d * xc;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        if (a.w[q] == 0.0f) continue;
        int ni, nj, nk;
        mg_neighbour(q, i, j, k, ni, nj, nk);
        if (ni < 0 || nj < 0 || nk < 0 || ni >= F.nx || nj >= F.ny || nk >= F.nz) continue;
        const size_t cn = idx3(ni, nj, nk, F.ny, F.nz);
        if (!mg_unknown(st, F, l, cn)) continue;
        lx -= a.w[q] * (F.x[cn] + kMgCorrScale * corr[idx3(ni >> 1, nj >> 1, nk >> 1, cny, cnz)]);
    }
    return xc + kMgOmega * (bc - lx) / a.d;
}
// The same on a level l >= 1 with every load issued at once (as mg_down_cell_coarse): the
// operator records for the cell's weights, its x and correction, and each neighbour's record, x
// and correction from indices clamped into the grid; a neighbour that is outside, not an unknown
// or joined by a zero weight adds its term times 0 (every value finite).
__device__ __forceinline__ float mg_up_cell_coarse(const MgLevel &F, const float *corr, int cny, int cnz, int c,
                                                   float bc) {
    const int k = c % F.nz, j = (c / F.nz) % F.ny, i = c / (F.ny * F.nz);
    const int sx = F.ny * F.nz, sy = F.nz;
    const float4 o = F.op[c];
    const float4 oxm = F.op[i > 0 ? c - sx : c], oym = F.op[j > 0 ? c - sy : c], ozm = F.op[k > 0 ? c - 1 : c];
    const float w[6] = {oxm.y * (float)(i > 0), o.y, oym.z * (float)(j > 0), o.z, ozm.w * (float)(k > 0), o.w};
    const float x0 = F.x[c], cc = corr[((i >> 1) * cny + (j >> 1)) * cnz + (k >> 1)];
    float dn[6], xn[6], crn[6];
    bool in[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        int ni, nj, nk;
        mg_neighbour(q, i, j, k, ni, nj, nk);
        in[q] = ni >= 0 && nj >= 0 && nk >= 0 && ni < F.nx && nj < F.ny && nk < F.nz;
        ni = clampi(ni, 0, F.nx - 1);
        nj = clampi(nj, 0, F.ny - 1);
        nk = clampi(nk, 0, F.nz - 1);
        const int cn = (ni * F.ny + nj) * F.nz + nk;
        dn[q] = F.op[cn].x;
        xn[q] = F.x[cn];
        crn[q] = corr[((ni >> 1) * cny + (nj >> 1)) * cnz + (nk >> 1)];
    }
    if (!(o.x > 0.0f)) return 0.0f;
    const float xc = x0 + kMgCorrScale * cc;
    float lx = o.x * xc;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const bool use = w[q] != 0.0f && in[q] && dn[q] > 0.0f;
        lx -= (w[q] * (xn[q] + kMgCorrScale * crn[q])) * (float)use;
    }
    return xc + kMgOmega * (bc - lx) / o.x;
}
// levels >= 1: one thread per cell, writes xs
__device__ __forceinline__ void mg_up_body(const PcgState &st, const MgLevel &F, const float *corr, int cny, int cnz,
                                           int l, int blk) {
    if (solve_done(st.scal)) return;
    const size_t n = (size_t)F.nx * F.ny * F.nz;
    const size_t c = (size_t)blk * kPB + threadIdx.x;
    if (c < n) F.xs[c] = l >= 1 ? mg_up_cell_coarse(F, corr, cny, cnz, (int)c, F.b[c]) : mg_up_cell(st, F, corr, cny, cnz, l, c, F.b[c]);
}
__global__ void __launch_bounds__(kPB) mg_up_kernel(PcgState st, MgLevel F, const float *corr, int cny, int cnz,
                                                    int l) {
    mg_up_body(st, F, corr, cny, cnz, l, blockIdx.x);
}
// Level 0 of the V-cycle on the solve's tiles (tile_sweep: the 2.5-D sweep of the CG kernels).
// Down: v = x = w r / d (the first smoothing sweep from zero), stored, and the residual r - L x
// into the level's residual buffer (b), summed over the 2 x 2 x 2 aggregates by
// mg_restrict0_kernel.  Up: v = x + the aggregate's correction, z = v + w (r - L v) / d and
// partial sums of r.z.
struct MgDown0Raw {
    unsigned f;
    double r;
};
struct MgDown0Op {
    using Raw = MgDown0Raw;
    const PcgState &st;
    float *x, *res;
    __device__ __forceinline__ Raw load(int x0, int y, int z, bool) const {
        const size_t c = clamped_cell(st.g, x0, y, z);
        const unsigned f = st.flags[c];
        const double r = st.r[c];
        const bool in = in_grid(st.g, x0, y, z);
        return Raw{in ? f : 0u, in ? r : 0.0};
    }
    __device__ __forceinline__ double value(const Raw &a) const {
        return (a.f & kPcgUnknown) ? (double)(kMgOmega * (float)a.r / (float)__builtin_popcount(a.f & 63u)) : 0.0;
    }
    __device__ __forceinline__ void cell(size_t c, const Raw &a, double v, double q, double *) const {
        x[c] = (float)v;
        res[c] = (a.f & kPcgUnknown) ? (float)(a.r - q) : 0.0f;
    }
};
struct MgUp0Raw {
    unsigned f;
    float x, corr;
    double r;
};
struct MgUp0Op {
    using Raw = MgUp0Raw;
    const PcgState &st;
    const float *x, *corr;
    int cny, cnz;
    // the loaded values as stored (x and the correction exist for every cell: loaded with the flag
    // byte, not after it); combined only where the sweep uses them.  Unlike the other sweeps' loads
    // this one keeps its branch: the clamped form needs 140 VGPRs (3 waves per SIMD instead of 4).
    // For the same reason its plane reads stay under selects (the zero-slot reads need 130).
    static constexpr bool kSelectReads = true;
    __device__ __forceinline__ Raw load(int x0, int y, int z, bool own) const {
        Raw a{0u, 0.0f, 0.0f, 0.0};
        if (in_grid(st.g, x0, y, z)) {
            const size_t c = idx3(x0, y, z, st.g.ny, st.g.nz);
            a.x = x[c];
            a.corr = corr[idx3(x0 >> 1, y >> 1, z >> 1, cny, cnz)];
            a.f = st.flags[c];
            if (own) a.r = st.r[c];
        }
        return a;
    }
    __device__ __forceinline__ double value(const Raw &a) const {
        return (a.f & kPcgUnknown) ? (double)(a.x + kMgCorrScale * a.corr) : 0.0;
    }
    __device__ __forceinline__ void cell(size_t c, const Raw &a, double v, double q, double *acc) const {
        double zc = 0.0;
        if (a.f & kPcgUnknown)
            zc = (double)((float)v + kMgOmega * (float)(a.r - q) / (float)__builtin_popcount(a.f & 63u));
        st.z[c] = zc;
        acc[0] += a.r * zc;
    }
};
__device__ __forceinline__ void mg_down0_body(const PcgState &st, float *x, float *res, int blk) {
    if (solve_done(st.scal)) return;
    double acc[1] = {0.0};
    tile_sweep(st, blk, MgDown0Op{st, x, res}, acc);
}
__global__ void __launch_bounds__(kPB) mg_down0_kernel(PcgState st, float *x, float *res) {
    mg_down0_body(st, x, res, blockIdx.x);
}
// b of level 1: the sum of each aggregate's children's residuals (fixed order)
__device__ __forceinline__ void mg_restrict0_body(const PcgState &st, const MgLevel &F, const MgLevel &C, int blk) {
    if (solve_done(st.scal)) return;
    const size_t n = (size_t)C.nx * C.ny * C.nz;
    const size_t I = (size_t)blk * kPB + threadIdx.x;
    if (I >= n) return;
    const int K = (int)(I % C.nz), J = (int)((I / C.nz) % C.ny), II = (int)(I / ((size_t)C.ny * C.nz));
    // the eight children loaded at once from clamped indices, a child outside the grid added
    // times 0 (a load under the bounds test was waited for before the next was issued)
    float v[8], m[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int i = 2 * II + (e >> 2), j = 2 * J + ((e >> 1) & 1), k = 2 * K + (e & 1);
        m[e] = (float)(i < F.nx && j < F.ny && k < F.nz);
        v[e] = F.b[idx3(min(i, F.nx - 1), min(j, F.ny - 1), min(k, F.nz - 1), F.ny, F.nz)];
    }
    float sum = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; ++e) sum += v[e] * m[e];
    C.b[I] = C.op[I].x > 0.0f ? sum : 0.0f;
}
__global__ void __launch_bounds__(kPB) mg_restrict0_kernel(PcgState st, MgLevel F, MgLevel C) {
    mg_restrict0_body(st, F, C, blockIdx.x);
}
__device__ __forceinline__ void mg_up0_body(const PcgState &st, const float *x, const float *corr, int cny, int cnz,
                                            int blk) {
    if (solve_done(st.scal)) return;
    double acc[1] = {0.0};
    tile_sweep(st, blk, MgUp0Op{st, x, corr, cny, cnz}, acc);
    write_partials<1>(acc, st.part_mg, st.dir_blocks, blk);
}
__global__ void __launch_bounds__(kPB) mg_up0_tile_kernel(PcgState st, const float *x, const float *corr, int cny,
                                                          int cnz) {
    mg_up0_body(st, x, corr, cny, cnz, blockIdx.x);
}

// ... and the V-cycles of many solves per launch (fia_scenes_step_batched): map[b] = (solve s,
// block of that solve's launch), each solve's levels in st.mg_lv (st.mg_levels of them); a level
// launch covers the solves that have that level, the coarsest launch is one block per solve.
__global__ void __launch_bounds__(kPB) mg_down0_batch_kernel(const PcgState *__restrict__ sts,
                                                             const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const PcgState &st = sts[m.x];
    mg_down0_body(st, st.mg_lv[0].x, st.mg_lv[0].b, m.y);
}
__global__ void __launch_bounds__(kPB) mg_restrict0_batch_kernel(const PcgState *__restrict__ sts,
                                                                 const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const PcgState &st = sts[m.x];
    mg_restrict0_body(st, st.mg_lv[0], st.mg_lv[1], m.y);
}
__global__ void __launch_bounds__(kPB) mg_down_batch_kernel(const PcgState *__restrict__ sts,
                                                            const int2 *__restrict__ map, int l) {
    const int2 m = map[blockIdx.x];
    const PcgState &st = sts[m.x];
    mg_down_body(st, st.mg_lv[l], st.mg_lv[l + 1], l, m.y);
}
__global__ void __launch_bounds__(kPB) mg_coarsest_batch_kernel(const PcgState *__restrict__ sts) {
    const PcgState &st = sts[blockIdx.x];
    mg_coarsest_body(st, st.mg_lv[st.mg_levels - 1]);
}
__global__ void __launch_bounds__(kPB) mg_up_batch_kernel(const PcgState *__restrict__ sts,
                                                          const int2 *__restrict__ map, int l) {
    const int2 m = map[blockIdx.x];
    const PcgState &st = sts[m.x];
    const MgLevel &C = st.mg_lv[l + 1];
    mg_up_body(st, st.mg_lv[l], l + 1 == st.mg_levels - 1 ? C.x : C.xs, C.ny, C.nz, l, m.y);
}
__global__ void __launch_bounds__(kPB) mg_up0_batch_kernel(const PcgState *__restrict__ sts,
                                                           const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const PcgState &st = sts[m.x];
    const MgLevel &C = st.mg_lv[1];
    mg_up0_body(st, st.mg_lv[0].x, st.mg_levels == 2 ? C.x : C.xs, C.ny, C.nz, m.y);
}

// One solve per launch ...
__global__ void __launch_bounds__(kPB) pcg_start_kernel(PcgState st, double tol2) {
    pcg_start_body(st, blockIdx.x, tol2);
}
__global__ void __launch_bounds__(kPB) pcg_direction_kernel(PcgState st, int k) {
    pcg_direction_body(st, blockIdx.x, k);
}
template <bool RECOMPUTE> __global__ void __launch_bounds__(kPB) pcg_update_kernel(PcgState st, int k) {
    pcg_update_body<RECOMPUTE>(st, blockIdx.x, k);
}
__global__ void __launch_bounds__(kPB) pcg_check_kernel(PcgState st, int k) { pcg_check_body(st, k); }
template <bool XUPD>
__global__ void __launch_bounds__(kPB) __attribute__((amdgpu_waves_per_eu(3))) pcg_fused_kernel(PcgState st, int k) {
    pcg_fused_body<XUPD>(st, blockIdx.x, k);
}

// The x step a fused solve still owes when it stopped after an odd iteration I (device count):
// x += alpha_(I-1) p_(I-1), with p_(I-1) in p[1] and alpha_(I-1) stored by sweep I.
__global__ void __launch_bounds__(kPB) pcg_fused_flush_kernel(PcgState st) {
    const long long iters = (long long)st.scal[kPcgIters];
    if (!(iters & 1)) return;
    const double a = st.scal[kPcgAlpha2];
    const float *__restrict__ p = st.p[1];
    double *__restrict__ x = st.x;
    const size_t n = (size_t)st.g.nx * st.g.ny * st.g.nz;
    const CellRange cr = block_cells(n, st.nblocks, blockIdx.x);
    for (size_t c = cr.first; c < cr.end; c += cr.step) x[c] += a * (double)p[c];
}

// ... or many solves per launch, all at the same iteration: map[b] = (s, j) says that block b of
// the launch is block j of solve s, so each solve's arithmetic is exactly that of its own launch.
__global__ void __launch_bounds__(kPB) pcg_direction_batch_kernel(const PcgState *__restrict__ st,
                                                                  const int2 *__restrict__ map, int k) {
    const int2 m = map[blockIdx.x];
    pcg_direction_body(st[m.x], m.y, k);
}
template <bool RECOMPUTE>
__global__ void __launch_bounds__(kPB) pcg_update_batch_kernel(const PcgState *__restrict__ st,
                                                               const int2 *__restrict__ map, int k) {
    const int2 m = map[blockIdx.x];
    pcg_update_body<RECOMPUTE>(st[m.x], m.y, k);
}
// one block per solve; the solves' polled scalars are then copied to out[s * kPcgScalars ...], so
// that the host reads all of them with one copy
__global__ void __launch_bounds__(kPB) pcg_check_batch_kernel(const PcgState *__restrict__ st, int k, double *out) {
    pcg_check_body(st[blockIdx.x], k);
    if (threadIdx.x == 0) {   // the thread that wrote the scalars in pcg_check_body
        const double *sc = st[blockIdx.x].scal;
        for (int i = 0; i < kPcgScalars; ++i) out[(size_t)blockIdx.x * kPcgScalars + i] = sc[i];
    }
}

static inline unsigned blocks(size_t n) { return (unsigned)((n + kPB - 1) / kPB); }

// Blocks of the start and update kernels: about kPcgCellsPerBlock cells each (every block of
// the next kernel re-sums all partials, so small grids get few blocks), at most kPcgMaxBlocks,
// a multiple of 8 (one share per XCD).
constexpr size_t kPcgCellsPerBlock = 2048;
static size_t target_blocks(size_t n) {
    return std::min<size_t>(kPcgMaxBlocks, std::max<size_t>(8, (n + kPcgCellsPerBlock - 1) / kPcgCellsPerBlock));
}
int pcg_blocks(size_t n) {
    const size_t b = target_blocks(n);
    return (int)((b + 7) / 8 * 8);
}

// Blocks of the tile sweeps resident on the device at once: blocks per CU that the registers and
// LDS allow (the fused kernel's, or the lesser of the direction and recomputing update kernels'),
// times the CUs; per device and form, computed once.  0 when the runtime cannot say.
int pcg_resident_blocks(bool fused) {
    static std::atomic<int> cache[64][2];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    const int known = cache[dev][fused].load();
    if (known > 0) return known;
    int cus = 0, a = 0, b = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (fused) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, pcg_fused_kernel<true>, kPB, 0) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, pcg_fused_kernel<false>, kPB, 0) != hipSuccess)
            return 0;
    } else if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, pcg_direction_kernel, kPB, 0) != hipSuccess ||
               hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, pcg_update_kernel<true>, kPB, 0) != hipSuccess) {
        return 0;
    }
    const int v = std::max(1, std::min(a, b)) * std::max(1, cus);
    cache[dev][fused].store(v);
    return v;
}

// Tiling of the tile sweeps: kTileZ x kTileY column tiles, x cut into chunks so that the launch
// has about as many blocks as fit on the device at once (one full round: on 256^3 a 22-plane
// chunk, 768 blocks at 3 per CU, ran the iteration in 191-203 us against 236-246 us for 16-plane
// chunks and 1024 blocks, whose second round ran a third full), else as target_blocks; rounded
// up to a multiple of 8.
void pcg_direction_tiling(int nx, int ny, int nz, PcgState &st, int chunk_x, int recompute, bool shared) {
    st.dir_tiles_y = (ny + kTileY - 1) / kTileY;
    st.dir_tiles_z = (nz + kTileZ - 1) / kTileZ;
    const long tiles = (long)st.dir_tiles_y * st.dir_tiles_z;
    // a solve that shares its launches with others (batched scenes) has no round of its own to fill
    const int resident = shared ? 0 : pcg_resident_blocks(st.fused);
    // chunks per tile column: as many as fit in one round of resident blocks (rounded down: a few
    // blocks over the round would run alone in a second one)
    const long want = resident > 0 ? std::max<long>(1, std::min<long>(nx, resident / tiles))
                                   : std::max<long>(1, std::min<long>(nx, ((long)target_blocks((size_t)nx * ny * nz) +
                                                                           tiles - 1) / tiles));
    st.dir_chunk_x = chunk_x > 0 ? std::min(chunk_x, nx) : (int)((nx + want - 1) / want);
    st.dir_chunks = (nx + st.dir_chunk_x - 1) / st.dir_chunk_x;
    long nb = tiles * st.dir_chunks;
    if (nb > (long)1 << 30) {   // a chunk too short for the grid: fall back to the sized one
        st.dir_chunk_x = (int)((nx + want - 1) / want);
        st.dir_chunks = (nx + st.dir_chunk_x - 1) / st.dir_chunk_x;
        nb = tiles * st.dir_chunks;
    }
    st.dir_blocks = (int)((nb + 7) / 8 * 8);
    // the update recomputes q on these tiles on large grids with long chunks
    st.recompute_q = recompute == 0 || recompute == 1
                         ? recompute == 1
                         : (size_t)nx * ny * nz >= kRecomputeCells && st.dir_chunk_x >= kRecomputeChunk;
    st.upd_blocks = st.recompute_q ? st.dir_blocks : pcg_blocks((size_t)nx * ny * nz);
}

hipError_t launch_pcg_setup(const PcgGrid &g, const float *u, const float *v, const float *w, double h, double *b,
                            uint8_t *flags, hipStream_t s) {
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(pcg_setup_kernel, dim3(blocks(n)), dim3(kPB), 0, s, g, u, v, w, h, b, flags);
    return hipGetLastError();
}
hipError_t launch_pcg_guess(const PcgState &st, const float *p_in, hipStream_t s) {
    const size_t n = (size_t)st.g.nx * st.g.ny * st.g.nz;
    hipLaunchKernelGGL(pcg_guess_kernel, dim3(blocks(n)), dim3(kPB), 0, s, st, p_in);
    return hipGetLastError();
}
hipError_t launch_pcg_scale_guess(const PcgState &st, hipStream_t s) {
    hipLaunchKernelGGL(pcg_energy_kernel, dim3(st.nblocks), dim3(kPB), 0, s, st);
    hipLaunchKernelGGL(pcg_scale_kernel, dim3(st.nblocks), dim3(kPB), 0, s, st);
    return hipGetLastError();
}
hipError_t launch_pcg_start(PcgState &st, double tol2, hipStream_t s) {
    st.it = 0;
    hipLaunchKernelGGL(pcg_start_kernel, dim3(st.nblocks), dim3(kPB), 0, s, st, tol2);
    return hipGetLastError();
}
hipError_t launch_pcg_iterations(PcgState &st, int count, hipStream_t s, const MgHier *mg) {
    if (st.fused) {
        if (mg) return hipErrorInvalidValue;   // the fused sweep is the Jacobi iteration
        for (int k = 0; k < count; ++k) {
            // x steps on even sweeps from 2 on (a captured batch never holds sweep 0)
            if (st.it >= 2 && !(st.it & 1))
                hipLaunchKernelGGL(pcg_fused_kernel<true>, dim3(st.dir_blocks), dim3(kPB), 0, s, st, st.it);
            else
                hipLaunchKernelGGL(pcg_fused_kernel<false>, dim3(st.dir_blocks), dim3(kPB), 0, s, st, st.it);
            ++st.it;
        }
        return hipGetLastError();
    }
    for (int k = 0; k < count; ++k) {
        hipLaunchKernelGGL(pcg_direction_kernel, dim3(st.dir_blocks), dim3(kPB), 0, s, st, st.it);
        if (st.recompute_q) hipLaunchKernelGGL(pcg_update_kernel<true>, dim3(st.upd_blocks), dim3(kPB), 0, s, st, st.it);
        else hipLaunchKernelGGL(pcg_update_kernel<false>, dim3(st.upd_blocks), dim3(kPB), 0, s, st, st.it);
        if (mg) {
            const hipError_t e = launch_mg_vcycle(st, *mg, s);
            if (e != hipSuccess) return e;
        }
        ++st.it;
    }
    return hipGetLastError();
}
hipError_t launch_pcg_fused_flush(const PcgState &st, hipStream_t s) {
    if (!st.fused) return hipSuccess;
    hipLaunchKernelGGL(pcg_fused_flush_kernel, dim3(st.nblocks), dim3(kPB), 0, s, st);
    return hipGetLastError();
}
hipError_t launch_pcg_check(const PcgState &st, hipStream_t s) {
    hipLaunchKernelGGL(pcg_check_kernel, dim3(1), dim3(kPB), 0, s, st, st.it);
    return hipGetLastError();
}
hipError_t launch_pcg_batch_iterations(PcgBatch &b, int count, hipStream_t s, const MgBatch *mb) {
    if (b.n == 0) return hipSuccess;
    for (int k = 0; k < count; ++k) {
        hipLaunchKernelGGL(pcg_direction_batch_kernel, dim3(b.dir_blocks), dim3(kPB), 0, s, b.states, b.dir_map, b.it);
        if (b.stream_blocks)
            hipLaunchKernelGGL(pcg_update_batch_kernel<false>, dim3(b.stream_blocks), dim3(kPB), 0, s, b.states,
                               b.stream_map, b.it);
        if (b.tile_blocks)
            hipLaunchKernelGGL(pcg_update_batch_kernel<true>, dim3(b.tile_blocks), dim3(kPB), 0, s, b.states, b.tile_map,
                               b.it);
        if (mb) {
            const hipError_t e = launch_mg_batch_vcycle(b, *mb, s);
            if (e != hipSuccess) return e;
        }
        ++b.it;
    }
    return hipGetLastError();
}
hipError_t launch_pcg_batch_check(const PcgBatch &b, double *out, hipStream_t s) {
    if (b.n == 0) return hipSuccess;
    hipLaunchKernelGGL(pcg_check_batch_kernel, dim3(b.n), dim3(kPB), 0, s, b.states, b.it, out);
    return hipGetLastError();
}
// ---- graph replay of iteration batches ----------------------------------------------------------

// Field-wise equality of everything the iteration launches pass (the iteration index excepted).
static bool same_launch(const PcgState &a, const PcgState &b) {
    return a.g.nx == b.g.nx && a.g.ny == b.g.ny && a.g.nz == b.g.nz && a.g.solid == b.g.solid && a.g.pin == b.g.pin &&
           a.flags == b.flags && a.b == b.b && a.x == b.x && a.r == b.r && a.r_alt == b.r_alt && a.p[0] == b.p[0] &&
           a.p[1] == b.p[1] &&
           a.q == b.q && a.scal == b.scal && a.part_start == b.part_start && a.part_dir == b.part_dir &&
           a.part_update == b.part_update && a.part_fused == b.part_fused && a.fused == b.fused &&
           a.nblocks == b.nblocks && a.dir_blocks == b.dir_blocks &&
           a.dir_tiles_y == b.dir_tiles_y && a.dir_tiles_z == b.dir_tiles_z && a.dir_chunk_x == b.dir_chunk_x &&
           a.dir_chunks == b.dir_chunks && a.recompute_q == b.recompute_q && a.upd_blocks == b.upd_blocks &&
           a.z == b.z && a.part_mg == b.part_mg && a.mg_lv == b.mg_lv && a.mg_levels == b.mg_levels;
}
static bool same_launch(const MgHier &a, const MgHier &b) {
    if (a.levels != b.levels) return false;
    for (int l = 0; l < a.levels; ++l) {
        const MgLevel &x = a.lv[l], &y = b.lv[l];
        if (x.nx != y.nx || x.ny != y.ny || x.nz != y.nz || x.op != y.op || x.b != y.b || x.x != y.x || x.xs != y.xs)
            return false;
    }
    return true;
}
static bool same_launch(const PcgBatch &a, const PcgBatch &b) {
    return a.states == b.states && a.dir_map == b.dir_map && a.stream_map == b.stream_map && a.tile_map == b.tile_map &&
           a.n == b.n && a.dir_blocks == b.dir_blocks && a.stream_blocks == b.stream_blocks &&
           a.tile_blocks == b.tile_blocks;
}
static bool same_launch(const MgBatch &a, const MgBatch &b) {
    if (a.r0_map != b.r0_map || a.r0_blocks != b.r0_blocks) return false;
    for (int l = 0; l < kMgMaxLevels; ++l)
        if (a.down_map[l] != b.down_map[l] || a.down_blocks[l] != b.down_blocks[l] || a.up_map[l] != b.up_map[l] ||
            a.up_blocks[l] != b.up_blocks[l])
            return false;
    return true;
}

void PcgGraphCache::release() {
    for (Entry &e : slots_) {
        if (e.exec) (void)hipGraphExecDestroy(e.exec);
        e = Entry{};
    }
}

PcgGraphCache::Entry *PcgGraphCache::find(const Entry &k) {
    for (Entry &e : slots_) {
        if (!e.exec || e.batch != k.batch || e.count != k.count || e.parity != k.parity || e.has_mg != k.has_mg) continue;
        if (k.batch ? same_launch(e.b, k.b) && (!k.has_mg || same_launch(e.mb, k.mb))
                    : same_launch(e.st, k.st) && (!k.has_mg || same_launch(e.mg, k.mg)))
            return &e;
    }
    return nullptr;
}

// Captures key's batch on stream s (thread-local capture: other threads' contexts keep launching),
// stores the executable graph in the least recently used slot and launches it.
hipError_t PcgGraphCache::capture_and_launch(Entry &key, hipStream_t s) {
    Entry *slot = &slots_[0];
    for (Entry &e : slots_)
        if (!e.exec) { slot = &e; break; } else if (e.used < slot->used) slot = &e;
    if (slot->exec) (void)hipGraphExecDestroy(slot->exec);
    *slot = Entry{};
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    if (key.batch) {
        PcgBatch b = key.b;
        e = launch_pcg_batch_iterations(b, key.count, s, key.has_mg ? &key.mb : nullptr);
    } else {
        PcgState st = key.st;
        e = launch_pcg_iterations(st, key.count, s, key.has_mg ? &key.mg : nullptr);
    }
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(s, &graph);
    if (e == hipSuccess) e = e2;
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return e;
    *slot = key;
    slot->exec = exec;
    slot->used = ++tick_;
    ++captures;
    ++replays;
    return hipGraphLaunch(exec, s);
}

hipError_t PcgGraphCache::iterations(PcgState &st, int count, hipStream_t s, const MgHier *mg) {
    if (!enabled || st.it == 0 || s == nullptr || count <= 0) return launch_pcg_iterations(st, count, s, mg);
    Entry key;
    key.batch = false;
    key.count = count;
    key.parity = st.it & 1;
    key.st = st;
    key.has_mg = mg != nullptr;
    if (mg) key.mg = *mg;
    hipError_t e;
    if (Entry *hit = find(key)) {
        hit->used = ++tick_;
        ++replays;
        e = hipGraphLaunch(hit->exec, s);
    } else {
        e = capture_and_launch(key, s);
    }
    if (e == hipSuccess) st.it += count;
    return e;
}

hipError_t PcgGraphCache::batch_iterations(PcgBatch &b, int count, hipStream_t s, const MgBatch *mb) {
    if (!enabled || b.it == 0 || s == nullptr || count <= 0 || b.n == 0)
        return launch_pcg_batch_iterations(b, count, s, mb);
    Entry key;
    key.batch = true;
    key.count = count;
    key.parity = b.it & 1;
    key.b = b;
    key.has_mg = mb != nullptr;
    if (mb) key.mb = *mb;
    hipError_t e;
    if (Entry *hit = find(key)) {
        hit->used = ++tick_;
        ++replays;
        e = hipGraphLaunch(hit->exec, s);
    } else {
        e = capture_and_launch(key, s);
    }
    if (e == hipSuccess) b.it += count;
    return e;
}

void mg_shape(int nx, int ny, int nz, MgHier &h) {
    h.levels = 0;
    int x = nx, y = ny, z = nz;
    while (h.levels < kMgMaxLevels) {
        MgLevel &L = h.lv[h.levels++];
        L.nx = x; L.ny = y; L.nz = z;
        L.op = nullptr; L.b = nullptr; L.x = nullptr; L.xs = nullptr;
        if (h.levels >= 2 && (size_t)x * y * z <= (size_t)kMgCoarsestCells) break;
        x = (x + 1) / 2; y = (y + 1) / 2; z = (z + 1) / 2;
    }
}
hipError_t launch_mg_build(const PcgState &st, const MgHier &h, hipStream_t s) {
    for (int l = 0; l + 1 < h.levels; ++l) {
        const size_t n = (size_t)h.lv[l + 1].nx * h.lv[l + 1].ny * h.lv[l + 1].nz;
        hipLaunchKernelGGL(mg_build_kernel, dim3(blocks(n)), dim3(kPB), 0, s, st, h.lv[l], h.lv[l + 1], l);
    }
    return hipGetLastError();
}
hipError_t launch_mg_vcycle(const PcgState &st, const MgHier &h, hipStream_t s) {
    const int L = h.levels;
    hipLaunchKernelGGL(mg_down0_kernel, dim3(st.dir_blocks), dim3(kPB), 0, s, st, h.lv[0].x, h.lv[0].b);
    {
        const size_t n1 = (size_t)h.lv[1].nx * h.lv[1].ny * h.lv[1].nz;
        hipLaunchKernelGGL(mg_restrict0_kernel, dim3(blocks(n1)), dim3(kPB), 0, s, st, h.lv[0], h.lv[1]);
    }
    for (int l = 1; l + 1 < L; ++l) {
        const MgLevel &F = h.lv[l];
        hipLaunchKernelGGL(mg_down_kernel, dim3((unsigned)mg_down_tiles(F)), dim3(kPB), 0, s, st, F, h.lv[l + 1], l);
    }
    hipLaunchKernelGGL(mg_coarsest_kernel, dim3(1), dim3(kPB), 0, s, st, h.lv[L - 1]);
    for (int l = L - 2; l >= 0; --l) {
        const MgLevel &C = h.lv[l + 1];
        const float *corr = l + 1 == L - 1 ? C.x : C.xs;
        if (l == 0) {
            hipLaunchKernelGGL(mg_up0_tile_kernel, dim3(st.dir_blocks), dim3(kPB), 0, s, st, h.lv[0].x, corr, C.ny,
                               C.nz);
        } else {
            const size_t n = (size_t)h.lv[l].nx * h.lv[l].ny * h.lv[l].nz;
            hipLaunchKernelGGL(mg_up_kernel, dim3(blocks(n)), dim3(kPB), 0, s, st, h.lv[l], corr, C.ny, C.nz, l);
        }
    }
    return hipGetLastError();
}
hipError_t launch_mg_batch_vcycle(const PcgBatch &b, const MgBatch &mb, hipStream_t s) {
    if (b.n == 0) return hipSuccess;
    hipLaunchKernelGGL(mg_down0_batch_kernel, dim3(b.dir_blocks), dim3(kPB), 0, s, b.states, b.dir_map);
    hipLaunchKernelGGL(mg_restrict0_batch_kernel, dim3(mb.r0_blocks), dim3(kPB), 0, s, b.states, mb.r0_map);
    for (int l = 1; l < kMgMaxLevels; ++l)
        if (mb.down_blocks[l])
            hipLaunchKernelGGL(mg_down_batch_kernel, dim3(mb.down_blocks[l]), dim3(kPB), 0, s, b.states, mb.down_map[l], l);
    hipLaunchKernelGGL(mg_coarsest_batch_kernel, dim3(b.n), dim3(kPB), 0, s, b.states);
    for (int l = kMgMaxLevels - 1; l >= 1; --l)
        if (mb.up_blocks[l])
            hipLaunchKernelGGL(mg_up_batch_kernel, dim3(mb.up_blocks[l]), dim3(kPB), 0, s, b.states, mb.up_map[l], l);
    hipLaunchKernelGGL(mg_up0_batch_kernel, dim3(b.dir_blocks), dim3(kPB), 0, s, b.states, b.dir_map);
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

}  // namespace fia
