// fia_internal.hpp - kernel parameter blocks and launchers shared by fia_kernels.hip and
// fia_api.hip.  Not part of the public ABI (include/fia.h is).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>

#include "../../include/fia.h"

namespace fia {

// Device buffer that grows on demand (contents are not kept when it grows).
template <class T> struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t want) {
        if (want <= n && p) return hipSuccess;
        if (p) { hipError_t e = hipFree(p); if (e != hipSuccess) return e; p = nullptr; n = 0; }
        hipError_t e = hipMalloc((void **)&p, std::max<size_t>(want, 1) * sizeof(T));
        if (e == hipSuccess) n = want;
        return e;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
    size_t bytes() const { return n * sizeof(T); }
};


constexpr int kMaxGauss = 49;
// work counters are kept as kStatSlots triples (pieces, density points, 1D WENO6 evaluations),
// summed on read
constexpr int kStatSlots = 1024;
constexpr int kStatWords = 3;

// n / d for 0 <= n < 2^31 by a multiply and a shift (unsigned division by an invariant divisor,
// Granlund-Montgomery / Hacker's Delight 10-9): with l = ceil(log2 d), p = 31 + l and
// m = ceil(2^p / d), 2^p <= m d < 2^p + 2^l, so floor(n m / 2^p) = floor(n / d), and m < 2^32.
struct FastDiv {
    uint32_t d = 1, m = 0x80000000u;
    int p = 31;
    FastDiv() = default;
    __host__ __device__ explicit FastDiv(uint32_t dd) : d(dd) {
        int l = 0;
        while (((uint64_t)1 << l) < dd) ++l;
        p = 31 + l;
        m = (uint32_t)((((uint64_t)1 << p) + dd - 1) / dd);
    }
    __host__ __device__ uint32_t div(uint32_t n) const { return (uint32_t)(((uint64_t)n * m) >> p); }
};

// A component grid G: density (comp = -1, dims n) or one MAC velocity component (comp = 0..2,
// dims n + e_comp).  Its vertex lattice has dims m + 1 and sits at lattice - e_comp/2 (cells).
struct CompGrid {
    int comp;
    int n[3];   // base grid
    int m[3];   // G dims
    int e[3];
    FastDiv fy[3], fz[3];   // face family f: its y and z extents (face ids decode by two divisions)
    __host__ __device__ void init(int c, int nx, int ny, int nz) {
        comp = c; n[0] = nx; n[1] = ny; n[2] = nz;
        for (int a = 0; a < 3; ++a) { e[a] = (a == c); m[a] = n[a] + e[a]; }
        for (int f = 0; f < 3; ++f) {
            fy[f] = FastDiv((uint32_t)(m[1] + (f == 1)));
            fz[f] = FastDiv((uint32_t)(m[2] + (f == 2)));
        }
    }
    __host__ __device__ size_t cells() const { return (size_t)m[0] * m[1] * m[2]; }
    __host__ __device__ size_t lattice() const { return (size_t)(m[0] + 1) * (m[1] + 1) * (m[2] + 1); }
    __host__ __device__ size_t psi_size() const { return (size_t)m[0] * (m[1] + 1) * m[2]; }
    __host__ __device__ size_t faces(int fam) const {
        return (size_t)(m[0] + (fam == 0)) * (m[1] + (fam == 1)) * (m[2] + (fam == 2));
    }
};

struct GaussTable {
    int n;
    double x[kMaxGauss];   // nodes on [-1, 1]
    double w[kMaxGauss];   // weights, sum 1
};

struct PsiParams {
    CompGrid g;
    const float *q;
    double *psi;
    int mode;       // 0 const, 1 linear, 2 WENO6
    int integral;
    float kappa;    // WENO smoothness scale for q (physical units: 1)
    GaussTable gauss;
};

struct VertexParams {
    CompGrid g;
    const float *u, *v, *w;
    float dt_over_h;
    int interp;
    float *disp[3];
};

struct FaceParams {
    CompGrid g;
    const float *q;          // advected field on G (density points)
    const double *psi;       // psi on the y-faces of G, in units of q * cells
    const float *disp[3];    // vertex displacements (cells); null = zero velocity
    double *flux[3];         // per family face fluxes (normalised: q * cells^3 / cells^3)
    double *corr[3];         // per family: (upper cell's own evaluation) + flux for the faces the
                             // general cutter handled, whose bit is set in corr_mask; zero elsewhere
    unsigned *corr_mask[3];  // per family: bit id of word id / 32 set where corr holds a correction
    unsigned *twisted;       // list of faces for the general cutter: fam << 30 | index (unsigned)
    int *twisted_count;
    unsigned *degen;         // degenerate faces (segment cutter): cut per face by the clipping pipeline
    int *degen_count;
    int *error;              // first error code seen (0 = none)
    int split, interp, integral;
    int weno_opt;            // interp IM_WENO6 with the "WENO6-optimized" quadrature interpolation
    int fam_mask;            // families computed (bit 0 x, 1 y, 2 z)
    int general_all;         // every face goes through the general cutter (minimal_area < 1)
    int fuse_value;          // the cutters store piece values, not records (non-integral rule with
                             // linear or 1Dy_linear psi: one interpolation per piece, no quadrature
                             // launch)
    float kappa_q, kappa_psi;
    double min_area;
    unsigned long long *stats;   // optional work counters: kStatSlots x kStatWords
    // phased face path: per face slot counts and starts over the three families (fam_base
    // offsets), piece records (centre in cells, y area) and piece values
    unsigned *face_n;
    unsigned long long *face_start;
    size_t fam_base[3];
    double *piece_c[3];
    double *piece_ay;
    double *piece_val;
    // twisted faces (general cutter): pieces per face and orientation (2 per listed face), their
    // exclusive scan (2 per face, then the total), and the piece records and values
    unsigned *gen_n;
    unsigned long long *gen_start;
    double *gen_c[3];
    double *gen_ay;
    double *gen_val;
    GaussTable gauss;
    // threads per block of the face kernels (count, cutter, sum, zero velocity) and of the
    // quadrature kernel: 64, 128 or 256, anything else 256.  Each face and piece is one thread's
    // work, so the results do not depend on them.
    int face_block, piece_block;
};
// the general cutter's piece budget of one flux pass (records of 40 bytes): more pieces than this
// report FIA_ERR_CAPACITY instead of running
constexpr unsigned long long kGenMaxPieces = 1ull << 25;

// errors reported by kernels through FaceParams::error
// KE_CAPACITY: a caller's piece array (fia_split_faces); KE_CAP_*: the general cutter's polygon
// vertices, stack of pending polygons, list of pending triangles
enum KernelError {
    KE_NONE = 0, KE_CAPACITY = 3, KE_NONFINITE = 4, KE_CAP_VERTS = 5, KE_CAP_STACK = 6, KE_CAP_TRIS = 7, KE_CUT = 8
};

hipError_t launch_psi(const PsiParams &p, hipStream_t s);
hipError_t launch_vertices(const VertexParams &p, hipStream_t s);
hipError_t launch_face_count(const FaceParams &p, hipStream_t s);
// segment cutter: the face kernels of the three families (pieces into the slots of the scan)
hipError_t launch_seg_faces(const FaceParams &p, hipStream_t s);
// 1 or 2: the WENO6 integral quadrature with that many Gauss points per half (piece_kernel<1|2>);
// 0: the general quadrature (piece_kernel<0>)
int piece_quadrature_mode(const FaceParams &p);
hipError_t launch_pieces(const FaceParams &p, unsigned long long nslots, hipStream_t s);
hipError_t launch_face_sum(const FaceParams &p, hipStream_t s);
// exclusive scan of the slot counts in (n faces) into out; bounds[1], bounds[2] = first slots of
// families 1 and 2 (at p.fam_base), bounds[3] = total; tile_tmp holds scan_temp_words(n) words
size_t scan_temp_words(size_t n);
hipError_t launch_scan(unsigned long long *tile_tmp, const unsigned *in, unsigned long long *out, size_t n,
                       const FaceParams &p, unsigned long long *bounds, hipStream_t s);
// the twisted faces' pieces: count per face and orientation, and their scan (p.gen_n, p.gen_start;
// gen_start[2 count] = the total), then with the piece records sized from that total: the pieces
// stored, integrated and summed into the faces' flux and corr (count: the listed faces)
hipError_t launch_general_count(const FaceParams &p, int count, hipStream_t s);
hipError_t launch_general_pieces(const FaceParams &p, int count, unsigned long long npieces, hipStream_t s);
// degenerate faces of the segment-cutter mode, cut per face into their slots by the pipeline
hipError_t launch_faces_degenerate(const FaceParams &p, int count, hipStream_t s);
// zero-velocity pass: fluxes of the y faces of undeformed lattice faces (minimal_area >= 1)
hipError_t launch_zero_velocity_faces(const FaceParams &p, hipStream_t s);
// out = sum over families of flux differences (scaled); mode 0: out = G1; 1: out = G1 + G2;
// 2: out = base - G1 (y family only).
// block: cells per block (64, 128 or 256, anything else 256)
// a flux set's fluxes, corrections and correction masks per family (FluxSet on the device)
struct FluxView {
    double *f[3], *c[3];
    unsigned *m[3];
};
hipError_t launch_gather(const CompGrid &g, const FluxView &s1, const FluxView *s2, const float *base, float *out,
                         int mode, int fam_mask, hipStream_t s, int block = 0);
// the corrections of the last pass back to zero (the entries whose mask bit is set) and the masks
// cleared: words[a] mask words of family a
hipError_t launch_corr_reset(const FluxView &v, const size_t words[3], hipStream_t s);
hipError_t launch_interpolate(const float *field, int mx, int my, int mz, double ox, double oy, double oz,
                              int method, float kappa, const float *pts_cells, size_t n, float *out,
                              hipStream_t s);
hipError_t launch_split_faces(const double *quads_cells, size_t n, int split, double min_area, int max_pieces,
                              int *count, double *cen, double *area, int *error, hipStream_t s);
hipError_t launch_absmax(const float *a, size_t n, unsigned int *out_bits, hipStream_t s);

// Jacobi-PCG (FP64 vectors, FP64 reductions in a fixed order)
struct PcgGrid {
    int nx, ny, nz;
    const uint8_t *solid;   // may be null
    long pin;               // pinned cell (p = 0)
};
// per-cell flag byte: bits 0-5 = the neighbour at -x, +x, -y, +y, -z, +z is fluid; this bit = the
// cell is an unknown of the system (fluid, not pinned, with at least one fluid neighbour)
constexpr unsigned kPcgUnknown = 0x80u;
constexpr int kPcgMaxBlocks = 1024;
// device scalars of one solve: rz of even and odd iterations, last residual r.r, b.b, tol^2,
// converged flag, iterations done, scale applied to the initial guess, direction launches that
// have run (the iteration index as the device counts it: a replayed graph passes the index it was
// captured with), alpha of the x step an odd fused sweep leaves to the next one
enum {
    kPcgRZ = 0, kPcgRR = 3, kPcgBB = 4, kPcgTol2 = 5, kPcgDone = 6, kPcgIters = 7, kPcgGamma = 8, kPcgNext = 9,
    kPcgAlpha2 = 10, kPcgScalars = 16
};
struct MgLevel;
// one CG solve in flight: its vectors and device scalars
struct PcgState {
    PcgGrid g;
    const uint8_t *flags;
    const double *b;
    double *x, *r;
    double *r_alt;        // fused solves: the residual of odd sweeps (a sweep reads the residual of
                          // neighbouring tiles, so it cannot be updated in place)
    float *p[2];          // the search direction (FP32: it only steers the solve; x and r are FP64)
    double *q;            // L p stored by the direction kernel (streaming update), or null (recompute)
    double *scal;         // kPcgScalars doubles
    double *part_start;   // 3 * nblocks partial sums (start kernel: rz, rr, bb)
    double *part_dir;     // dir_blocks (direction kernel: p.q)
    double *part_update;  // 2 * upd_blocks (update kernel: rz, rr)
    double *part_fused;   // fused solves: 2 x 5 * dir_blocks (sweeps of even and odd k)
    int nblocks;          // pcg_blocks(n): start kernel, guess scaling, streaming update
    int dir_blocks, dir_tiles_y, dir_tiles_z, dir_chunk_x, dir_chunks;   // pcg_direction_tiling
    bool recompute_q;     // pcg_direction_tiling: the update recomputes q on the direction tiles
    int upd_blocks;       // blocks of the update kernel: dir_blocks (recompute) or nblocks (streaming)
    bool fused;           // Jacobi only: one sweep per iteration (pcg_fused_kernel) instead of the
                          // direction and update kernels; launch k is sweep k, and sweep k >= 1
                          // runs iteration k after testing iteration k - 1
    // multigrid preconditioner (null for Jacobi): z = M r written by the last V-cycle kernel, with
    // dir_blocks partial sums of r.z in part_mg
    double *z;
    double *part_mg;
    const MgLevel *mg_lv;   // the solve's levels in device memory (batched V-cycles), mg_levels of them
    int mg_levels;
    int it;             // iterations enqueued since the start (selects the p buffer)
};
int pcg_blocks(size_t n);
int pcg_resident_blocks(bool fused);
// chunk_x > 0: the x planes one tile sweeps (else one round of the device's resident blocks of
// st.fused's form, or about 2048 cells per block for a solve sharing the device: shared);
// recompute 0 / 1: the update streams q / recomputes it (else chosen by grid size and chunk)
void pcg_direction_tiling(int nx, int ny, int nz, PcgState &st, int chunk_x = 0, int recompute = -1,
                          bool shared = false);

// Geometric multigrid preconditioner for the pressure system: level 0 is the system itself (flag
// bytes), level l + 1 aggregates 2 x 2 x 2 cells of level l with the exact Galerkin operator of
// piecewise-constant prolongation (coarse face weight = fluid face pairs between the aggregates,
// diagonal keeping the pinned cell's coupling), stored as float4 {diag, w+x, w+y, w+z}; a cell
// with diag 0 is not an unknown.  One V(1,1) cycle with damped Jacobi (symmetric, so CG keeps its
// guarantees) and a fixed number of Jacobi sweeps on the coarsest level.
constexpr int kMgMaxLevels = 12;
struct MgLevel {
    int nx, ny, nz;
    float4 *op;       // levels >= 1
    float *b;         // right-hand side (level 0: the residual of the first sweep; it reads r)
    float *x;         // pre-smoothed iterate (the coarsest level: its solution)
    float *xs;        // post-smoothed iterate (levels 1 .. L-2)
};
struct MgHier {
    int levels = 0;
    MgLevel lv[kMgMaxLevels];
};
// the down step of levels >= 1 runs on tiles of 2 x 2 x kMgTileZ cells
constexpr int kMgTileZ = 64;
__host__ __device__ inline int mg_down_tiles(const MgLevel &F) {
    return ((F.nz + kMgTileZ - 1) / kMgTileZ) * ((F.ny + 1) / 2) * ((F.nx + 1) / 2);
}
// block maps (solve, block of that solve's launch) of a batched V-cycle: level 1's restriction,
// and per level l >= 1 the down and up launches of the solves that have level l + 1
struct MgBatch {
    const int2 *r0_map;
    int r0_blocks;
    const int2 *down_map[kMgMaxLevels];
    int down_blocks[kMgMaxLevels];
    const int2 *up_map[kMgMaxLevels];
    int up_blocks[kMgMaxLevels];
};
// dims of every level for a grid (coarsest: at most 256 cells or kMgMaxLevels levels)
void mg_shape(int nx, int ny, int nz, MgHier &h);
// coarse operators of levels 1 .. L-1 from the solve's flag bytes
hipError_t launch_mg_build(const PcgState &st, const MgHier &h, hipStream_t s);
// z = M r and the partial sums of r.z (a no-op once the solve has converged)
hipError_t launch_mg_vcycle(const PcgState &st, const MgHier &h, hipStream_t s);
hipError_t launch_pcg_setup(const PcgGrid &g, const float *u, const float *v, const float *w, double h,
                            double *b, uint8_t *flags, hipStream_t s);
// x = p_in on the unknowns (0 elsewhere); p_in null: x = 0
hipError_t launch_pcg_guess(const PcgState &st, const float *p_in, hipStream_t s);
// r = b - L x, scalars; then `count` iterations (each a no-op once converged); the check updates
// the polled scalars (residual, iterations, converged flag) for the iterations enqueued so far
// x *= x.b / x.L x (the energy-minimising multiple of the initial guess in x)
hipError_t launch_pcg_scale_guess(const PcgState &st, hipStream_t s);
hipError_t launch_pcg_start(PcgState &st, double tol2, hipStream_t s);
hipError_t launch_pcg_iterations(PcgState &st, int count, hipStream_t s, const MgHier *mg = nullptr);
hipError_t launch_pcg_check(const PcgState &st, hipStream_t s);
// after a fused solve's last check: the x step it owes (nothing for the two-sweep form)
hipError_t launch_pcg_fused_flush(const PcgState &st, hipStream_t s);
// Many started solves advanced together (all at the same iteration), one launch per kernel and
// iteration: their states and, per block of the launches, (solve, block of that solve), all in
// device memory.
struct PcgBatch {
    const PcgState *states;
    const int2 *dir_map;      // dir_blocks entries (solve, block)
    const int2 *stream_map;   // update blocks of the streaming solves
    const int2 *tile_map;     // update blocks of the recomputing solves
    int n, dir_blocks, stream_blocks, tile_blocks;
    int it;
};
// partial-sum words of a solve: 3 nblocks (start) + dir_blocks (p.q) + 2 upd_blocks (rz, rr)
// (+ 10 dir_blocks for the fused sweeps)
inline size_t pcg_partial_words(const PcgState &st) {
    return 3 * (size_t)st.nblocks + (size_t)st.dir_blocks + 2 * (size_t)st.upd_blocks +
           (st.fused ? 10 * (size_t)st.dir_blocks : 0);
}
// CG iterations completed after st.it launches, and the launches that complete `iters` more
inline int pcg_iters_done(const PcgState &st) { return st.fused ? std::max(0, st.it - 1) : st.it; }
inline int pcg_launches_for(const PcgState &st, int iters) {
    return iters + (st.fused && st.it == 0 && iters > 0 ? 1 : 0);
}
hipError_t launch_pcg_batch_iterations(PcgBatch &b, int count, hipStream_t s, const MgBatch *mb = nullptr);
hipError_t launch_mg_batch_vcycle(const PcgBatch &b, const MgBatch &mb, hipStream_t s);
// the convergence test of every solve, then out[s * kPcgScalars ...] = scalars of solve s
hipError_t launch_pcg_batch_check(const PcgBatch &b, double *out, hipStream_t s);

// Batches of CG iterations replayed from captured HIP graphs.  The iteration kernels read the
// index k only as k & 1 and k == 0 (the counts they report come from kPcgNext), so a batch of
// `count` iterations starting at k >= 1 launches the same kernels with the same arguments as any
// other batch of the same solve starting at an index of the same parity: one capture serves every
// later batch.  Captures are keyed by everything the launches pass (buffers, tiling, multigrid
// levels, count, parity); the cache keeps kPcgGraphSlots of them and evicts the least recently
// used.  A batch at k = 0, on the null stream, or with graphs disabled is launched directly.
constexpr int kPcgGraphSlots = 16;
class PcgGraphCache {
  public:
    PcgGraphCache() = default;
    PcgGraphCache(const PcgGraphCache &) = delete;
    PcgGraphCache &operator=(const PcgGraphCache &) = delete;
    ~PcgGraphCache() { release(); }
    void release();
    // st's (or every batched solve's) next `count` iterations; st.it (b.it) advances by count
    hipError_t iterations(PcgState &st, int count, hipStream_t s, const MgHier *mg);
    hipError_t batch_iterations(PcgBatch &b, int count, hipStream_t s, const MgBatch *mb);
    bool enabled = true;
    long long captures = 0, replays = 0;

  private:
    struct Entry {
        hipGraphExec_t exec = nullptr;
        bool batch = false, has_mg = false;
        int count = 0, parity = 0;
        PcgState st{};
        MgHier mg{};
        PcgBatch b{};
        MgBatch mb{};
        unsigned long long used = 0;
    };
    Entry *find(const Entry &key);
    hipError_t capture_and_launch(Entry &key, hipStream_t s);
    Entry slots_[kPcgGraphSlots];
    unsigned long long tick_ = 0;
};
hipError_t launch_project(const PcgGrid &g, const double *p, double h, float *u, float *v, float *w,
                          float *pressure_out, hipStream_t s);
hipError_t launch_zero_solid_faces(const PcgGrid &g, float *u, float *v, float *w, hipStream_t s);

// context accessors and the calling thread's error message (fia_api.hip), for other host files
}  // namespace fia
struct fia_ctx;
namespace fia {
int ctx_device(const fia_ctx *c);
hipStream_t ctx_stream(const fia_ctx *c);
int ctx_precond(const fia_ctx *c);
// marks a context as one of several sharing the device for a step of several scenes (its solves
// then tile and iterate the same way batched or not)
void ctx_set_shared(fia_ctx *c, bool shared);
fia_status set_error(fia_status status, const std::string &msg);   // returns status
// batched pressure solves of many scenes (fia_api.hip): buffers on the main context, set-up of
// one scene on a worker context, then the joint solve and projection on the main context
fia_status scenes_solve_alloc(fia_ctx *main, const fia_scene *scenes, int n);
fia_status scene_solve_setup(fia_ctx *worker, fia_ctx *main, int s, fia_scene *scene, fia_unet *net, double rel_tol,
                             int max_iters);
fia_status scenes_solve_run(fia_ctx *main, fia_scene *scenes, const int *idx, int m, int max_iters);

// diagnostics (fia_io.hip): FP64 sums of up to four fields (out[0..3]) and of the pressure over
// the six boundary layers of cells (out[0..5]); part holds io_partial_words() doubles
size_t io_partial_words();
// the allocation-free totals' fixed device scratch (io_partial_words() doubles): its address on
// `device` looked up once with that device current (context creation), and that address
hipError_t io_scratch_init(int device);
double *io_scratch(int device);
hipError_t launch_field_sums(const float *const fields[4], const size_t sizes[4], double *part, double *out,
                             hipStream_t s);
hipError_t launch_wall_sums(const float *p, int nx, int ny, int nz, double *part, double *out, hipStream_t s);

}  // namespace fia
