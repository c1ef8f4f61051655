// fia_api.hip - the C ABI of libfia.so (include/fia.h): context, device scratch, argument
// checks and the order of kernels for each entry point.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/fia.h"
#include "fia_internal.hpp"
#include "fia_unet.hpp"

using namespace fia;

namespace {

thread_local std::string g_last_error = "no error";

fia_status fail(fia_status s, const std::string &msg) {
    g_last_error = msg;
    return s;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return fail(FIA_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e));    \
    } while (0)


const int kGaussSizes[] = {1, 3, 5, 7, 9, 11, 13, 25, 49};

bool gauss_supported(int n) {
    for (int s : kGaussSizes) if (s == n) return true;
    return false;
}

// Gauss-Legendre rule with weights summing to 1 (reference src/gauss_quadrature_points.cpp:18-299)
GaussTable make_gauss(int n) {
    GaussTable g;
    memset(&g, 0, sizeof(g));
    g.n = n;
    if (n == 1) { g.x[0] = 0.0; g.w[0] = 1.0; return g; }
    for (int i = 0; i < n; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = 0.0;
            for (int k = 1; k <= n; ++k) { double p2 = p1; p1 = p0; p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k; }
            dp = n * (z * p0 - p1) / (z * z - 1.0);
            const double dz = p0 / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = 0.0;
        for (int k = 1; k <= n; ++k) { double p2 = p1; p1 = p0; p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k; }
        dp = n * (z * p0 - p1) / (z * z - 1.0);
        g.x[n - 1 - i] = z;
        g.w[n - 1 - i] = 1.0 / ((1.0 - z * z) * dp * dp);
    }
    return g;
}

}  // namespace

// Per-family face fluxes of one pass and the corrections of faces whose two evaluations differ.
// Between passes every correction is zero and every mask bit clear: a pass's general cutter sets
// the bits of its faces (dirty), and the next pass on the set clears them first.
struct FluxSet {
    DevBuf<double> f[3], c[3];
    DevBuf<unsigned> m[3];
    bool dirty = false;
    size_t bytes() const {
        size_t b = 0;
        for (int a = 0; a < 3; ++a) b += f[a].bytes() + c[a].bytes() + m[a].bytes();
        return b;
    }
    void release() { for (int a = 0; a < 3; ++a) { f[a].release(); c[a].release(); m[a].release(); } dirty = false; }
    FluxView view() const {
        FluxView v;
        for (int a = 0; a < 3; ++a) { v.f[a] = f[a].p; v.c[a] = c[a].p; v.m[a] = m[a].p; }
        return v;
    }
};

struct fia_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;   // false: the caller's stream (fia_ctx_create_on_stream)
    DevBuf<double> psi;
    DevBuf<float> disp[3];
    FluxSet fsa, fsb;      // face fluxes (and orientation corrections) of two passes
    DevBuf<unsigned> face_n;               // piece slots per face
    DevBuf<unsigned long long> face_start; // exclusive scan of face_n
    DevBuf<unsigned long long> scan_tmp;   // tile sums of the slot scan
    DevBuf<unsigned long long> bounds;     // slot bounds of the families (device side)
    unsigned long long *host_bounds = nullptr;   // pinned: [0] 0, [1] [2] family starts, [3] total
    int *host_flags = nullptr;                   // pinned: twisted count, kernel error
    double *host_scal = nullptr;                 // pinned: PCG scalars of two solves read back while iterating
    DevBuf<double> piece_c[3], piece_ay, piece_val;
    DevBuf<unsigned> gen_n;                 // twisted faces: pieces per face and orientation
    DevBuf<unsigned long long> gen_start;   // their exclusive scan and total
    DevBuf<double> gen_c[3], gen_ay, gen_val;   // the twisted faces' piece records and values
    DevBuf<float> work[3];
    DevBuf<unsigned> twisted;
    DevBuf<unsigned> degen;     // degenerate faces of the segment-cutter mode
    DevBuf<int> counters;   // [0] twisted count, [1] error, [2] degenerate count
    DevBuf<float> grid[4];  // device mirrors of host grids: rho, u, v, w
    DevBuf<float> gen[3];   // scratch for test entry points
    DevBuf<double> genD[2];
    DevBuf<int> genI;
    DevBuf<double> pcg_b;         // right-hand side of the pressure system
    DevBuf<uint8_t> pcg_flags;    // per-cell neighbour flags of the pressure system
    struct PcgBufs {              // one CG solve: x, r, p (two buffers), q (streaming update), z and r.z
                                  // partials (multigrid), partials, scalars
        DevBuf<double> x, r, r_alt, q, z, part_mg, part, scal;
        DevBuf<float> p0, p1;   // the search direction
        size_t bytes() const {
            return x.bytes() + r.bytes() + r_alt.bytes() + p0.bytes() + p1.bytes() + q.bytes() + z.bytes() +
                   part_mg.bytes() + part.bytes() + scal.bytes();
        }
        void release() {
            x.release(); r.release(); r_alt.release(); p0.release(); p1.release(); q.release(); z.release();
            part_mg.release(); part.release(); scal.release();
        }
    } pcgs[2];
    // multigrid preconditioner: the levels' operators and vectors, and the hierarchy's shape
    int precond = FIA_PRECOND_MG;
    struct MgBufs {
        DevBuf<float4> op[kMgMaxLevels];
        DevBuf<float> b[kMgMaxLevels], x[kMgMaxLevels], xs[kMgMaxLevels];
        size_t bytes() const {
            size_t t = 0;
            for (int l = 0; l < kMgMaxLevels; ++l) t += op[l].bytes() + b[l].bytes() + x[l].bytes() + xs[l].bytes();
            return t;
        }
        void release() {
            for (int l = 0; l < kMgMaxLevels; ++l) { op[l].release(); b[l].release(); x[l].release(); xs[l].release(); }
        }
    } mgb;
    MgHier mgh;
    DevBuf<double> scal;
    DevBuf<unsigned int> bits;
    DevBuf<uint8_t> solid;
    // pressure systems of scenes solved together (fia_scenes_step_batched)
    struct SceneSolve {           // one scene's system and its CG solves (two while probing)
        DevBuf<double> b;
        DevBuf<uint8_t> flags;
        PcgBufs slot[2];
        PcgGrid pg;
        PcgState st;              // the solve the joint iterations continue
        int used_network = 0;
        MgBufs mgb;               // multigrid levels (main context's preconditioner FIA_PRECOND_MG)
        MgHier mgh;
        DevBuf<MgLevel> lv_dev;   // mgh.lv in device memory, for the batched V-cycle
        size_t bytes() const {
            return b.bytes() + flags.bytes() + slot[0].bytes() + slot[1].bytes() + mgb.bytes() + lv_dev.bytes();
        }
        void release() {
            b.release(); flags.release(); slot[0].release(); slot[1].release(); mgb.release(); lv_dev.release();
        }
    };
    std::vector<SceneSolve> solves;
    DevBuf<PcgState> batch_states;
    DevBuf<int2> batch_map;
    DevBuf<double> batch_scal;
    double *host_batch = nullptr;   // pinned: the scalars of every batched solve
    size_t host_batch_n = 0;
    PcgGraphCache graphs;           // captured CG iteration batches (fia_ctx_set_graphs)
    bool shared_device = false;     // inside a step of several scenes: other solves share the device
    // launch shapes per cell-grid shape (fia_ctx_set_tuning, fia_ctx_autotune) and for the rest
    struct ShapeTuning {
        int nx, ny, nz;
        fia_tuning t;
    };
    std::vector<ShapeTuning> tunings;
    fia_tuning tuning_default{0, 0, 0, 0, -1, -1};
    long long graph_captures0 = 0, graph_replays0 = 0;   // the cache's counters at the last stats reset
    // instrumentation
    bool timing = false;
    DevBuf<unsigned long long> stats;   // [0] pieces, [1] density points
    struct Span { hipEvent_t a, b; int kind; };
    std::vector<Span> spans;            // recorded, not yet read
    fia_kernel_stats acc{};
};

struct fia_unet {
    int device = 0;
    fia::UNet net;
    DevBuf<float> y, dy;
    DevBuf<uint8_t> flags, solid;
    DevBuf<double> rhs;
};

namespace fia {
int ctx_device(const fia_ctx *c) { return c->device; }
hipStream_t ctx_stream(const fia_ctx *c) { return c->stream; }
int ctx_precond(const fia_ctx *c) { return c->precond; }
void ctx_set_shared(fia_ctx *c, bool shared) { c->shared_device = shared; }
fia_status set_error(fia_status status, const std::string &msg) { return fail(status, msg); }
}  // namespace fia

namespace {

size_t ctx_bytes(const fia_ctx *c) {
    size_t b = c->psi.bytes() + c->twisted.bytes() + c->degen.bytes() + c->counters.bytes() + c->scal.bytes() + c->bits.bytes() +
               c->solid.bytes() + c->genI.bytes();
    b += c->fsa.bytes() + c->fsb.bytes() + c->face_n.bytes() + c->face_start.bytes() + c->scan_tmp.bytes() + c->bounds.bytes() +
         c->piece_ay.bytes() + c->piece_val.bytes();
    for (int a = 0; a < 3; ++a) b += c->piece_c[a].bytes() + c->gen_c[a].bytes();
    b += c->gen_n.bytes() + c->gen_start.bytes() + c->gen_ay.bytes() + c->gen_val.bytes();
    for (int a = 0; a < 3; ++a) b += c->disp[a].bytes() + c->work[a].bytes() + c->gen[a].bytes();
    for (int a = 0; a < 4; ++a) b += c->grid[a].bytes();
    for (int a = 0; a < 2; ++a) b += c->genD[a].bytes();
    b += c->pcg_b.bytes() + c->pcg_flags.bytes() + c->pcgs[0].bytes() + c->pcgs[1].bytes() + c->mgb.bytes();
    for (const auto &sv : c->solves) b += sv.bytes();
    b += c->batch_states.bytes() + c->batch_map.bytes() + c->batch_scal.bytes();
    return b;
}

hipError_t set_device(fia_ctx *c) { return hipSetDevice(c->device); }

// The launch shapes for grids of nx x ny x nz cells (and their staggered components).
const fia_tuning &tuning_for(const fia_ctx *c, int nx, int ny, int nz) {
    for (const auto &s : c->tunings)
        if (s.nx == nx && s.ny == ny && s.nz == nz) return s.t;
    return c->tuning_default;
}

enum SpanKind { SK_FACE = 0, SK_VERTEX = 1, SK_PSI = 2, SK_GATHER = 3, SK_QUAD = 4, SK_PCG = 5 };

// Opens a timed span on the context's stream (no-op unless timing is enabled); idx receives the
// span's index for span_end (spans may nest).
hipError_t span_begin(fia_ctx *c, int kind, int &idx) {
    idx = -1;
    if (!c->timing) return hipSuccess;
    fia_ctx::Span sp;
    sp.kind = kind;
    hipError_t e = hipEventCreate(&sp.a);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&sp.b);
    if (e != hipSuccess) { (void)hipEventDestroy(sp.a); return e; }
    e = hipEventRecord(sp.a, c->stream);
    idx = (int)c->spans.size();
    c->spans.push_back(sp);
    return e;
}
hipError_t span_end(fia_ctx *c, int idx) {
    if (!c->timing || idx < 0 || idx >= (int)c->spans.size()) return hipSuccess;
    return hipEventRecord(c->spans[idx].b, c->stream);
}

struct PassSpec {
    int split, interp, psi_mode, zero_velocity, integral, ng;
    double min_area;
};

// Reads the kernel error flag (synchronises the stream) and maps it to a status.
fia_status check_kernel_errors(fia_ctx *c) {
    int *h = c->host_flags;
    HIP_TRY(hipMemcpyAsync(h, c->counters.p, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (h[1] == KE_CAPACITY) return fail(FIA_ERR_CAPACITY, "more pieces than max_pieces");
    if (h[1] == KE_CAP_VERTS) return fail(FIA_ERR_CAPACITY, "face cutter: a piece has more than 30 vertices");
    if (h[1] == KE_CAP_STACK) return fail(FIA_ERR_CAPACITY, "face cutter: too many pending pieces");
    if (h[1] == KE_CAP_TRIS) return fail(FIA_ERR_CAPACITY, "face cutter: too many pending triangles (minimal_area)");
    if (h[1] == KE_CUT)
        return fail(FIA_ERR_CUT, "a cut line crossed a piece's boundary an odd number of times "
                                 "(the reference's cutter stops the program here: src/facecutter.h:190-203)");
    if (h[1] == KE_NONFINITE) return fail(FIA_ERR_NONFINITE, "non-finite backtraced vertex (NaN/Inf velocity or dt)");
    if (h[1] != 0) return fail(FIA_ERR_DEVICE, "kernel reported error " + std::to_string(h[1]));
    return FIA_OK;
}

__global__ void zero_ints_kernel(int *a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}
__global__ void zero_uints_kernel(unsigned int *a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0u;
}
__global__ void copy_floats_kernel(const float *a, float *b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

fia_status ensure_common(fia_ctx *c) {
    HIP_TRY(c->counters.ensure(4));
    return FIA_OK;
}

fia_status reset_error(fia_ctx *c) {
    hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, c->stream, c->counters.p, 4);
    HIP_TRY(hipGetLastError());
    return FIA_OK;
}

fia_status do_psi(fia_ctx *c, const CompGrid &g, const float *q, int mode, int integral, int ng) {
    HIP_TRY(c->psi.ensure(g.psi_size()));
    PsiParams p;
    memset(&p, 0, sizeof(p));
    p.g = g; p.q = q; p.psi = c->psi.p; p.mode = mode; p.integral = integral; p.kappa = 1.0f;
    p.gauss = make_gauss(ng);
    int sp;
    HIP_TRY(span_begin(c, SK_PSI, sp));
    HIP_TRY(launch_psi(p, c->stream));
    HIP_TRY(span_end(c, sp));
    return FIA_OK;
}

fia_status do_vertices(fia_ctx *c, const CompGrid &g, const float *u, const float *v, const float *w, double dt_over_h,
                       int interp) {
    for (int a = 0; a < 3; ++a) HIP_TRY(c->disp[a].ensure(g.lattice()));
    VertexParams p;
    memset(&p, 0, sizeof(p));
    // the vertex velocities of "WENO6-optimized" are WENO6's (reference src/grid_3d.cpp:557)
    p.g = g; p.u = u; p.v = v; p.w = w; p.dt_over_h = (float)dt_over_h;
    p.interp = interp == FIA_INTERP_WENO6_OPTIMIZED ? FIA_INTERP_WENO6 : interp;
    for (int a = 0; a < 3; ++a) p.disp[a] = c->disp[a].p;
    int sp;
    HIP_TRY(span_begin(c, SK_VERTEX, sp));
    HIP_TRY(launch_vertices(p, c->stream));
    HIP_TRY(span_end(c, sp));
    if (c->timing) c->acc.vertices += (long long)g.lattice();
    return FIA_OK;
}

fia_status do_faces(fia_ctx *c, const CompGrid &g, const float *q, bool moving, FluxSet &fs, const PassSpec &s,
                    int fam_mask, double h) {
    for (int f = 0; f < 3; ++f) {
        if (g.faces(f) >= (size_t(1) << 30)) return fail(FIA_ERR_INVALID_ARG, "grid too large for the face index");
        HIP_TRY(fs.f[f].ensure(g.faces(f)));
        if (!fs.c[f].p || fs.c[f].n < g.faces(f)) {   // new corrections and masks start at zero
            HIP_TRY(fs.c[f].ensure(g.faces(f)));
            HIP_TRY(fs.m[f].ensure((fs.c[f].n + 31) / 32));
            HIP_TRY(hipMemsetAsync(fs.c[f].p, 0, fs.c[f].bytes(), c->stream));
            HIP_TRY(hipMemsetAsync(fs.m[f].p, 0, fs.m[f].bytes(), c->stream));
        }
    }
    if (fs.dirty) {
        const size_t words[3] = {fs.m[0].n, fs.m[1].n, fs.m[2].n};
        HIP_TRY(launch_corr_reset(fs.view(), words, c->stream));
        fs.dirty = false;
    }
    HIP_TRY(c->twisted.ensure(g.faces(0) + g.faces(1) + g.faces(2)));
    HIP_TRY(c->degen.ensure(g.faces(0) + g.faces(1) + g.faces(2)));
    hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, c->stream, c->counters.p, 1);       // twisted count
    hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, c->stream, c->counters.p + 2, 1);   // degenerate count
    HIP_TRY(hipGetLastError());
    FaceParams p;
    memset(&p, 0, sizeof(p));
    p.g = g; p.q = q; p.psi = c->psi.p;
    for (int a = 0; a < 3; ++a) {
        p.disp[a] = moving ? c->disp[a].p : nullptr;
        p.flux[a] = fs.f[a].p;
        p.corr[a] = fs.c[a].p;
        p.corr_mask[a] = fs.m[a].p;
    }
    p.twisted = c->twisted.p; p.twisted_count = c->counters.p; p.error = c->counters.p + 1;
    p.degen = c->degen.p; p.degen_count = c->counters.p + 2;
    p.split = s.split; p.integral = s.integral; p.fam_mask = fam_mask;
    // "WENO6-optimized": the WENO6 step whose quadrature points are interpolated by the ISPC variant
    // (reference src/calc_density_in_cell_from_quadrature_point_list*.cpp)
    p.interp = s.interp == FIA_INTERP_WENO6_OPTIMIZED ? FIA_INTERP_WENO6 : s.interp;
    p.weno_opt = s.interp == FIA_INTERP_WENO6_OPTIMIZED ? 1 : 0;
    p.general_all = s.min_area < 1.0;
    p.fuse_value = !p.integral && (p.interp == FIA_INTERP_LINEAR || p.interp == FIA_INTERP_1DY_LINEAR);
    p.kappa_q = 1.0f;
    p.kappa_psi = (float)(h * h);   // psi is stored in units of q*cells; S in physical units
    p.min_area = s.min_area;
    p.gauss = make_gauss(s.ng);
    const fia_tuning &tu = tuning_for(c, g.n[0], g.n[1], g.n[2]);
    p.face_block = tu.face_block;
    p.piece_block = tu.piece_block;
    if (c->timing) {
        HIP_TRY(c->stats.ensure(kStatWords * kStatSlots));
        p.stats = c->stats.p;
    }
    if (!moving && !p.general_all) {
        // undeformed faces: closed-form pieces of the y faces (the only family with y area)
        int sp;
        HIP_TRY(span_begin(c, SK_FACE, sp));
        HIP_TRY(launch_zero_velocity_faces(p, c->stream));
        HIP_TRY(span_end(c, sp));
        if (c->timing) c->acc.face_launches += 1;
        return FIA_OK;
    }
    // slots per face -> exclusive scan -> piece records -> quadrature -> ordered sums
    const size_t nf = g.faces(0) + g.faces(1) + g.faces(2);
    HIP_TRY(c->face_n.ensure(nf));
    HIP_TRY(c->face_start.ensure(nf));
    HIP_TRY(c->scan_tmp.ensure(scan_temp_words(nf)));
    HIP_TRY(c->bounds.ensure(4));
    p.face_n = c->face_n.p;
    p.face_start = c->face_start.p;
    p.fam_base[0] = 0; p.fam_base[1] = g.faces(0); p.fam_base[2] = g.faces(0) + g.faces(1);
    int sp;
    HIP_TRY(span_begin(c, SK_FACE, sp));
    HIP_TRY(launch_face_count(p, c->stream));   // all families; inactive ones count zero
    HIP_TRY(launch_scan(c->scan_tmp.p, c->face_n.p, c->face_start.p, nf, p, c->bounds.p, c->stream));
    // slot bounds of the families into pinned host memory
    HIP_TRY(hipMemcpyAsync(c->host_bounds, c->bounds.p, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           c->stream));
    // with the twisted and degenerate face counts (complete after the count kernels), which size
    // the grids of their kernels
    HIP_TRY(hipMemcpyAsync(c->host_flags, c->counters.p, 3 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int n_twisted = c->host_flags[0], n_degen = c->host_flags[2];
    unsigned long long fam_slots[4] = {0, c->host_bounds[1], c->host_bounds[2], c->host_bounds[3]};
    const unsigned long long nslots = fam_slots[3];
    if (!(fam_slots[1] <= fam_slots[2] && fam_slots[2] <= fam_slots[3]))
        return fail(FIA_ERR_DEVICE, "inconsistent piece slot bounds");
    for (int a = 0; a < 3; ++a) HIP_TRY(c->piece_c[a].ensure(nslots));
    HIP_TRY(c->piece_ay.ensure(nslots));
    HIP_TRY(c->piece_val.ensure(nslots));
    for (int a = 0; a < 3; ++a) p.piece_c[a] = c->piece_c[a].p;
    p.piece_ay = c->piece_ay.p;
    p.piece_val = c->piece_val.p;
    // pieces: one-pass face kernels (segment cutter), then the degenerate faces (clipping pipeline)
    HIP_TRY(launch_seg_faces(p, c->stream));
    HIP_TRY(launch_faces_degenerate(p, n_degen, c->stream));
    // the WENO6 integral quadrature launches are timed on their own (the dominant kernel)
    const bool weno_quad = piece_quadrature_mode(p) != 0;
    int sq = -1;
    if (weno_quad) HIP_TRY(span_begin(c, SK_QUAD, sq));
    if (!p.fuse_value) HIP_TRY(launch_pieces(p, nslots, c->stream));   // fused: the cutters stored the values
    if (weno_quad) HIP_TRY(span_end(c, sq));
    HIP_TRY(launch_face_sum(p, c->stream));
    if (n_twisted > 0) {
        // twisted faces: pieces counted, then stored in records sized from the read-back total
        HIP_TRY(c->gen_n.ensure(2 * (size_t)n_twisted));
        HIP_TRY(c->gen_start.ensure(2 * (size_t)n_twisted + 1));
        p.gen_n = c->gen_n.p;
        p.gen_start = c->gen_start.p;
        HIP_TRY(launch_general_count(p, n_twisted, c->stream));
        HIP_TRY(hipMemcpyAsync(c->host_bounds, c->gen_start.p + 2 * (size_t)n_twisted, sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(c->host_flags + 1, c->counters.p + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        const unsigned long long npieces = c->host_bounds[0];
        if (c->host_flags[1] != 0) return check_kernel_errors(c);   // the cutter failed on a face
        if (npieces > kGenMaxPieces)
            return fail(FIA_ERR_CAPACITY, "twisted faces: " + std::to_string(npieces) + " pieces in one pass, more than "
                                          "the general cutter's budget of " + std::to_string(kGenMaxPieces));
        for (int a = 0; a < 3; ++a) HIP_TRY(c->gen_c[a].ensure(npieces));
        HIP_TRY(c->gen_ay.ensure(npieces));
        HIP_TRY(c->gen_val.ensure(npieces));
        for (int a = 0; a < 3; ++a) p.gen_c[a] = c->gen_c[a].p;
        p.gen_ay = c->gen_ay.p;
        p.gen_val = c->gen_val.p;
        fs.dirty = true;   // from here on the set may hold corrections
        HIP_TRY(launch_general_pieces(p, n_twisted, npieces, c->stream));
        if (c->timing) { c->acc.twisted_faces += n_twisted; c->acc.general_pieces += (long long)npieces; }
    }
    HIP_TRY(span_end(c, sp));
    if (c->timing) { c->acc.face_launches += 1; c->acc.quad_launches += weno_quad ? 1 : 0; }
    return FIA_OK;
}

hipError_t timed_gather(fia_ctx *c, const CompGrid &g, FluxSet &s1, FluxSet *s2, const float *base, float *out,
                        int mode, int fm) {
    int sp;
    hipError_t e = span_begin(c, SK_GATHER, sp);
    if (e != hipSuccess) return e;
    const FluxView v1 = s1.view();
    FluxView v2;
    if (s2) v2 = s2->view();
    e = launch_gather(g, v1, s2 ? &v2 : nullptr, base, out, mode, fm, c->stream,
                      tuning_for(c, g.n[0], g.n[1], g.n[2]).cell_block);
    if (e != hipSuccess) return e;
    if (c->timing) c->acc.gather_cells += (long long)g.cells();
    return span_end(c, sp);
}

// Complete pass: q_out = flux-advected q_in (q_out may alias q_in).
fia_status run_pass(fia_ctx *c, const CompGrid &g, const float *q_in, float *q_out, const float *u, const float *v,
                    const float *w, double h, double dt, const PassSpec &s) {
    fia_status st;
    if ((st = do_psi(c, g, q_in, s.psi_mode, s.integral, s.ng)) != FIA_OK) return st;
    if (!s.zero_velocity && (st = do_vertices(c, g, u, v, w, dt / h, s.interp)) != FIA_OK) return st;
    const int fm = s.zero_velocity ? 2 : 7;   // undeformed x and z faces have exactly zero y area
    if ((st = do_faces(c, g, q_in, !s.zero_velocity, c->fsa, s, fm, h)) != FIA_OK) return st;
    HIP_TRY(timed_gather(c, g, c->fsa, nullptr, nullptr, q_out, 0, fm));
    return FIA_OK;
}

// psi mode implied by the interpolation method (reference src/move_substances_3d.cpp:61-83,
// src/advect_velocity_flux_advection_3d.cpp:601-620)
int psi_mode_for(int interp, int &integral) {
    if (interp == FIA_INTERP_LINEAR) return FIA_PSI_LINEAR;
    if (interp == FIA_INTERP_1DY_LINEAR) { integral = 0; return FIA_PSI_CONST; }
    return FIA_PSI_WENO6;
}

// Full step on a component grid: main pass plus the zero-velocity correction (reference
// src/move_substances_3d.cpp:84-226 for density, src/advect_velocity_flux_advection_3d.cpp:
// 580-759 per velocity component).  q (device, G-sized) is read; result goes to out.
fia_status run_step(fia_ctx *c, const CompGrid &g, const float *q, float *out, const float *u, const float *v,
                    const float *w, double h, double dt, const fia_advect_opts *o, int split_main, int split_err) {
    fia_status st;
    int integral = o->use_integral ? 1 : 0;
    const int pm = psi_mode_for(o->interpolation_method, integral);
    PassSpec p1{split_main, (int)o->interpolation_method, pm, 0, integral, o->num_gauss_points_integrate_density,
                o->minimal_area};
    if ((st = do_psi(c, g, q, p1.psi_mode, p1.integral, p1.ng)) != FIA_OK) return st;
    if ((st = do_vertices(c, g, u, v, w, dt / h, p1.interp)) != FIA_OK) return st;
    if ((st = do_faces(c, g, q, true, c->fsa, p1, 7, h)) != FIA_OK) return st;
    if (!o->enable_zero_velocity_correction) {
        HIP_TRY(timed_gather(c, g, c->fsa, nullptr, nullptr, out, 0, 7));
        return FIA_OK;
    }
    // pass 2: zero velocity, non-integral psi (reference src/move_substances_3d.cpp:133-153)
    HIP_TRY(c->work[2].ensure(g.cells()));
    PassSpec p2{split_main, (int)o->interpolation_method, FIA_PSI_CONST, 1, 0, o->num_gauss_points_integrate_density,
                o->minimal_area};
    if ((st = do_psi(c, g, q, p2.psi_mode, p2.integral, p2.ng)) != FIA_OK) return st;
    if ((st = do_faces(c, g, q, false, c->fsb, p2, 2, h)) != FIA_OK) return st;
    float *err = c->work[2].p;
    HIP_TRY(timed_gather(c, g, c->fsb, nullptr, q, err, 2, 2));   // err = q - pass2(q)
    // pass 3: error advected with 1Dy_linear (reference src/move_substances_3d.cpp:177-225)
    PassSpec p3{split_err, FIA_INTERP_1DY_LINEAR, FIA_PSI_CONST, 0, 0, o->num_gauss_points_integrate_density,
                o->minimal_area};
    if ((st = do_psi(c, g, err, p3.psi_mode, p3.integral, p3.ng)) != FIA_OK) return st;
    if ((st = do_vertices(c, g, u, v, w, dt / h, p3.interp)) != FIA_OK) return st;
    if ((st = do_faces(c, g, err, true, c->fsb, p3, 7, h)) != FIA_OK) return st;
    HIP_TRY(timed_gather(c, g, c->fsa, &c->fsb, nullptr, out, 1, 7));
    return FIA_OK;
}

fia_status check_grid(const fia_grid3d *g, bool need_density) {
    if (!g) return fail(FIA_ERR_INVALID_ARG, "grid is NULL");
    if (g->nx < 1 || g->ny < 1 || g->nz < 1) return fail(FIA_ERR_INVALID_ARG, "grid sizes must be >= 1");
    if (!(g->dx > 0.0) || !isfinite(g->dx)) return fail(FIA_ERR_INVALID_ARG, "dx must be positive and finite");
    if (!g->u || !g->v || !g->w) return fail(FIA_ERR_INVALID_ARG, "velocity pointer is NULL");
    if (need_density && !g->density) return fail(FIA_ERR_INVALID_ARG, "density pointer is NULL");
    return FIA_OK;
}

fia_status check_opts(const fia_advect_opts *o) {
    if (!o) return fail(FIA_ERR_INVALID_ARG, "opts is NULL");
    if (!o->use_flux_advection) return fail(FIA_ERR_UNSUPPORTED, "semi-Lagrangian advection is not implemented");
    if (o->use_maccormack) return fail(FIA_ERR_UNSUPPORTED, "MacCormack is not implemented");
    if (o->enable_cell_volume_correction) return fail(FIA_ERR_UNSUPPORTED, "cell volume correction is not implemented");
    if (o->interpolation_method < FIA_INTERP_LINEAR || o->interpolation_method > FIA_INTERP_WENO6_OPTIMIZED)
        return fail(FIA_ERR_INVALID_ARG, "unknown interpolation method");
    if (o->interpolation_method == FIA_INTERP_1DY_WENO6)
        return fail(FIA_ERR_UNSUPPORTED, "1Dy_WENO6 has no psi rule in the step (reference leaves it undefined)");
    if (o->split_method < FIA_SPLIT_NONE || o->split_method > FIA_SPLIT_X_AXIS_Z_CELLFACE)
        return fail(FIA_ERR_INVALID_ARG, "unknown split method");
    if (!gauss_supported(o->num_gauss_points_integrate_density))
        return fail(FIA_ERR_INVALID_ARG, "num_gauss_points_integrate_density must be 1,3,5,7,9,11,13,25 or 49");
    if (!(o->minimal_area > 0.0)) return fail(FIA_ERR_INVALID_ARG, "minimal_area must be > 0");
    return FIA_OK;
}

// Device views of a grid: the caller's pointers or mirrors filled from host memory.
struct GridView {
    float *rho, *u, *v, *w;
};

fia_status upload_grid(fia_ctx *c, const fia_grid3d *g, GridView &gv, bool density) {
    const size_t nc = (size_t)g->nx * g->ny * g->nz;
    const size_t nu = (size_t)(g->nx + 1) * g->ny * g->nz, nv = (size_t)g->nx * (g->ny + 1) * g->nz,
                 nw = (size_t)g->nx * g->ny * (g->nz + 1);
    if (g->on_device) {
        gv.rho = g->density; gv.u = g->u; gv.v = g->v; gv.w = g->w;
        return FIA_OK;
    }
    HIP_TRY(c->grid[1].ensure(nu)); HIP_TRY(c->grid[2].ensure(nv)); HIP_TRY(c->grid[3].ensure(nw));
    HIP_TRY(hipMemcpyAsync(c->grid[1].p, g->u, nu * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->grid[2].p, g->v, nv * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->grid[3].p, g->w, nw * 4, hipMemcpyHostToDevice, c->stream));
    gv.u = c->grid[1].p; gv.v = c->grid[2].p; gv.w = c->grid[3].p; gv.rho = nullptr;
    if (density) {
        HIP_TRY(c->grid[0].ensure(nc));
        HIP_TRY(hipMemcpyAsync(c->grid[0].p, g->density, nc * 4, hipMemcpyHostToDevice, c->stream));
        gv.rho = c->grid[0].p;
    }
    return FIA_OK;
}

// Split of each velocity component's main pass and error pass (reference
// src/advect_velocity_flux_advection_3d.cpp:622-628, 804-810, 997-1003; 721, 913, 1096).
int velocity_split(int comp, bool one_dy, bool error_pass) {
    const bool flip = one_dy != error_pass;
    if (comp == 0) return flip ? FIA_SPLIT_X_CELLFACE_Z_AXIS : FIA_SPLIT_X_AXIS_Z_CELLFACE;
    if (comp == 1) return flip ? FIA_SPLIT_XZ_AXIS : FIA_SPLIT_XZ_CELLFACE;
    return flip ? FIA_SPLIT_X_AXIS_Z_CELLFACE : FIA_SPLIT_X_CELLFACE_Z_AXIS;
}

}  // namespace

// =============================================================================================

// The pressure system of a grid: right-hand side and neighbour flags in the context, obstacle
// faces zeroed, the pinned cell chosen (the last cell, or the last fluid cell with obstacles).
// The pressure system of the grid whose velocity is at gv, with obstacle mask solid_dev (device,
// or null), into b and flags; pg describes it.  Velocity on faces touching an obstacle is zeroed.
static fia_status assemble_system(fia_ctx *c, const fia_grid3d *g, const GridView &gv, const uint8_t *solid_dev,
                                  double *b, uint8_t *flags, PcgGrid &pg) {
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    pg.nx = g->nx; pg.ny = g->ny; pg.nz = g->nz; pg.solid = solid_dev; pg.pin = (long)n - 1;
    if (solid_dev) {
        // last fluid cell is pinned (the reference pins the last cell, which has no obstacles)
        std::vector<uint8_t> hs(n);
        HIP_TRY(hipMemcpyAsync(hs.data(), solid_dev, n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        long pin = -1;
        for (long i = (long)n - 1; i >= 0; --i) if (!hs[i]) { pin = i; break; }
        if (pin < 0) return fail(FIA_ERR_INVALID_ARG, "every cell is solid");
        pg.pin = pin;
        HIP_TRY(launch_zero_solid_faces(pg, gv.u, gv.v, gv.w, c->stream));
    }
    HIP_TRY(launch_pcg_setup(pg, gv.u, gv.v, gv.w, g->dx, b, flags, c->stream));
    return FIA_OK;
}

// The system into the context's own buffers (the mask, host or device, copied first).
static fia_status pressure_system(fia_ctx *c, const fia_grid3d *g, const GridView &gv, const uint8_t *solid, PcgGrid &pg) {
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    if (solid) {
        HIP_TRY(c->solid.ensure(n));
        HIP_TRY(hipMemcpyAsync(c->solid.p, solid, n, g->on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                               c->stream));
    }
    HIP_TRY(c->pcg_b.ensure(n));
    HIP_TRY(c->pcg_flags.ensure(n));
    return assemble_system(c, g, gv, solid ? c->solid.p : nullptr, c->pcg_b.p, c->pcg_flags.p, pg);
}

// Buffers B (allocated here for the grid of pg) bound into a CG state over the system (b, flags).
// batched: the solve joins fia_scenes_step_batched's launches, which run the two-kernel iteration.
// shared: other scenes' solves share the device (a step of several scenes): the tiles are sized by
// cells per block, not by one round of the device, and the iteration is the two-kernel one, so a
// scene's solve has the same bits whether its launches are batched or its own.
static fia_status bind_solve(fia_ctx::PcgBufs &B, const PcgGrid &pg, const double *b, const uint8_t *flags,
                             PcgState &ps, bool mg, const fia_tuning &tu, bool batched, bool shared) {
    const size_t n = (size_t)pg.nx * pg.ny * pg.nz;
    ps.nblocks = pcg_blocks(n);
    // fused by default from 4M cells (256^3: 191 against 203 us per iteration at the best chunking;
    // 128^3: even)
    ps.fused = !mg && !batched && !shared &&
               (tu.pcg_fused == 1 || (tu.pcg_fused == -1 && n >= ((size_t)1 << 22)));
    pcg_direction_tiling(pg.nx, pg.ny, pg.nz, ps, tu.pcg_chunk_x, tu.pcg_recompute, shared);
    HIP_TRY(B.x.ensure(n)); HIP_TRY(B.r.ensure(n)); HIP_TRY(B.p0.ensure(n)); HIP_TRY(B.p1.ensure(n));
    HIP_TRY(B.part.ensure(pcg_partial_words(ps))); HIP_TRY(B.scal.ensure(kPcgScalars));
    if (!ps.recompute_q && !ps.fused) HIP_TRY(B.q.ensure(n));
    if (ps.fused) HIP_TRY(B.r_alt.ensure(n));
    if (mg) { HIP_TRY(B.z.ensure(n)); HIP_TRY(B.part_mg.ensure((size_t)ps.dir_blocks)); }
    ps.g = pg; ps.flags = flags; ps.b = b;
    ps.x = B.x.p; ps.r = B.r.p; ps.p[0] = B.p0.p; ps.p[1] = B.p1.p;
    ps.r_alt = ps.fused ? B.r_alt.p : nullptr;
    ps.q = (ps.recompute_q || ps.fused) ? nullptr : B.q.p;
    ps.z = mg ? B.z.p : nullptr;
    ps.part_mg = mg ? B.part_mg.p : nullptr;
    ps.mg_lv = nullptr;
    ps.mg_levels = 0;
    ps.scal = B.scal.p; ps.part_start = B.part.p; ps.part_dir = B.part.p + 3 * (size_t)ps.nblocks;
    ps.part_update = ps.part_dir + ps.dir_blocks;
    ps.part_fused = ps.fused ? ps.part_update + 2 * (size_t)ps.upd_blocks : nullptr;
    ps.it = 0;
    return FIA_OK;
}

// A multigrid hierarchy for ps's grid in buffers B, its coarse operators built from ps's flags on
// stream st.
static fia_status mg_setup(hipStream_t stream, fia_ctx::MgBufs &B, MgHier &h, const PcgState &ps) {
    mg_shape(ps.g.nx, ps.g.ny, ps.g.nz, h);
    const MgLevel &top = h.lv[h.levels - 1];
    if ((size_t)top.nx * top.ny * top.nz > 256) return fail(FIA_ERR_UNSUPPORTED, "grid too large for the multigrid levels");
    for (int l = 0; l < h.levels; ++l) {
        MgLevel &L = h.lv[l];
        const size_t n = (size_t)L.nx * L.ny * L.nz;
        HIP_TRY(B.x[l].ensure(n));
        HIP_TRY(B.b[l].ensure(n));
        L.x = B.x[l].p;
        L.b = B.b[l].p;
        if (l > 0) {
            HIP_TRY(B.op[l].ensure(n)); HIP_TRY(B.xs[l].ensure(n));
            L.op = B.op[l].p; L.xs = B.xs[l].p;
        }
    }
    HIP_TRY(launch_mg_build(ps, h, stream));
    return FIA_OK;
}

// The multigrid hierarchy of the context's current system (ps's grid and flags) under
// FIA_PRECOND_MG, built on the context's stream; null under Jacobi.
static fia_status mg_prepare(fia_ctx *c, const PcgState &ps, const MgHier **out) {
    *out = nullptr;
    if (c->precond != FIA_PRECOND_MG) return FIA_OK;
    fia_status st;
    if ((st = mg_setup(c->stream, c->mgb, c->mgh, ps)) != FIA_OK) return st;
    *out = &c->mgh;
    return FIA_OK;
}

// Solve k's buffers bound into a CG state over the context's system.
static fia_status pcg_bind(fia_ctx *c, int k, const PcgGrid &pg, PcgState &ps) {
    return bind_solve(c->pcgs[k], pg, c->pcg_b.p, c->pcg_flags.p, ps, c->precond == FIA_PRECOND_MG,
                      tuning_for(c, pg.nx, pg.ny, pg.nz), false, c->shared_device);
}

// The start of a solve: r = b - L x and the scalars, then z = M r under multigrid.
static fia_status pcg_begin(fia_ctx *c, PcgState &ps, double tol2, const MgHier *mg) {
    HIP_TRY(launch_pcg_start(ps, tol2, c->stream));
    if (mg) HIP_TRY(launch_mg_vcycle(ps, *mg, c->stream));
    return FIA_OK;
}

// The learned start: sa starts from the network's guess scaled by its energy-optimal factor, sb
// from zero; both run `probe` iterations and the one with the smaller residual is chosen.  ha and
// hb (pinned) receive their scalars.
struct LearnedStart {
    bool use_net;
    double guess_res, gamma, ra, rb, bb;
};
static fia_status learned_start(fia_ctx *c, fia_unet *u, const PcgGrid &pg, PcgState &sa, PcgState &sb, double tol2,
                                int probe, double *ha, double *hb, LearnedStart &ls, const MgHier *mg) {
    HIP_TRY(u->net.forward(sa.b, pg.solid, u->y.p, nullptr, c->stream));
    HIP_TRY(u->net.guess(u->y.p, pg.pin, sa.flags, sa.x, c->stream));
    HIP_TRY(launch_pcg_scale_guess(sa, c->stream));
    fia_status st;
    if ((st = pcg_begin(c, sa, tol2, mg)) != FIA_OK) return st;
    HIP_TRY(launch_pcg_check(sa, c->stream));
    HIP_TRY(hipMemcpyAsync(ha, sa.scal, kPcgScalars * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    ls.bb = ha[kPcgBB] > 0 ? ha[kPcgBB] : 1.0;
    ls.guess_res = sqrt(ha[kPcgRR] / ls.bb);
    ls.gamma = ha[kPcgGamma];
    HIP_TRY(launch_pcg_iterations(sa, pcg_launches_for(sa, probe), c->stream, mg));
    HIP_TRY(launch_pcg_check(sa, c->stream));
    HIP_TRY(launch_pcg_guess(sb, nullptr, c->stream));
    if ((st = pcg_begin(c, sb, tol2, mg)) != FIA_OK) return st;
    HIP_TRY(launch_pcg_iterations(sb, pcg_launches_for(sb, probe), c->stream, mg));
    HIP_TRY(launch_pcg_check(sb, c->stream));
    HIP_TRY(hipMemcpyAsync(ha, sa.scal, kPcgScalars * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(hb, sb.scal, kPcgScalars * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    ls.ra = sqrt(ha[kPcgRR] / ls.bb);
    ls.rb = sqrt(hb[kPcgRR] / ls.bb);
    ls.use_net = ls.ra <= ls.rb;
    return FIA_OK;
}

// Iterates a started solve until it converges or has run max_iters iterations in total; hs
// (pinned, kPcgScalars doubles) receives its scalars.  The device skips iterations after
// convergence, so batches of them are enqueued between polls: 32 for Jacobi (thousands of
// iterations), 8 under multigrid (tens of iterations of ~19 launches, most of a 32-batch after
// convergence would be launches that return at once).
// With timing on, each batch of iteration launches is a timed span (the polls between batches are
// host round trips with the stream idle, outside the spans).
static fia_status pcg_run(fia_ctx *c, PcgState &ps, int max_iters, double *hs, const MgHier *mg) {
    const int poll = mg ? 8 : 32;
    const int it0 = pcg_iters_done(ps);
    while (true) {
        HIP_TRY(launch_pcg_check(ps, c->stream));
        HIP_TRY(hipMemcpyAsync(hs, ps.scal, kPcgScalars * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (hs[kPcgDone] != 0.0 || pcg_iters_done(ps) >= max_iters) break;
        int sp;
        HIP_TRY(span_begin(c, SK_PCG, sp));
        HIP_TRY(c->graphs.iterations(ps, pcg_launches_for(ps, std::min(poll, max_iters - pcg_iters_done(ps))),
                                     c->stream, mg));
        HIP_TRY(span_end(c, sp));
    }
    HIP_TRY(launch_pcg_fused_flush(ps, c->stream));
    if (c->timing) {
        c->acc.pcg_iters += std::max(0LL, (long long)hs[kPcgIters] - (long long)it0);
        c->acc.pcg_solves += 1;
    }
    return FIA_OK;
}

// ---- pressure systems of many scenes, solved together ------------------------------------------

namespace {
fia_status unet_grid(fia_unet *u, int nx, int ny, int nz);
}

namespace fia {

fia_status scenes_solve_alloc(fia_ctx *c, const fia_scene *scenes, int n) {
    HIP_TRY(set_device(c));
    if ((int)c->solves.size() < n) c->solves.resize(n);
    HIP_TRY(c->batch_states.ensure((size_t)std::max(n, 1)));
    HIP_TRY(c->batch_scal.ensure((size_t)std::max(n, 1) * kPcgScalars));
    if (c->host_batch_n < (size_t)n * kPcgScalars) {
        if (c->host_batch) HIP_TRY(hipHostFree(c->host_batch));
        c->host_batch = nullptr;
        c->host_batch_n = 0;
        HIP_TRY(hipHostMalloc((void **)&c->host_batch, (size_t)std::max(n, 1) * kPcgScalars * sizeof(double)));
        c->host_batch_n = (size_t)n * kPcgScalars;
    }
    return FIA_OK;
}

// On worker context cw, for scene s (a device grid): its pressure system into main's solve
// buffers, started as fia_pressure_project (zero guess) or, with a network, as
// fia_pressure_project_learned (network and zero starts probed, the better one kept).
fia_status scene_solve_setup(fia_ctx *cw, fia_ctx *main, int s, fia_scene *sc, fia_unet *net, double rel_tol,
                             int max_iters) {
    HIP_TRY(set_device(cw));
    fia_grid3d *g = &sc->grid;
    fia_ctx::SceneSolve &sv = main->solves[s];
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    fia_status st;
    if (net && (st = unet_grid(net, g->nx, g->ny, g->nz)) != FIA_OK) return st;
    GridView gv;
    if ((st = upload_grid(cw, g, gv, false)) != FIA_OK) return st;
    HIP_TRY(sv.b.ensure(n));
    HIP_TRY(sv.flags.ensure(n));
    if ((st = assemble_system(cw, g, gv, sc->solid, sv.b.p, sv.flags.p, sv.pg)) != FIA_OK) return st;
    const double tol2 = rel_tol * rel_tol;
    const bool use_mg = main->precond == FIA_PRECOND_MG;
    PcgState sa;
    const fia_tuning &tu = tuning_for(main, sv.pg.nx, sv.pg.ny, sv.pg.nz);
    if ((st = bind_solve(sv.slot[0], sv.pg, sv.b.p, sv.flags.p, sa, use_mg, tu, true, cw->shared_device)) != FIA_OK)
        return st;
    const MgHier *mg = nullptr;
    if (use_mg) {
        if ((st = mg_setup(cw->stream, sv.mgb, sv.mgh, sa)) != FIA_OK) return st;
        HIP_TRY(sv.lv_dev.ensure(kMgMaxLevels));
        HIP_TRY(hipMemcpyAsync(sv.lv_dev.p, sv.mgh.lv, sizeof(sv.mgh.lv), hipMemcpyHostToDevice, cw->stream));
        sa.mg_lv = sv.lv_dev.p;
        sa.mg_levels = sv.mgh.levels;
        mg = &sv.mgh;
    }
    if (net) {
        PcgState sb;
        if ((st = bind_solve(sv.slot[1], sv.pg, sv.b.p, sv.flags.p, sb, use_mg, tu, true, cw->shared_device)) !=
            FIA_OK)
            return st;
        sb.mg_lv = sa.mg_lv;
        sb.mg_levels = sa.mg_levels;
        LearnedStart ls;
        if ((st = learned_start(cw, net, sv.pg, sa, sb, tol2, std::min(5, max_iters), cw->host_scal,
                                cw->host_scal + kPcgScalars, ls, mg)) != FIA_OK)
            return st;
        sv.st = ls.use_net ? sa : sb;
        sv.used_network = ls.use_net ? 1 : 0;
    } else {
        HIP_TRY(launch_pcg_guess(sa, nullptr, cw->stream));
        if ((st = pcg_begin(cw, sa, tol2, mg)) != FIA_OK) return st;
        sv.st = sa;
        sv.used_network = 0;
    }
    HIP_TRY(hipStreamSynchronize(cw->stream));
    return FIA_OK;
}

// On the main context: the started solves of scenes idx[0..m) iterated together (one direction
// and one update launch per iteration for all of them) until every one has converged or run
// max_iters iterations, then each scene projected.  Every started solve has run the same number
// of iterations (0, or the probe).
fia_status scenes_solve_run(fia_ctx *c, fia_scene *scenes, const int *idx, int m, int max_iters) {
    HIP_TRY(set_device(c));
    if (m == 0) return FIA_OK;
    std::vector<PcgState> states(m);
    std::vector<int2> map;   // direction blocks of every solve, then update blocks
    for (int j = 0; j < m; ++j) {
        states[j] = c->solves[idx[j]].st;
        if (states[j].it != states[0].it) return fail(FIA_ERR_INVALID_ARG, "scenes started at different iterations");
        for (int k = 0; k < states[j].dir_blocks; ++k) map.push_back(make_int2(j, k));
    }
    const int dir_blocks = (int)map.size();
    for (int j = 0; j < m; ++j)
        if (!states[j].recompute_q)
            for (int k = 0; k < states[j].upd_blocks; ++k) map.push_back(make_int2(j, k));
    const int stream_end = (int)map.size();
    for (int j = 0; j < m; ++j)
        if (states[j].recompute_q)
            for (int k = 0; k < states[j].upd_blocks; ++k) map.push_back(make_int2(j, k));
    const int upd_end = (int)map.size();
    // the batched V-cycle's launches (multigrid): level 1's restriction, then per level the down and
    // up launches of the solves that have the level below
    const bool use_mg = m > 0 && states[0].mg_levels > 0;
    int r0_off = upd_end, down_off[kMgMaxLevels] = {0}, up_off[kMgMaxLevels] = {0};
    MgBatch mb{};
    if (use_mg) {
        for (int j = 0; j < m; ++j) {
            const MgHier &h = c->solves[idx[j]].mgh;
            const size_t n1 = (size_t)h.lv[1].nx * h.lv[1].ny * h.lv[1].nz;
            for (size_t k = 0; k < (n1 + 255) / 256; ++k) map.push_back(make_int2(j, (int)k));
        }
        mb.r0_blocks = (int)map.size() - r0_off;
        for (int l = 1; l < kMgMaxLevels; ++l) {
            down_off[l] = (int)map.size();
            for (int j = 0; j < m; ++j) {
                const MgHier &h = c->solves[idx[j]].mgh;
                if (l + 1 < h.levels)
                    for (int k = 0; k < mg_down_tiles(h.lv[l]); ++k) map.push_back(make_int2(j, k));
            }
            mb.down_blocks[l] = (int)map.size() - down_off[l];
        }
        for (int l = 1; l < kMgMaxLevels; ++l) {
            up_off[l] = (int)map.size();
            for (int j = 0; j < m; ++j) {
                const MgHier &h = c->solves[idx[j]].mgh;
                if (l + 1 < h.levels) {
                    const size_t n = (size_t)h.lv[l].nx * h.lv[l].ny * h.lv[l].nz;
                    for (size_t k = 0; k < (n + 255) / 256; ++k) map.push_back(make_int2(j, (int)k));
                }
            }
            mb.up_blocks[l] = (int)map.size() - up_off[l];
        }
    }
    HIP_TRY(c->batch_map.ensure(map.size()));
    HIP_TRY(hipMemcpyAsync(c->batch_states.p, states.data(), m * sizeof(PcgState), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->batch_map.p, map.data(), map.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
    PcgBatch b;
    b.states = c->batch_states.p;
    b.dir_map = c->batch_map.p; b.stream_map = c->batch_map.p + dir_blocks; b.tile_map = c->batch_map.p + stream_end;
    b.n = m; b.dir_blocks = dir_blocks; b.stream_blocks = stream_end - dir_blocks;
    b.tile_blocks = upd_end - stream_end; b.it = states[0].it;
    if (use_mg) {
        mb.r0_map = c->batch_map.p + r0_off;
        for (int l = 1; l < kMgMaxLevels; ++l) {
            mb.down_map[l] = c->batch_map.p + down_off[l];
            mb.up_map[l] = c->batch_map.p + up_off[l];
        }
    }
    double *hs = c->host_batch;
    const int poll = use_mg ? 8 : 32;
    while (true) {
        HIP_TRY(launch_pcg_batch_check(b, c->batch_scal.p, c->stream));
        HIP_TRY(hipMemcpyAsync(hs, c->batch_scal.p, (size_t)m * kPcgScalars * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        bool all = true;
        for (int j = 0; j < m && all; ++j) all = hs[(size_t)j * kPcgScalars + kPcgDone] != 0.0;
        if (all || b.it >= max_iters) break;
        HIP_TRY(c->graphs.batch_iterations(b, std::min(poll, max_iters - b.it), c->stream, use_mg ? &mb : nullptr));
    }
    for (int j = 0; j < m; ++j) {
        const double *h = hs + (size_t)j * kPcgScalars;
        fia_scene *sc = &scenes[idx[j]];
        const fia_ctx::SceneSolve &sv = c->solves[idx[j]];
        const double bb = h[kPcgBB] > 0 ? h[kPcgBB] : 1.0;
        sc->pcg_iters = (int)h[kPcgIters];
        sc->pcg_residual = sqrt(h[kPcgRR] / bb);
        sc->used_network = sv.used_network;
        sc->status = h[kPcgDone] != 0.0 ? FIA_OK : FIA_ERR_NOT_CONVERGED;
        HIP_TRY(launch_project(sv.pg, sv.st.x, sc->grid.dx, sc->grid.u, sc->grid.v, sc->grid.w, sc->pressure,
                               c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FIA_OK;
}

}  // namespace fia

extern "C" {

void fia_default_opts(fia_advect_opts *o) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->use_flux_advection = 1;
    o->use_maccormack = 0;
    o->use_clamping_in_maccormack = 1;
    o->num_gauss_quad_boundary = 1;
    o->num_gauss_quad_bulk = 1;
    o->enable_zero_velocity_correction = 1;
    o->split_method = FIA_SPLIT_NONE;
    o->num_gauss_quadrature_points = 1;
    o->interpolation_method = FIA_INTERP_WENO6;
    o->use_integral = 1;
    o->num_gauss_points_integrate_density = 1;
    o->enable_cell_volume_correction = 0;
    o->minimal_area = 1.0;
}

const char *fia_last_error(void) { return g_last_error.c_str(); }

int fia_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// A context on a new stream of its own (stream == nullptr) or on the caller's stream.
static fia_status ctx_create(int device, hipStream_t stream, fia_ctx **out) {
    if (!out) return fail(FIA_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(FIA_ERR_NO_DEVICE, "no GPU visible");
    int dev = device;
    if (dev < 0) HIP_TRY(hipGetDevice(&dev));
    if (dev >= n) return fail(FIA_ERR_INVALID_ARG, "device index out of range");
    fia_ctx *c = new fia_ctx();
    c->device = dev;
    c->own_stream = stream == nullptr;
    if (hipSetDevice(dev) != hipSuccess) {
        delete c;
        return fail(FIA_ERR_DEVICE, "cannot select the device");
    }
    if (c->own_stream) {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return fail(FIA_ERR_DEVICE, "cannot create the stream");
        }
    } else {
        c->stream = stream;
    }
    auto drop_stream = [&]() { if (c->own_stream) (void)hipStreamDestroy(c->stream); };
    if (hipHostMalloc((void **)&c->host_bounds, 4 * sizeof(unsigned long long)) != hipSuccess ||
        hipHostMalloc((void **)&c->host_flags, 4 * sizeof(int)) != hipSuccess ||
        hipHostMalloc((void **)&c->host_scal, 2 * kPcgScalars * sizeof(double)) != hipSuccess) {
        if (c->host_bounds) (void)hipHostFree(c->host_bounds);
        if (c->host_flags) (void)hipHostFree(c->host_flags);
        drop_stream();
        delete c;
        return fail(FIA_ERR_DEVICE, "cannot allocate pinned host memory");
    }
    if (io_scratch_init(dev) != hipSuccess) {
        (void)hipHostFree(c->host_bounds); (void)hipHostFree(c->host_flags); (void)hipHostFree(c->host_scal);
        drop_stream(); delete c;
        return fail(FIA_ERR_DEVICE, "cannot find the device scratch of the allocation-free entry points");
    }
    memset(c->host_bounds, 0, 4 * sizeof(unsigned long long));
    memset(c->host_flags, 0, 4 * sizeof(int));
    memset(c->host_scal, 0, 2 * kPcgScalars * sizeof(double));
    if (ensure_common(c) != FIA_OK) {
        (void)hipHostFree(c->host_bounds); (void)hipHostFree(c->host_flags); (void)hipHostFree(c->host_scal);
        drop_stream(); delete c; return FIA_ERR_DEVICE;
    }
    *out = c;
    return FIA_OK;
}

fia_status fia_ctx_create(int device, fia_ctx **out) { return ctx_create(device, nullptr, out); }

fia_status fia_ctx_create_on_stream(int device, void *stream, fia_ctx **out) {
    if (!stream) return fail(FIA_ERR_INVALID_ARG, "stream is NULL (use fia_ctx_create for a stream of its own)");
    return ctx_create(device, (hipStream_t)stream, out);
}

void *fia_ctx_stream(fia_ctx *c) { return c ? (void *)c->stream : nullptr; }

fia_status fia_ctx_destroy(fia_ctx *c) {
    if (!c) return FIA_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->psi.release(); c->twisted.release(); c->degen.release(); c->counters.release(); c->scal.release(); c->bits.release();
    c->solid.release(); c->genI.release();
    c->fsa.release(); c->fsb.release(); c->face_n.release(); c->face_start.release(); c->scan_tmp.release();
    c->piece_ay.release(); c->piece_val.release();
    for (int a = 0; a < 3; ++a) { c->piece_c[a].release(); c->gen_c[a].release(); }
    c->gen_n.release(); c->gen_start.release(); c->gen_ay.release(); c->gen_val.release();
    for (int a = 0; a < 3; ++a) { c->disp[a].release(); c->work[a].release(); c->gen[a].release(); }
    for (int a = 0; a < 4; ++a) c->grid[a].release();
    for (int a = 0; a < 2; ++a) c->genD[a].release();
    c->pcg_b.release(); c->pcg_flags.release(); c->pcgs[0].release(); c->pcgs[1].release(); c->mgb.release();
    for (auto &sv : c->solves) sv.release();
    c->solves.clear();
    c->batch_states.release(); c->batch_map.release(); c->batch_scal.release();
    c->graphs.release();
    if (c->host_batch) (void)hipHostFree(c->host_batch);
    c->bounds.release(); c->stats.release();
    for (auto &sp : c->spans) { (void)hipEventDestroy(sp.a); (void)hipEventDestroy(sp.b); }
    c->spans.clear();
    if (c->host_bounds) (void)hipHostFree(c->host_bounds);
    if (c->host_flags) (void)hipHostFree(c->host_flags);
    if (c->host_scal) (void)hipHostFree(c->host_scal);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return FIA_OK;
}

size_t fia_ctx_device_bytes(const fia_ctx *c) { return c ? ctx_bytes(c) : 0; }

fia_status fia_ctx_set_timing(fia_ctx *c, int enable) {
    if (!c) return fail(FIA_ERR_INVALID_ARG, "ctx is NULL");
    HIP_TRY(set_device(c));
    c->timing = enable != 0;
    if (c->timing) {
        HIP_TRY(c->stats.ensure(kStatWords * kStatSlots));
        HIP_TRY(hipMemsetAsync(c->stats.p, 0, kStatWords * kStatSlots * sizeof(unsigned long long), c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return FIA_OK;
}

fia_status fia_ctx_set_pressure_preconditioner(fia_ctx *c, fia_precond kind) {
    if (!c) return fail(FIA_ERR_INVALID_ARG, "ctx is NULL");
    if (kind != FIA_PRECOND_JACOBI && kind != FIA_PRECOND_MG) return fail(FIA_ERR_INVALID_ARG, "unknown preconditioner");
    c->precond = (int)kind;
    return FIA_OK;
}

fia_status fia_ctx_set_graphs(fia_ctx *c, int enable) {
    if (!c) return fail(FIA_ERR_INVALID_ARG, "ctx is NULL");
    HIP_TRY(set_device(c));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->graphs.enabled = enable != 0;
    if (!enable) c->graphs.release();
    return FIA_OK;
}

static bool tuning_valid(const fia_tuning *t) {
    auto blk = [](int b) { return b == 0 || b == 64 || b == 128 || b == 256; };
    return blk(t->face_block) && blk(t->piece_block) && blk(t->cell_block) && t->pcg_chunk_x >= 0 &&
           t->pcg_recompute >= -1 && t->pcg_recompute <= 1 && t->pcg_fused >= -1 && t->pcg_fused <= 1;
}

fia_status fia_ctx_set_tuning(fia_ctx *c, int nx, int ny, int nz, const fia_tuning *t) {
    if (!c || !t) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    if (!tuning_valid(t)) return fail(FIA_ERR_INVALID_ARG, "block sizes must be 0, 64, 128 or 256, the chunk >= 0 "
                                                           "and pcg_recompute and pcg_fused -1, 0 or 1");
    if (nx == 0 && ny == 0 && nz == 0) {
        c->tuning_default = *t;
        return FIA_OK;
    }
    if (nx <= 0 || ny <= 0 || nz <= 0) return fail(FIA_ERR_INVALID_ARG, "grid shape must be positive (or all zero)");
    for (auto &s : c->tunings)
        if (s.nx == nx && s.ny == ny && s.nz == nz) {
            s.t = *t;
            return FIA_OK;
        }
    if (c->tunings.size() >= 64) return fail(FIA_ERR_CAPACITY, "64 grid shapes already have their own launch shapes");
    c->tunings.push_back({nx, ny, nz, *t});
    return FIA_OK;
}

fia_status fia_ctx_get_tuning(fia_ctx *c, int nx, int ny, int nz, fia_tuning *t) {
    if (!c || !t) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    *t = tuning_for(c, nx, ny, nz);
    return FIA_OK;
}

namespace {
// The autotuner's copies of a grid and its timing events, released on every path.
struct TuneScratch {
    DevBuf<float> src[4], work[4];
    DevBuf<uint8_t> solid;
    hipEvent_t a = nullptr, b = nullptr;
    ~TuneScratch() {
        for (int k = 0; k < 4; ++k) { src[k].release(); work[k].release(); }
        solid.release();
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};
int effective_block(int b) { return (b == 64 || b == 128) ? b : 256; }
}  // namespace

// Coordinate descent over the launch shapes of g's grid shape, timed on copies of g: each block
// size in turn (faces, pieces, cells) over {64, 128, 256} by whole density steps, then the CG
// chunking by iterations of g's pressure system.  A candidate replaces the current choice only
// when it is at least 2% faster (timer noise does not move the choice back and forth).
fia_status fia_ctx_autotune(fia_ctx *c, const fia_grid3d *g, const uint8_t *solid, double dt,
                            const fia_advect_opts *o, int reps, fia_tuning *out) {
    if (!c) return fail(FIA_ERR_INVALID_ARG, "ctx is NULL");
    fia_status st;
    if ((st = check_grid(g, true)) != FIA_OK) return st;
    if ((st = check_opts(o)) != FIA_OK) return st;
    if (!(dt > 0.0) || !isfinite(dt)) return fail(FIA_ERR_INVALID_ARG, "dt must be positive and finite");
    if (reps < 1 || reps > 50) return fail(FIA_ERR_INVALID_ARG, "reps must be in 1..50");
    HIP_TRY(set_device(c));
    const size_t nc = (size_t)g->nx * g->ny * g->nz;
    const size_t len[4] = {nc, (size_t)(g->nx + 1) * g->ny * g->nz, (size_t)g->nx * (g->ny + 1) * g->nz,
                           (size_t)g->nx * g->ny * (g->nz + 1)};
    const float *fields[4] = {g->density, g->u, g->v, g->w};
    const hipMemcpyKind kind = g->on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    TuneScratch ts;
    for (int k = 0; k < 4; ++k) {
        HIP_TRY(ts.src[k].ensure(len[k]));
        HIP_TRY(ts.work[k].ensure(len[k]));
        HIP_TRY(hipMemcpyAsync(ts.src[k].p, fields[k], len[k] * 4, kind, c->stream));
    }
    if (solid) {
        HIP_TRY(ts.solid.ensure(nc));
        HIP_TRY(hipMemcpyAsync(ts.solid.p, solid, nc, kind, c->stream));
    }
    HIP_TRY(hipEventCreate(&ts.a));
    HIP_TRY(hipEventCreate(&ts.b));
    fia_grid3d wg = *g;
    wg.on_device = 1;
    wg.density = ts.work[0].p; wg.u = ts.work[1].p; wg.v = ts.work[2].p; wg.w = ts.work[3].p;
    auto restore = [&]() -> hipError_t {
        for (int k = 0; k < 4; ++k) {
            const hipError_t e = hipMemcpyAsync(ts.work[k].p, ts.src[k].p, len[k] * 4, hipMemcpyDeviceToDevice, c->stream);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    auto elapsed = [&](double &ms) -> hipError_t {
        hipError_t e = hipEventSynchronize(ts.b);
        float f = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&f, ts.a, ts.b);
        ms = f;
        return e;
    };
    const bool timing = c->timing;
    c->timing = false;
    struct RestoreTiming {
        fia_ctx *c;
        bool on;
        ~RestoreTiming() { c->timing = on; }
    } keep{c, timing};
    fia_tuning best = tuning_for(c, g->nx, g->ny, g->nz);
    if ((st = fia_ctx_set_tuning(c, g->nx, g->ny, g->nz, &best)) != FIA_OK) return st;
    // one density step on fresh copies: the fastest of reps timed runs after an untimed one
    auto time_step = [&](double &ms) -> fia_status {
        ms = INFINITY;
        for (int r = 0; r <= reps; ++r) {
            HIP_TRY(restore());
            HIP_TRY(hipEventRecord(ts.a, c->stream));
            fia_status s2 = fia_move_substances_3d(c, &wg, 0, dt, o);
            if (s2 != FIA_OK) return s2;
            HIP_TRY(hipEventRecord(ts.b, c->stream));
            double t;
            HIP_TRY(elapsed(t));
            if (r > 0) ms = std::min(ms, t);
        }
        return FIA_OK;
    };
    double best_ms;
    if ((st = time_step(best_ms)) != FIA_OK) return st;
    int fia_tuning::*const knobs[3] = {&fia_tuning::face_block, &fia_tuning::piece_block, &fia_tuning::cell_block};
    for (int k = 0; k < 3; ++k) {
        for (int b : {64, 128, 256}) {
            if (effective_block(best.*knobs[k]) == b) continue;
            fia_tuning t = best;
            t.*knobs[k] = b;
            if ((st = fia_ctx_set_tuning(c, g->nx, g->ny, g->nz, &t)) != FIA_OK) return st;
            double ms;
            if ((st = time_step(ms)) != FIA_OK) return st;
            if (ms < 0.98 * best_ms) { best = t; best_ms = ms; }
        }
        if ((st = fia_ctx_set_tuning(c, g->nx, g->ny, g->nz, &best)) != FIA_OK) return st;
    }
    // the CG sweep: iterations of g's pressure system (never converging: tolerance 0) under the
    // context's preconditioner, per chunk length and q policy
    HIP_TRY(restore());
    GridView gv{wg.density, wg.u, wg.v, wg.w};
    PcgGrid pg;
    if ((st = pressure_system(c, &wg, gv, solid ? ts.solid.p : nullptr, pg)) != FIA_OK) return st;
    const bool use_mg = c->precond == FIA_PRECOND_MG;
    const int iters = use_mg ? 8 : 24;
    auto time_cg = [&](const fia_tuning &t, double &ms) -> fia_status {
        PcgState ps;
        fia_status s2;
        if ((s2 = bind_solve(c->pcgs[0], pg, c->pcg_b.p, c->pcg_flags.p, ps, use_mg, t, false, false)) != FIA_OK)
            return s2;
        const MgHier *mg = nullptr;
        if ((s2 = mg_prepare(c, ps, &mg)) != FIA_OK) return s2;
        HIP_TRY(launch_pcg_guess(ps, nullptr, c->stream));
        if ((s2 = pcg_begin(c, ps, 0.0, mg)) != FIA_OK) return s2;
        HIP_TRY(launch_pcg_iterations(ps, 2, c->stream, mg));
        ms = INFINITY;
        for (int r = 0; r < reps; ++r) {
            HIP_TRY(hipEventRecord(ts.a, c->stream));
            HIP_TRY(launch_pcg_iterations(ps, iters, c->stream, mg));
            HIP_TRY(hipEventRecord(ts.b, c->stream));
            double t2;
            HIP_TRY(elapsed(t2));
            ms = std::min(ms, t2 / iters);
        }
        return FIA_OK;
    };
    double best_cg;
    if ((st = time_cg(best, best_cg)) != FIA_OK) return st;
    // a candidate's launches: chunk length, and the q policy of the two-sweep iteration or the
    // fused sweep (Jacobi only)
    const size_t ncells = (size_t)pg.nx * pg.ny * pg.nz;
    auto launch_key = [&](const fia_tuning &t) {
        PcgState ps;
        ps.fused = !use_mg && (t.pcg_fused == 1 || (t.pcg_fused == -1 && ncells >= ((size_t)1 << 22)));
        pcg_direction_tiling(pg.nx, pg.ny, pg.nz, ps, t.pcg_chunk_x, t.pcg_recompute);
        return std::make_pair(ps.dir_chunk_x, ps.fused ? 2 : (int)ps.recompute_q);
    };
    std::vector<std::pair<int, int>> tried{launch_key(best)};
    // chunk lengths: powers of two, and those that fill the device's resident blocks a half, one,
    // two or three times
    std::vector<int> chunks{2, 4, 8, 16, 32, 64};
    for (bool fused : {false, true}) {
        PcgState ps;
        ps.fused = fused;
        pcg_direction_tiling(pg.nx, pg.ny, pg.nz, ps, 0, -1);
        const long tiles = (long)ps.dir_tiles_y * ps.dir_tiles_z, res = std::max(1, pcg_resident_blocks(fused));
        for (long num : {1L, 2L, 4L, 6L}) {   // half-rounds
            const long want = std::max<long>(1, std::min<long>(pg.nx, num * res / 2 / tiles));
            chunks.push_back((int)((pg.nx + want - 1) / want));
        }
    }
    std::sort(chunks.begin(), chunks.end());
    chunks.erase(std::unique(chunks.begin(), chunks.end()), chunks.end());
    for (int chunk : chunks) {
        if (chunk > pg.nx) break;
        for (int mode : {0, 1, 2}) {   // stream q, recompute q, fused sweep
            if (mode == 2 && use_mg) continue;
            fia_tuning t = best;
            t.pcg_chunk_x = chunk;
            t.pcg_recompute = mode == 2 ? -1 : mode;
            t.pcg_fused = mode == 2 ? 1 : 0;
            PcgState ps;
            ps.fused = mode == 2;
            pcg_direction_tiling(pg.nx, pg.ny, pg.nz, ps, chunk, t.pcg_recompute);
            if (ps.dir_blocks > 4 * kPcgMaxBlocks) continue;   // every block re-sums the partials
            const auto key = launch_key(t);
            if (std::find(tried.begin(), tried.end(), key) != tried.end()) continue;
            tried.push_back(key);
            double ms;
            if ((st = time_cg(t, ms)) != FIA_OK) return st;
            if (ms < 0.98 * best_cg) { best = t; best_cg = ms; }
        }
    }
    if ((st = fia_ctx_set_tuning(c, g->nx, g->ny, g->nz, &best)) != FIA_OK) return st;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (out) *out = best;
    return FIA_OK;
}

fia_status fia_ctx_reset_stats(fia_ctx *c) {
    if (!c) return fail(FIA_ERR_INVALID_ARG, "ctx is NULL");
    HIP_TRY(set_device(c));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (auto &sp : c->spans) { (void)hipEventDestroy(sp.a); (void)hipEventDestroy(sp.b); }
    c->spans.clear();
    c->acc = fia_kernel_stats{};
    c->graph_captures0 = c->graphs.captures;
    c->graph_replays0 = c->graphs.replays;
    if (c->stats.p) {
        HIP_TRY(hipMemsetAsync(c->stats.p, 0, kStatWords * kStatSlots * sizeof(unsigned long long), c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return FIA_OK;
}

fia_status fia_ctx_get_stats(fia_ctx *c, fia_kernel_stats *out) {
    if (!c || !out) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    HIP_TRY(set_device(c));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (auto &sp : c->spans) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, sp.a, sp.b));
        if (sp.kind == SK_FACE) c->acc.face_ms += ms;
        else if (sp.kind == SK_VERTEX) c->acc.vertex_ms += ms;
        else if (sp.kind == SK_PSI) c->acc.psi_ms += ms;
        else if (sp.kind == SK_QUAD) c->acc.quad_ms += ms;
        else if (sp.kind == SK_PCG) c->acc.pcg_ms += ms;
        else c->acc.gather_ms += ms;
        (void)hipEventDestroy(sp.a);
        (void)hipEventDestroy(sp.b);
    }
    c->spans.clear();
    *out = c->acc;
    out->graph_captures = c->graphs.captures - c->graph_captures0;
    out->graph_replays = c->graphs.replays - c->graph_replays0;
    if (c->stats.p) {
        std::vector<unsigned long long> h(kStatWords * kStatSlots);
        HIP_TRY(hipMemcpy(h.data(), c->stats.p, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        long long sum[kStatWords] = {0, 0, 0};
        for (int i = 0; i < kStatSlots; ++i)
            for (int w = 0; w < kStatWords; ++w) sum[w] += (long long)h[kStatWords * i + w];
        out->pieces = sum[0];
        out->psi_points = sum[0];
        out->density_points = sum[1];
        out->weno_evals = sum[2];
    }
    return FIA_OK;
}

fia_status fia_move_substances_3d(fia_ctx *c, fia_grid3d *g, int i_frame, double dt, const fia_advect_opts *o) {
    (void)i_frame;
    if (!c) return fail(FIA_ERR_INVALID_ARG, "ctx is NULL");
    fia_status st;
    if ((st = check_grid(g, true)) != FIA_OK) return st;
    if ((st = check_opts(o)) != FIA_OK) return st;
    if (!isfinite(dt)) return fail(FIA_ERR_NONFINITE, "dt is not finite");
    HIP_TRY(set_device(c));
    if ((st = reset_error(c)) != FIA_OK) return st;
    GridView gv;
    if ((st = upload_grid(c, g, gv, true)) != FIA_OK) return st;
    CompGrid cg;
    cg.init(-1, g->nx, g->ny, g->nz);
    if ((st = run_step(c, cg, gv.rho, gv.rho, gv.u, gv.v, gv.w, g->dx, dt, o, o->split_method, FIA_SPLIT_XZ_AXIS)) != FIA_OK)
        return st;
    if (!g->on_device)
        HIP_TRY(hipMemcpyAsync(g->density, gv.rho, cg.cells() * 4, hipMemcpyDeviceToHost, c->stream));
    return check_kernel_errors(c);
}

fia_status fia_advect_flux_pass(fia_ctx *c, const fia_grid3d *g, float *q, double dt, fia_split split,
                                fia_interp interp, fia_psi_mode psi_mode, int zero_velocity, int use_integral,
                                int num_gauss_density, double minimal_area) {
    if (!c) return fail(FIA_ERR_INVALID_ARG, "ctx is NULL");
    fia_status st;
    if ((st = check_grid(g, false)) != FIA_OK) return st;
    if (!q) return fail(FIA_ERR_INVALID_ARG, "q is NULL");
    if (!gauss_supported(num_gauss_density)) return fail(FIA_ERR_INVALID_ARG, "unsupported Gauss rule");
    if (interp < FIA_INTERP_LINEAR || interp > FIA_INTERP_WENO6_OPTIMIZED)
        return fail(FIA_ERR_INVALID_ARG, "unknown interpolation");
    if (split < FIA_SPLIT_NONE || split > FIA_SPLIT_X_AXIS_Z_CELLFACE) return fail(FIA_ERR_INVALID_ARG, "unknown split");
    if (psi_mode < FIA_PSI_CONST || psi_mode > FIA_PSI_WENO6) return fail(FIA_ERR_INVALID_ARG, "unknown psi mode");
    if (!(minimal_area > 0.0)) return fail(FIA_ERR_INVALID_ARG, "minimal_area must be > 0");
    if (!isfinite(dt)) return fail(FIA_ERR_NONFINITE, "dt is not finite");
    HIP_TRY(set_device(c));
    if ((st = reset_error(c)) != FIA_OK) return st;
    GridView gv;
    if ((st = upload_grid(c, g, gv, false)) != FIA_OK) return st;
    CompGrid cg;
    cg.init(-1, g->nx, g->ny, g->nz);
    float *qd = q;
    if (!g->on_device) {
        HIP_TRY(c->work[0].ensure(cg.cells()));
        HIP_TRY(hipMemcpyAsync(c->work[0].p, q, cg.cells() * 4, hipMemcpyHostToDevice, c->stream));
        qd = c->work[0].p;
    }
    PassSpec s{(int)split, (int)interp, (int)psi_mode, zero_velocity ? 1 : 0, use_integral ? 1 : 0,
               num_gauss_density, minimal_area};
    if ((st = run_pass(c, cg, qd, qd, gv.u, gv.v, gv.w, g->dx, dt, s)) != FIA_OK) return st;
    if (!g->on_device) HIP_TRY(hipMemcpyAsync(q, qd, cg.cells() * 4, hipMemcpyDeviceToHost, c->stream));
    return check_kernel_errors(c);
}

fia_status fia_advect_velocity_3d(fia_ctx *c, fia_grid3d *g, int dim, double dt, const fia_advect_opts *o) {
    if (!c) return fail(FIA_ERR_INVALID_ARG, "ctx is NULL");
    fia_status st;
    if ((st = check_grid(g, false)) != FIA_OK) return st;
    if ((st = check_opts(o)) != FIA_OK) return st;
    if (dim < -1 || dim > 2) return fail(FIA_ERR_INVALID_ARG, "dim must be -1, 0, 1 or 2");
    if (o->interpolation_method == FIA_INTERP_WENO6_OPTIMIZED)
        return fail(FIA_ERR_UNSUPPORTED, "WENO6-optimized velocity: the reference's velocity quadrature has no branch "
                                         "for it (src/calc_velocity_from_quadrature_point_list_use_integral.cpp:85-111)");
    if (!isfinite(dt)) return fail(FIA_ERR_NONFINITE, "dt is not finite");
    HIP_TRY(set_device(c));
    if ((st = reset_error(c)) != FIA_OK) return st;
    GridView gv;
    if ((st = upload_grid(c, g, gv, false)) != FIA_OK) return st;
    float *comp_ptr[3] = {gv.u, gv.v, gv.w};
    float *host_ptr[3] = {g->u, g->v, g->w};
    const bool one_dy = o->interpolation_method == FIA_INTERP_1DY_LINEAR;
    for (int comp = 0; comp < 3; ++comp) {
        if (dim >= 0 && comp != dim) continue;
        CompGrid cg;
        cg.init(comp, g->nx, g->ny, g->nz);
        HIP_TRY(c->work[comp].ensure(cg.cells()));
        if ((st = run_step(c, cg, comp_ptr[comp], c->work[comp].p, gv.u, gv.v, gv.w, g->dx, dt, o,
                           velocity_split(comp, one_dy, false), velocity_split(comp, false, true))) != FIA_OK)
            return st;
    }
    // write back after all components: every component backtraces through the pre-step velocity
    // (reference src/advect_velocity_flux_advection_3d.cpp:1247-1271)
    for (int comp = 0; comp < 3; ++comp) {
        if (dim >= 0 && comp != dim) continue;
        CompGrid cg;
        cg.init(comp, g->nx, g->ny, g->nz);
        const size_t n = cg.cells();
        if (g->on_device) {
            hipLaunchKernelGGL(copy_floats_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                               c->work[comp].p, comp_ptr[comp], n);
            HIP_TRY(hipGetLastError());
        } else {
            HIP_TRY(hipMemcpyAsync(host_ptr[comp], c->work[comp].p, n * 4, hipMemcpyDeviceToHost, c->stream));
        }
    }
    return check_kernel_errors(c);
}

fia_status fia_calc_psi_y(fia_ctx *c, int nx, int ny, int nz, double dx, const float *q, double *psi,
                          fia_psi_mode mode, int use_integral, int num_gauss, int on_device) {
    if (!c || !q || !psi) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    if (nx < 1 || ny < 1 || nz < 1 || !(dx > 0)) return fail(FIA_ERR_INVALID_ARG, "bad grid");
    if (!gauss_supported(num_gauss)) return fail(FIA_ERR_INVALID_ARG, "unsupported Gauss rule");
    if (mode < FIA_PSI_CONST || mode > FIA_PSI_WENO6) return fail(FIA_ERR_INVALID_ARG, "unknown psi mode");
    HIP_TRY(set_device(c));
    CompGrid cg;
    cg.init(-1, nx, ny, nz);
    const float *qd = q;
    if (!on_device) {
        HIP_TRY(c->gen[0].ensure(cg.cells()));
        HIP_TRY(hipMemcpyAsync(c->gen[0].p, q, cg.cells() * 4, hipMemcpyHostToDevice, c->stream));
        qd = c->gen[0].p;
    }
    fia_status st;
    if ((st = do_psi(c, cg, qd, mode, use_integral, num_gauss)) != FIA_OK) return st;
    HIP_TRY(hipMemcpyAsync(psi, c->psi.p, cg.psi_size() * 8, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FIA_OK;
}

fia_status fia_interpolate(fia_ctx *c, const float *field, int mx, int my, int mz, double ox, double oy, double oz,
                           double dx, fia_interp method, const float *points_xyz, size_t n, float *out,
                           int on_device) {
    if (!c || !field || (!points_xyz && n) || (!out && n)) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    if (mx < 1 || my < 1 || mz < 1 || !(dx > 0)) return fail(FIA_ERR_INVALID_ARG, "bad field");
    if (method < FIA_INTERP_LINEAR || method > FIA_INTERP_WENO6_OPTIMIZED) return fail(FIA_ERR_INVALID_ARG, "unknown method");
    HIP_TRY(set_device(c));
    const size_t nf = (size_t)mx * my * mz;
    HIP_TRY(c->gen[0].ensure(nf)); HIP_TRY(c->gen[1].ensure(3 * n)); HIP_TRY(c->gen[2].ensure(n));
    HIP_TRY(hipMemcpyAsync(c->gen[0].p, field, nf * 4, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    // points to cell units
    std::vector<float> pts;
    if (!on_device) {
        pts.resize(3 * n);
        for (size_t i = 0; i < 3 * n; ++i) pts[i] = (float)((double)points_xyz[i] / dx);
        HIP_TRY(hipMemcpyAsync(c->gen[1].p, pts.data(), 3 * n * 4, hipMemcpyHostToDevice, c->stream));
    } else {
        return fail(FIA_ERR_UNSUPPORTED, "fia_interpolate takes host points");
    }
    HIP_TRY(launch_interpolate(c->gen[0].p, mx, my, mz, ox, oy, oz, (int)method, 1.0f, c->gen[1].p, n, c->gen[2].p,
                               c->stream));
    HIP_TRY(hipMemcpyAsync(out, c->gen[2].p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FIA_OK;
}

fia_status fia_backtrace_vertices(fia_ctx *c, const fia_grid3d *g, int dim, double dt, fia_interp interp,
                                  float *disp_xyz, int on_device) {
    if (!c || !disp_xyz) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    fia_status st;
    if ((st = check_grid(g, false)) != FIA_OK) return st;
    if (dim < -1 || dim > 2) return fail(FIA_ERR_INVALID_ARG, "dim must be -1..2");
    if (interp < FIA_INTERP_LINEAR || interp > FIA_INTERP_WENO6_OPTIMIZED) return fail(FIA_ERR_INVALID_ARG, "unknown interp");
    HIP_TRY(set_device(c));
    GridView gv;
    if ((st = upload_grid(c, g, gv, false)) != FIA_OK) return st;
    CompGrid cg;
    cg.init(dim, g->nx, g->ny, g->nz);
    if ((st = do_vertices(c, cg, gv.u, gv.v, gv.w, dt / g->dx, (int)interp)) != FIA_OK) return st;
    const size_t nl = cg.lattice();
    for (int a = 0; a < 3; ++a)
        HIP_TRY(hipMemcpyAsync(disp_xyz + a * nl, c->disp[a].p, nl * 4,
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FIA_OK;
}

fia_status fia_split_faces(fia_ctx *c, const float *quads_xyz, size_t n, double dx, fia_split split,
                           double minimal_area, int max_pieces, int *piece_count, double *centroid_xyz,
                           double *area_xyz, int on_device) {
    if (!c || (n && (!quads_xyz || !piece_count || !centroid_xyz || !area_xyz)))
        return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    if (max_pieces < 1 || !(dx > 0) || !(minimal_area > 0)) return fail(FIA_ERR_INVALID_ARG, "bad argument");
    if (on_device) return fail(FIA_ERR_UNSUPPORTED, "fia_split_faces takes host arrays");
    if (n == 0) return FIA_OK;
    HIP_TRY(set_device(c));
    fia_status st;
    if ((st = reset_error(c)) != FIA_OK) return st;
    std::vector<double> q(12 * n);
    for (size_t i = 0; i < 12 * n; ++i) q[i] = (double)quads_xyz[i] / dx;
    HIP_TRY(c->pcg_b.ensure(12 * n));
    HIP_TRY(c->genI.ensure(n));
    HIP_TRY(c->genD[0].ensure(3 * n * max_pieces));
    HIP_TRY(c->genD[1].ensure(3 * n * max_pieces));
    HIP_TRY(hipMemcpyAsync(c->pcg_b.p, q.data(), 12 * n * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_split_faces(c->pcg_b.p, n, (int)split, minimal_area, max_pieces, c->genI.p, c->genD[0].p,
                               c->genD[1].p, c->counters.p + 1, c->stream));
    HIP_TRY(hipMemcpyAsync(piece_count, c->genI.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    std::vector<double> cen(3 * n * max_pieces), ar(3 * n * max_pieces);
    HIP_TRY(hipMemcpyAsync(cen.data(), c->genD[0].p, cen.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(ar.data(), c->genD[1].p, ar.size() * 8, hipMemcpyDeviceToHost, c->stream));
    if ((st = check_kernel_errors(c)) != FIA_OK) return st;
    for (size_t i = 0; i < 3 * n * max_pieces; ++i) { centroid_xyz[i] = cen[i] * dx; area_xyz[i] = ar[i] * dx * dx; }
    return FIA_OK;
}

fia_status fia_cfl_time_step(fia_ctx *c, const fia_grid3d *g, double now_time, int i_movie_frame, int cfl_number,
                             double movie_fps, double *dt_out, int *write_frame) {
    if (!c || !dt_out || !write_frame) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    fia_status st;
    if ((st = check_grid(g, false)) != FIA_OK) return st;
    if (!(movie_fps > 0)) return fail(FIA_ERR_INVALID_ARG, "movie_fps must be > 0");
    HIP_TRY(set_device(c));
    GridView gv;
    if ((st = upload_grid(c, g, gv, false)) != FIA_OK) return st;
    HIP_TRY(c->bits.ensure(1));
    hipLaunchKernelGGL(zero_uints_kernel, dim3(1), dim3(64), 0, c->stream, c->bits.p, 1);
    HIP_TRY(hipGetLastError());
    const size_t nu = (size_t)(g->nx + 1) * g->ny * g->nz, nv = (size_t)g->nx * (g->ny + 1) * g->nz,
                 nw = (size_t)g->nx * g->ny * (g->nz + 1);
    HIP_TRY(launch_absmax(gv.u, nu, c->bits.p, c->stream));
    HIP_TRY(launch_absmax(gv.v, nv, c->bits.p, c->stream));
    HIP_TRY(launch_absmax(gv.w, nw, c->bits.p, c->stream));
    unsigned int bits = 0;
    HIP_TRY(hipMemcpyAsync(&bits, c->bits.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float m;
    memcpy(&m, &bits, 4);
    if (m != m) return fail(FIA_ERR_NONFINITE, "velocity contains NaN");
    // reference src/calc_time_step_length_from_CFL_number.cpp:122-131 (FP32 there)
    const double dt = (double)cfl_number * g->dx / (double)m;
    const double next = (i_movie_frame + 1) / movie_fps;
    if (now_time + dt > next) { *write_frame = 1; *dt_out = next - now_time; }
    else { *write_frame = 0; *dt_out = dt; }
    return FIA_OK;
}

fia_status fia_pressure_project(fia_ctx *c, fia_grid3d *g, const uint8_t *solid, float *pressure, int warm_start,
                                double rel_tol, int max_iters, int *iters, double *rel_residual) {
    if (!c) return fail(FIA_ERR_INVALID_ARG, "ctx is NULL");
    fia_status st;
    if ((st = check_grid(g, false)) != FIA_OK) return st;
    if (!(rel_tol > 0) || max_iters < 0) return fail(FIA_ERR_INVALID_ARG, "bad tolerance or iteration limit");
    HIP_TRY(set_device(c));
    GridView gv;
    if ((st = upload_grid(c, g, gv, false)) != FIA_OK) return st;
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    PcgGrid pg;
    if ((st = pressure_system(c, g, gv, solid, pg)) != FIA_OK) return st;
    PcgState ps;
    if ((st = pcg_bind(c, 0, pg, ps)) != FIA_OK) return st;
    float *pd = nullptr;
    if (pressure) {
        HIP_TRY(c->gen[0].ensure(n));
        pd = c->gen[0].p;
        if (g->on_device) pd = pressure;
        else if (warm_start) HIP_TRY(hipMemcpyAsync(pd, pressure, n * 4, hipMemcpyHostToDevice, c->stream));
    }
    const MgHier *mg = nullptr;
    if ((st = mg_prepare(c, ps, &mg)) != FIA_OK) return st;
    HIP_TRY(launch_pcg_guess(ps, (pressure && warm_start) ? pd : nullptr, c->stream));
    if ((st = pcg_begin(c, ps, rel_tol * rel_tol, mg)) != FIA_OK) return st;
    double *hs = c->host_scal;   // pinned
    if ((st = pcg_run(c, ps, max_iters, hs, mg)) != FIA_OK) return st;
    const int iters_done = (int)hs[kPcgIters];
    const double bb = hs[kPcgBB] > 0 ? hs[kPcgBB] : 1.0;
    const double res = sqrt(hs[kPcgRR] / bb);
    const bool converged = hs[kPcgDone] != 0.0;
    HIP_TRY(launch_project(pg, ps.x, g->dx, gv.u, gv.v, gv.w, pd, c->stream));
    if (!g->on_device) {
        const size_t nu = (size_t)(g->nx + 1) * g->ny * g->nz, nv = (size_t)g->nx * (g->ny + 1) * g->nz,
                     nw = (size_t)g->nx * g->ny * (g->nz + 1);
        HIP_TRY(hipMemcpyAsync(g->u, gv.u, nu * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(g->v, gv.v, nv * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(g->w, gv.w, nw * 4, hipMemcpyDeviceToHost, c->stream));
        if (pressure) HIP_TRY(hipMemcpyAsync(pressure, pd, n * 4, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (iters) *iters = iters_done;
    if (rel_residual) *rel_residual = res;
    if (!converged) return fail(FIA_ERR_NOT_CONVERGED, "PCG reached max_iters before rel_tol");
    return FIA_OK;
}


// ---- learned warm start -----------------------------------------------------------------------

size_t fia_unet_param_count(void) { return unet_param_count(); }

fia_status fia_unet_create(fia_ctx *c, unsigned long long seed, fia_unet **out) {
    if (!c || !out) return fail(FIA_ERR_INVALID_ARG, "ctx or out is NULL");
    *out = nullptr;
    HIP_TRY(set_device(c));
    fia_unet *u = new fia_unet;
    u->device = c->device;
    hipError_t e = u->net.init(seed, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        delete u;
        return fail(FIA_ERR_DEVICE, std::string("U-Net initialisation: ") + hipGetErrorString(e));
    }
    *out = u;
    return FIA_OK;
}

fia_status fia_unet_destroy(fia_unet *u) {
    if (!u) return FIA_OK;
    (void)hipSetDevice(u->device);
    u->y.release(); u->dy.release(); u->flags.release(); u->solid.release(); u->rhs.release();
    delete u;
    return FIA_OK;
}

fia_status fia_unet_set_params(fia_ctx *c, fia_unet *u, const float *params, size_t n) {
    if (!c || !u || !params) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    if (n != unet_param_count()) return fail(FIA_ERR_INVALID_ARG, "parameter count does not match the network");
    HIP_TRY(set_device(c));
    HIP_TRY(u->net.set_params(params, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FIA_OK;
}

fia_status fia_unet_get_params(fia_ctx *c, fia_unet *u, float *params, size_t n) {
    if (!c || !u || !params) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    if (n != unet_param_count()) return fail(FIA_ERR_INVALID_ARG, "parameter count does not match the network");
    HIP_TRY(set_device(c));
    HIP_TRY(u->net.get_params(params, c->stream));
    return FIA_OK;
}

namespace {

fia_status unet_grid(fia_unet *u, int nx, int ny, int nz) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx % 8 || ny % 8 || nz % 8)
        return fail(FIA_ERR_INVALID_ARG, "network grid sizes must be positive multiples of 8");
    const size_t n = (size_t)nx * ny * nz;
    HIP_TRY(u->net.ensure_grid(nx, ny, nz));
    HIP_TRY(u->y.ensure(n));
    HIP_TRY(u->dy.ensure(n));
    HIP_TRY(u->flags.ensure(n));
    return FIA_OK;
}

// device copies of host inputs (or the caller's device pointers)
fia_status unet_inputs(fia_ctx *c, fia_unet *u, size_t n, const double *rhs, const uint8_t *solid, int on_device,
                       const double *&rd, const uint8_t *&sd) {
    if (on_device) { rd = rhs; sd = solid; return FIA_OK; }
    HIP_TRY(u->rhs.ensure(n));
    HIP_TRY(hipMemcpyAsync(u->rhs.p, rhs, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    rd = u->rhs.p;
    sd = nullptr;
    if (solid) {
        HIP_TRY(u->solid.ensure(n));
        HIP_TRY(hipMemcpyAsync(u->solid.p, solid, n, hipMemcpyHostToDevice, c->stream));
        sd = u->solid.p;
    }
    return FIA_OK;
}

}  // namespace

fia_status fia_unet_forward(fia_ctx *c, fia_unet *u, int nx, int ny, int nz, const double *rhs, const uint8_t *solid,
                            float *y, double *scale, int on_device) {
    if (!c || !u || !rhs || !y) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    HIP_TRY(set_device(c));
    fia_status st;
    if ((st = unet_grid(u, nx, ny, nz)) != FIA_OK) return st;
    const size_t n = (size_t)nx * ny * nz;
    const double *rd; const uint8_t *sd;
    if ((st = unet_inputs(c, u, n, rhs, solid, on_device, rd, sd)) != FIA_OK) return st;
    double sc = 0.0;
    HIP_TRY(u->net.forward(rd, sd, on_device ? y : u->y.p, &sc, c->stream));
    if (!on_device) HIP_TRY(hipMemcpyAsync(y, u->y.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (scale) *scale = sc;
    return FIA_OK;
}

fia_status fia_unet_loss_grad(fia_ctx *c, fia_unet *u, int nx, int ny, int nz, const double *rhs, const uint8_t *solid,
                              int loss_mode, double *loss, double *residual, float *grad, int on_device) {
    if (!c || !u || !rhs) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    if (loss_mode != FIA_UNET_LOSS_RESIDUAL && loss_mode != FIA_UNET_LOSS_ENERGY)
        return fail(FIA_ERR_INVALID_ARG, "unknown loss mode");
    HIP_TRY(set_device(c));
    fia_status st;
    if ((st = unet_grid(u, nx, ny, nz)) != FIA_OK) return st;
    const size_t n = (size_t)nx * ny * nz;
    const double *rd; const uint8_t *sd;
    if ((st = unet_inputs(c, u, n, rhs, solid, on_device, rd, sd)) != FIA_OK) return st;
    HIP_TRY(launch_fluid_flags(sd, nx, ny, nz, u->flags.p, c->stream));
    HIP_TRY(u->net.forward(rd, sd, u->y.p, nullptr, c->stream));
    double l = 0.0, r = 0.0;
    HIP_TRY(u->net.loss(loss_mode, u->flags.p, rd, u->y.p, grad ? u->dy.p : nullptr, &l, &r, c->stream));
    if (grad) {
        HIP_TRY(u->net.zero_grad(c->stream));
        HIP_TRY(u->net.backward(u->dy.p, c->stream));
        HIP_TRY(hipMemcpyAsync(grad, u->net.grad(), unet_param_count() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (loss) *loss = l;
    if (residual) *residual = r;
    return FIA_OK;
}

fia_status fia_unet_train_step(fia_ctx *c, fia_unet *u, int nx, int ny, int nz, int batch, const double *rhs,
                               const uint8_t *solid, int loss_mode, float lr, double *mean_loss, double *mean_residual) {
    if (!c || !u || !rhs || batch <= 0 || !(lr > 0)) return fail(FIA_ERR_INVALID_ARG, "bad argument");
    if (loss_mode != FIA_UNET_LOSS_RESIDUAL && loss_mode != FIA_UNET_LOSS_ENERGY)
        return fail(FIA_ERR_INVALID_ARG, "unknown loss mode");
    HIP_TRY(set_device(c));
    fia_status st;
    if ((st = unet_grid(u, nx, ny, nz)) != FIA_OK) return st;
    const size_t n = (size_t)nx * ny * nz;
    HIP_TRY(u->net.zero_grad(c->stream));
    double sum = 0.0, sum_r = 0.0;
    for (int b = 0; b < batch; ++b) {
        const double *rd = rhs + (size_t)b * n;
        const uint8_t *sd = solid ? solid + (size_t)b * n : nullptr;
        HIP_TRY(launch_fluid_flags(sd, nx, ny, nz, u->flags.p, c->stream));
        HIP_TRY(u->net.forward(rd, sd, u->y.p, nullptr, c->stream));
        double l = 0.0, r = 0.0;
        HIP_TRY(u->net.loss(loss_mode, u->flags.p, rd, u->y.p, u->dy.p, &l, &r, c->stream));
        sum += l;
        sum_r += r;
        HIP_TRY(u->net.backward(u->dy.p, c->stream));
    }
    HIP_TRY(u->net.adam_step(lr, 0.9f, 0.999f, 1e-8f, 1.0f / (float)batch, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (mean_loss) *mean_loss = sum / batch;
    if (mean_residual) *mean_residual = sum_r / batch;
    return FIA_OK;
}

fia_status fia_pressure_rhs(fia_ctx *c, fia_grid3d *g, const uint8_t *solid, double *rhs) {
    if (!c || !rhs) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    fia_status st;
    if ((st = check_grid(g, false)) != FIA_OK) return st;
    HIP_TRY(set_device(c));
    GridView gv;
    if ((st = upload_grid(c, g, gv, false)) != FIA_OK) return st;
    // the system is assembled from a copy: obstacle faces are zeroed in the context's mirror only
    // when the grid lives on the host, so work on a device copy of the velocity either way
    const size_t nu = (size_t)(g->nx + 1) * g->ny * g->nz, nv = (size_t)g->nx * (g->ny + 1) * g->nz,
                 nw = (size_t)g->nx * g->ny * (g->nz + 1);
    if (g->on_device) {
        HIP_TRY(c->grid[1].ensure(nu)); HIP_TRY(c->grid[2].ensure(nv)); HIP_TRY(c->grid[3].ensure(nw));
        HIP_TRY(hipMemcpyAsync(c->grid[1].p, gv.u, nu * 4, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->grid[2].p, gv.v, nv * 4, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->grid[3].p, gv.w, nw * 4, hipMemcpyDeviceToDevice, c->stream));
        gv.u = c->grid[1].p; gv.v = c->grid[2].p; gv.w = c->grid[3].p;
    }
    PcgGrid pg;
    if ((st = pressure_system(c, g, gv, solid, pg)) != FIA_OK) return st;
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    HIP_TRY(hipMemcpyAsync(rhs, c->pcg_b.p, n * sizeof(double),
                           g->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FIA_OK;
}

fia_status fia_pressure_project_learned(fia_ctx *c, fia_grid3d *g, const uint8_t *solid, float *pressure, fia_unet *u,
                                        double rel_tol, int max_iters, int *iters, double *rel_residual,
                                        fia_learned_info *info) {
    if (!c || !u) return fail(FIA_ERR_INVALID_ARG, "ctx or net is NULL");
    fia_status st;
    if ((st = check_grid(g, false)) != FIA_OK) return st;
    if (!(rel_tol > 0) || max_iters < 0) return fail(FIA_ERR_INVALID_ARG, "bad tolerance or iteration limit");
    HIP_TRY(set_device(c));
    if ((st = unet_grid(u, g->nx, g->ny, g->nz)) != FIA_OK) return st;
    GridView gv;
    if ((st = upload_grid(c, g, gv, false)) != FIA_OK) return st;
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    PcgGrid pg;
    if ((st = pressure_system(c, g, gv, solid, pg)) != FIA_OK) return st;
    PcgState sa, sb;
    if ((st = pcg_bind(c, 0, pg, sa)) != FIA_OK) return st;
    if ((st = pcg_bind(c, 1, pg, sb)) != FIA_OK) return st;
    const double tol2 = rel_tol * rel_tol;
    const int probe = std::min(5, max_iters);
    double *ha = c->host_scal, *hb = c->host_scal + kPcgScalars;
    LearnedStart ls;
    const MgHier *mg = nullptr;
    if ((st = mg_prepare(c, sa, &mg)) != FIA_OK) return st;
    if ((st = learned_start(c, u, pg, sa, sb, tol2, probe, ha, hb, ls, mg)) != FIA_OK) return st;
    const bool use_net = ls.use_net;
    const double bb = ls.bb;
    PcgState &ps = use_net ? sa : sb;
    double *hs = use_net ? ha : hb;
    if ((st = pcg_run(c, ps, max_iters, hs, mg)) != FIA_OK) return st;
    const int iters_done = (int)hs[kPcgIters];
    const double res = sqrt(hs[kPcgRR] / bb);
    const bool converged = hs[kPcgDone] != 0.0;
    float *pd = nullptr;
    if (pressure) {
        HIP_TRY(c->gen[0].ensure(n));
        pd = g->on_device ? pressure : c->gen[0].p;
    }
    HIP_TRY(launch_project(pg, ps.x, g->dx, gv.u, gv.v, gv.w, pd, c->stream));
    if (!g->on_device) {
        const size_t nu = (size_t)(g->nx + 1) * g->ny * g->nz, nv = (size_t)g->nx * (g->ny + 1) * g->nz,
                     nw = (size_t)g->nx * g->ny * (g->nz + 1);
        HIP_TRY(hipMemcpyAsync(g->u, gv.u, nu * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(g->v, gv.v, nv * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(g->w, gv.w, nw * 4, hipMemcpyDeviceToHost, c->stream));
        if (pressure) HIP_TRY(hipMemcpyAsync(pressure, pd, n * 4, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (iters) *iters = iters_done;
    if (rel_residual) *rel_residual = res;
    if (info) {
        info->used_network = use_net ? 1 : 0;
        info->probe_iters = probe;
        info->guess_rel_residual = ls.guess_res;
        info->guess_scale = ls.gamma;
        info->probe_res_network = ls.ra;
        info->probe_res_zero = ls.rb;
    }
    if (!converged) return fail(FIA_ERR_NOT_CONVERGED, "PCG reached max_iters before rel_tol");
    return FIA_OK;
}

// ---- diagnostics of the reference's output series --------------------------------------------

fia_status fia_grid_totals(fia_ctx *c, fia_grid3d *g, double totals[4]) {
    if (!c || !totals) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    fia_status st;
    if ((st = check_grid(g, false)) != FIA_OK) return st;
    HIP_TRY(set_device(c));
    GridView gv;
    if ((st = upload_grid(c, g, gv, g->density != nullptr)) != FIA_OK) return st;
    const size_t nc = (size_t)g->nx * g->ny * g->nz;
    const float *f[4] = {g->density ? gv.rho : nullptr, gv.u, gv.v, gv.w};
    const size_t n[4] = {nc, (size_t)(g->nx + 1) * g->ny * g->nz, (size_t)g->nx * (g->ny + 1) * g->nz,
                         (size_t)g->nx * g->ny * (g->nz + 1)};
    HIP_TRY(c->genD[1].ensure(io_partial_words()));
    double *out = c->genD[1].p + io_partial_words() - 8;
    HIP_TRY(launch_field_sums(f, n, c->genD[1].p, out, c->stream));
    HIP_TRY(hipMemcpyAsync(c->host_scal, out, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double vol = g->dx * g->dx * g->dx;
    for (int k = 0; k < 4; ++k) totals[k] = c->host_scal[k] * vol;
    return FIA_OK;
}

fia_status fia_grid_totals_noalloc(fia_ctx *c, const fia_grid3d *g, double totals[4]) {
    if (!c || !totals) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    fia_status st;
    if ((st = check_grid(g, false)) != FIA_OK) return st;
    if (!g->on_device)
        return fail(FIA_ERR_INVALID_ARG, "fia_grid_totals_noalloc needs device fields (on_device = 1): host fields "
                                         "would need device copies");
    HIP_TRY(set_device(c));
    double *part = io_scratch(c->device);
    if (!part) return fail(FIA_ERR_DEVICE, "the device scratch of this GPU was not set up");
    const size_t nc = (size_t)g->nx * g->ny * g->nz;
    const float *f[4] = {g->density, g->u, g->v, g->w};
    const size_t n[4] = {nc, (size_t)(g->nx + 1) * g->ny * g->nz, (size_t)g->nx * (g->ny + 1) * g->nz,
                         (size_t)g->nx * g->ny * (g->nz + 1)};
    double *out = part + io_partial_words() - 8;
    // every partial and result is written by the call itself: nothing in the scratch needs clearing
    HIP_TRY(launch_field_sums(f, n, part, out, c->stream));
    HIP_TRY(hipMemcpyAsync(c->host_scal, out, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double vol = g->dx * g->dx * g->dx;
    for (int k = 0; k < 4; ++k) totals[k] = c->host_scal[k] * vol;
    return FIA_OK;
}

fia_status fia_wall_momentum_loss(fia_ctx *c, int nx, int ny, int nz, double dx, const float *pressure, int on_device,
                                  double loss[3]) {
    if (!c || !pressure || !loss) return fail(FIA_ERR_INVALID_ARG, "NULL argument");
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(dx > 0)) return fail(FIA_ERR_INVALID_ARG, "bad grid size");
    HIP_TRY(set_device(c));
    const size_t n = (size_t)nx * ny * nz;
    const float *pd = pressure;
    if (!on_device) {
        HIP_TRY(c->gen[0].ensure(n));
        HIP_TRY(hipMemcpyAsync(c->gen[0].p, pressure, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        pd = c->gen[0].p;
    }
    HIP_TRY(c->genD[1].ensure(io_partial_words()));
    double *out = c->genD[1].p + io_partial_words() - 8;
    HIP_TRY(launch_wall_sums(pd, nx, ny, nz, c->genD[1].p, out, c->stream));
    HIP_TRY(hipMemcpyAsync(c->host_scal, out, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double a = dx * dx;
    for (int k = 0; k < 3; ++k) loss[k] = a * (c->host_scal[2 * k + 1] - c->host_scal[2 * k]);
    return FIA_OK;
}

}  // extern "C"
