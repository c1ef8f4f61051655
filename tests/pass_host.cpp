// Host replay of the library's flux passes: the __host__ __device__ kernel bodies of
// flux-interpolated-advection_amd/csrc/fia_kernels.hip (the same source the GPU build compiles),
// driven over host arrays in the order fia_api.hip launches them, on host threads.
//
// The replay context mirrors fia_ctx's device scratch: buffers only grow, their contents are not
// kept when they grow, and nothing is cleared between calls.  New memory is filled with 0xFF bytes
// (NaN for floats, 0xffffffff for counts), so a result that reads an entry no kernel of the call
// wrote comes out NaN instead of silently right.  Each call checks the face lists, slot counts
// and slot ranges the kernels produced against their buffers, and reports the work of every pass
// (twisted and degenerate faces, piece slots, pieces of the general cutter) in a stats array.
//
// Built by tests/test_pass_host.py with hipcc (optionally with AddressSanitizer on the host code);
// only the host code runs.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "../flux-interpolated-advection_amd/csrc/fia_kernels.hip"

using namespace fia;

namespace {

int g_threads = 8;

template <class F> void parallel_for(long n, F fn) {
    if (n <= 0) return;
    const int t = (int)std::min<long>(g_threads, (n + 1023) / 1024);
    if (t <= 1) {
        for (long i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> th;
    for (int k = 0; k < t; ++k)
        th.emplace_back([&, k]() {
            const long a = n * k / t, b = n * (k + 1) / t;
            for (long i = a; i < b; ++i) fn(i);
        });
    for (auto &x : th) x.join();
}

// grow-only buffer (DevBuf's rules), new memory poisoned
template <class T> struct HostBuf {
    std::vector<T> v;
    T *p = nullptr;
    size_t n = 0;
    void ensure(size_t want) {
        if (want <= n && p) return;
        v.clear();
        v.shrink_to_fit();
        v.resize(std::max<size_t>(want, 1));
        memset((void *)v.data(), 0xFF, v.size() * sizeof(T));
        p = v.data();
        n = want;
    }
};

struct FluxSetH {
    HostBuf<double> f[3], c[3];
    HostBuf<unsigned> m[3];
    bool dirty = false;
    FluxView view() {
        FluxView v;
        for (int a = 0; a < 3; ++a) { v.f[a] = f[a].p; v.c[a] = c[a].p; v.m[a] = m[a].p; }
        return v;
    }
};

// replay counters per pass (stats rows of kStatCols doubles)
enum { ST_TWISTED, ST_DEGEN, ST_SLOTS, ST_GEN_PIECES, ST_GEN_MAX, ST_MULTI_FACES, ST_MAX_SLOTS, kStatCols };

struct HostCtx {
    HostBuf<double> psi;
    HostBuf<float> disp[3];
    FluxSetH fsa, fsb;
    HostBuf<unsigned> face_n;
    HostBuf<unsigned long long> face_start;
    HostBuf<double> piece_c[3], piece_ay, piece_val;
    HostBuf<float> work[3];
    HostBuf<unsigned> twisted, degen;
    HostBuf<unsigned> gen_n;
    HostBuf<unsigned long long> gen_start;
    HostBuf<double> gen_c[3], gen_ay, gen_val;
    int counters[4] = {0, 0, 0, 0};
    std::vector<double> stats;   // kStatCols per face pass
    char msg[256] = {0};
};

struct PassSpec {
    int split, interp, psi_mode, zero_velocity, integral, ng;
    double min_area;
};

GaussTable gauss_table(int n) {
    // Gauss-Legendre with weights summing to 1 (the same Newton iteration as fia_api.hip make_gauss)
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

void do_psi(HostCtx &c, const CompGrid &g, const float *q, int mode, int integral, int ng) {
    c.psi.ensure(g.psi_size());
    PsiParams p;
    memset(&p, 0, sizeof(p));
    p.g = g; p.q = q; p.psi = c.psi.p; p.mode = mode; p.integral = integral; p.kappa = 1.0f;
    p.gauss = gauss_table(ng);
    parallel_for((long)g.m[0] * g.m[2], [&](long col) { psi_body(p, col); });
}

void do_vertices(HostCtx &c, const CompGrid &g, const float *u, const float *v, const float *w, double dt_over_h,
                 int interp) {
    for (int a = 0; a < 3; ++a) c.disp[a].ensure(g.lattice());
    VertexParams p;
    memset(&p, 0, sizeof(p));
    p.g = g; p.u = u; p.v = v; p.w = w; p.dt_over_h = (float)dt_over_h; p.interp = interp;
    for (int a = 0; a < 3; ++a) p.disp[a] = c.disp[a].p;
    parallel_for((long)g.lattice(), [&](long id) { vertex_body(p, id); });
}

// returns 0, or a kernel error code, or -1 with c.msg set when a list or slot check fails
int do_faces(HostCtx &c, const CompGrid &g, const float *q, bool moving, FluxSetH &fs, const PassSpec &s, int fam_mask,
             double h) {
    double st[kStatCols] = {0};
    for (int f = 0; f < 3; ++f) {
        fs.f[f].ensure(g.faces(f));
        if (!fs.c[f].p || fs.c[f].n < g.faces(f)) {   // new corrections and masks start at zero
            fs.c[f].ensure(g.faces(f));
            fs.m[f].ensure((fs.c[f].n + 31) / 32);
            memset((void *)fs.c[f].p, 0, fs.c[f].n * sizeof(double));
            memset((void *)fs.m[f].p, 0, fs.m[f].n * sizeof(unsigned));
        }
    }
    if (fs.dirty) {   // the last pass's corrections back to zero (corr_reset_kernel)
        for (int f = 0; f < 3; ++f)
            parallel_for((long)fs.m[f].n, [&](long w) { corr_reset_word(fs.m[f].p, fs.c[f].p, (size_t)w); });
        fs.dirty = false;
    }
    const size_t nf = g.faces(0) + g.faces(1) + g.faces(2);
    c.twisted.ensure(nf);
    c.degen.ensure(nf);
    c.counters[0] = 0; c.counters[2] = 0;
    FaceParams p;
    memset(&p, 0, sizeof(p));
    p.g = g; p.q = q; p.psi = c.psi.p;
    for (int a = 0; a < 3; ++a) {
        p.disp[a] = moving ? c.disp[a].p : nullptr;
        p.flux[a] = fs.f[a].p;
        p.corr[a] = fs.c[a].p;
        p.corr_mask[a] = fs.m[a].p;
    }
    p.twisted = c.twisted.p; p.twisted_count = &c.counters[0]; p.error = &c.counters[1];
    p.degen = c.degen.p; p.degen_count = &c.counters[2];
    p.split = s.split; p.interp = s.interp; p.integral = s.integral; p.fam_mask = fam_mask;
    p.general_all = s.min_area < 1.0;
    p.fuse_value = !p.integral && (p.interp == FIA_INTERP_LINEAR || p.interp == FIA_INTERP_1DY_LINEAR);
    p.kappa_q = 1.0f;
    p.kappa_psi = (float)(h * h);
    p.min_area = s.min_area;
    p.gauss = gauss_table(s.ng);
    if (!moving && !p.general_all) {
        const bool fast = p.interp == IM_WENO6 && !p.integral && zero_velocity_cuts_both(p);
        parallel_for((long)g.faces(1), [&](long id) {
            if (fast) zero_velocity_face_body<true>(p, id);
            else zero_velocity_face_body<false>(p, id);
        });
        c.stats.insert(c.stats.end(), st, st + kStatCols);
        return c.counters[1];
    }
    c.face_n.ensure(nf);
    c.face_start.ensure(nf);
    p.face_n = c.face_n.p;
    p.face_start = c.face_start.p;
    p.fam_base[0] = 0; p.fam_base[1] = g.faces(0); p.fam_base[2] = g.faces(0) + g.faces(1);
    parallel_for((long)g.faces(0), [&](long id) { face_count_body<0>(p, id); });
    parallel_for((long)g.faces(1), [&](long id) { face_count_body<1>(p, id); });
    parallel_for((long)g.faces(2), [&](long id) { face_count_body<2>(p, id); });
    const int n_twisted = c.counters[0], n_degen = c.counters[2];
    if (n_twisted < 0 || (size_t)n_twisted > nf || n_degen < 0 || (size_t)n_degen > nf) {
        snprintf(c.msg, sizeof(c.msg), "face list counts out of range: twisted %d, degenerate %d, faces %zu", n_twisted,
                 n_degen, nf);
        return -1;
    }
    // exclusive scan of the slot counts (the three scan kernels' result)
    unsigned long long run = 0, bounds[4] = {0, 0, 0, 0};
    unsigned max_slots = 0;
    for (size_t k = 0; k < nf; ++k) {
        if (k == p.fam_base[1]) bounds[1] = run;
        if (k == p.fam_base[2]) bounds[2] = run;
        c.face_start.p[k] = run;
        run += c.face_n.p[k];
        max_slots = std::max(max_slots, c.face_n.p[k]);
    }
    bounds[3] = run;
    const unsigned long long nslots = bounds[3];
    for (int a = 0; a < 3; ++a) c.piece_c[a].ensure(nslots);
    c.piece_ay.ensure(nslots);
    c.piece_val.ensure(nslots);
    for (int a = 0; a < 3; ++a) p.piece_c[a] = c.piece_c[a].p;
    p.piece_ay = c.piece_ay.p;
    p.piece_val = c.piece_val.p;
    // every listed face is a face of the grid, listed once
    std::vector<char> seen(nf, 0);
    for (int t = 0; t < n_twisted + n_degen; ++t) {
        const unsigned code = t < n_twisted ? c.twisted.p[t] : c.degen.p[t - n_twisted];
        const unsigned fam = code >> 30, id = code & ((1u << 30) - 1u);
        if (fam > 2 || id >= g.faces(fam) || seen[p.fam_base[fam] + id]) {
            snprintf(c.msg, sizeof(c.msg), "bad face list entry %u (family %u, id %u)", code, fam, id);
            return -1;
        }
        seen[p.fam_base[fam] + id] = 1;
        if ((t < n_twisted) != (c.face_n.p[p.fam_base[fam] + id] == 0)) {
            snprintf(c.msg, sizeof(c.msg), "twisted face with slots, or degenerate face without");
            return -1;
        }
    }
    // the one-pass kernels, the block's multi-cut faces piece by piece
    std::atomic<long> multi{0};
    auto seg = [&](auto famc, long id) {
        constexpr int F = decltype(famc)::value;
        const unsigned n = seg_face_body<F>(p, id);
        if (n) {
            ++multi;
            for (unsigned l = 0; l < n; ++l) seg_piece_body<F>(p, id, l);
        }
    };
    if (fam_mask & 1) parallel_for((long)g.faces(0), [&](long id) { seg(std::integral_constant<int, 0>(), id); });
    if (fam_mask & 2) parallel_for((long)g.faces(1), [&](long id) { seg(std::integral_constant<int, 1>(), id); });
    if (fam_mask & 4) parallel_for((long)g.faces(2), [&](long id) { seg(std::integral_constant<int, 2>(), id); });
    parallel_for(n_degen, [&](long t) { degenerate_body(p, (int)t); });
    if (p.fuse_value) parallel_for(n_degen, [&](long t) { degenerate_value_body(p, (int)t); });
    if (!p.fuse_value) {
        const int qm = piece_quadrature_mode(p);
        parallel_for((long)nslots, [&](long q) {
            int a = 0, b = 0, e = 0;
            if (qm == 1) piece_body<1>(p, (unsigned long long)q, a, b, e);
            else if (qm == 2) piece_body<2>(p, (unsigned long long)q, a, b, e);
            else piece_body<0>(p, (unsigned long long)q, a, b, e);
        });
    }
    if (fam_mask & 1) parallel_for((long)g.faces(0), [&](long id) { face_sum_body<0>(p, id); });
    if (fam_mask & 2) parallel_for((long)g.faces(1), [&](long id) { face_sum_body<1>(p, id); });
    if (fam_mask & 4) parallel_for((long)g.faces(2), [&](long id) { face_sum_body<2>(p, id); });
    // twisted faces: pieces counted per face and orientation, scanned, stored, integrated, summed
    long gen_pieces = 0, gen_max = 0;
    if (n_twisted > 0) {
        c.gen_n.ensure(2 * (size_t)n_twisted);
        c.gen_start.ensure(2 * (size_t)n_twisted + 1);
        p.gen_n = c.gen_n.p;
        p.gen_start = c.gen_start.p;
        parallel_for(n_twisted, [&](long t) { general_count_body(p, (int)t); });
        if (c.counters[1] != 0) return c.counters[1];
        unsigned long long run2 = 0;
        for (int t = 0; t < 2 * n_twisted; ++t) {
            c.gen_start.p[t] = run2;
            run2 += c.gen_n.p[t];
            if (t % 2 == 1) gen_max = std::max(gen_max, (long)(c.gen_n.p[t - 1] + c.gen_n.p[t]));
        }
        c.gen_start.p[2 * n_twisted] = run2;
        gen_pieces = (long)run2;
        for (int a = 0; a < 3; ++a) c.gen_c[a].ensure(run2);
        c.gen_ay.ensure(run2);
        c.gen_val.ensure(run2);
        for (int a = 0; a < 3; ++a) p.gen_c[a] = c.gen_c[a].p;
        p.gen_ay = c.gen_ay.p;
        p.gen_val = c.gen_val.p;
        parallel_for(n_twisted, [&](long t) { general_emit_body(p, (int)t); });
        parallel_for((long)run2, [&](long q) { general_piece_body(p, (unsigned long long)q); });
        fs.dirty = true;
        parallel_for(n_twisted, [&](long t) { general_sum_body(p, (int)t); });
    }
    st[ST_TWISTED] = n_twisted; st[ST_DEGEN] = n_degen; st[ST_SLOTS] = (double)nslots;
    st[ST_GEN_PIECES] = (double)gen_pieces; st[ST_GEN_MAX] = (double)gen_max; st[ST_MULTI_FACES] = (double)multi.load();
    st[ST_MAX_SLOTS] = max_slots;
    c.stats.insert(c.stats.end(), st, st + kStatCols);
    return c.counters[1];
}

void gather(const CompGrid &g, FluxSetH &s1, FluxSetH *s2, const float *base, float *out, int mode, int fm) {
    GatherArgs A;
    A.g = g;
    A.s1 = s1.view();
    if (s2) A.s2 = s2->view();
    else memset(&A.s2, 0, sizeof(A.s2));
    A.base = base; A.out = out; A.mode = mode; A.fam_mask = fm;
    parallel_for((long)g.cells(), [&](long id) { gather_body(A, id); });
}

int psi_mode_for(int interp, int &integral) {
    if (interp == FIA_INTERP_LINEAR) return FIA_PSI_LINEAR;
    if (interp == FIA_INTERP_1DY_LINEAR) { integral = 0; return FIA_PSI_CONST; }
    return FIA_PSI_WENO6;
}

// fia_api.hip run_step
int run_step(HostCtx &c, const CompGrid &g, const float *q, float *out, const float *u, const float *v, const float *w,
             double h, double dt, int interp, int use_integral, int corr, int ng, double min_area, int split_main,
             int split_err) {
    int integral = use_integral ? 1 : 0;
    const int pm = psi_mode_for(interp, integral);
    PassSpec p1{split_main, interp, pm, 0, integral, ng, min_area};
    int e;
    do_psi(c, g, q, p1.psi_mode, p1.integral, p1.ng);
    do_vertices(c, g, u, v, w, dt / h, p1.interp);
    if ((e = do_faces(c, g, q, true, c.fsa, p1, 7, h)) != 0) return e;
    if (!corr) {
        gather(g, c.fsa, nullptr, nullptr, out, 0, 7);
        return 0;
    }
    c.work[2].ensure(g.cells());
    PassSpec p2{split_main, interp, FIA_PSI_CONST, 1, 0, ng, min_area};
    do_psi(c, g, q, p2.psi_mode, p2.integral, p2.ng);
    if ((e = do_faces(c, g, q, false, c.fsb, p2, 2, h)) != 0) return e;
    float *err = c.work[2].p;
    gather(g, c.fsb, nullptr, q, err, 2, 2);
    PassSpec p3{split_err, FIA_INTERP_1DY_LINEAR, FIA_PSI_CONST, 0, 0, ng, min_area};
    do_psi(c, g, err, p3.psi_mode, p3.integral, p3.ng);
    do_vertices(c, g, u, v, w, dt / h, p3.interp);
    if ((e = do_faces(c, g, err, true, c.fsb, p3, 7, h)) != 0) return e;
    gather(g, c.fsa, &c.fsb, nullptr, out, 1, 7);
    return 0;
}

int velocity_split(int comp, bool one_dy, bool error_pass) {
    const bool flip = one_dy != error_pass;
    if (comp == 0) return flip ? FIA_SPLIT_X_CELLFACE_Z_AXIS : FIA_SPLIT_X_AXIS_Z_CELLFACE;
    if (comp == 1) return flip ? FIA_SPLIT_XZ_AXIS : FIA_SPLIT_XZ_CELLFACE;
    return flip ? FIA_SPLIT_X_AXIS_Z_CELLFACE : FIA_SPLIT_X_CELLFACE_Z_AXIS;
}

void copy_stats(const HostCtx &c, double *stats, int max_rows, int *rows) {
    const int n = (int)(c.stats.size() / kStatCols);
    if (rows) *rows = n;
    if (stats) memcpy(stats, c.stats.data(), sizeof(double) * kStatCols * std::min(n, max_rows));
}

}  // namespace

extern "C" {

int host_stat_cols(void) { return kStatCols; }
void host_set_threads(int n) { g_threads = std::max(1, n); }

void *host_ctx_create(void) { return new HostCtx(); }
void host_ctx_destroy(void *c) { delete (HostCtx *)c; }
const char *host_ctx_message(void *c) { return ((HostCtx *)c)->msg; }

// fia_move_substances_3d on host arrays, in place; returns 0 or a kernel error code (-1: a list or
// slot check failed, see host_ctx_message).  stats: up to max_rows rows of host_stat_cols() values
// (one per face pass of the call), rows: how many there were.
int host_move_substances(void *ctx, int nx, int ny, int nz, double h, float *rho, const float *u, const float *v,
                         const float *w, double dt, int split, int interp, int use_integral, int corr, int ng,
                         double min_area, double *stats, int max_rows, int *rows) {
    HostCtx &c = *(HostCtx *)ctx;
    c.stats.clear();
    c.counters[1] = 0;
    CompGrid cg;
    cg.init(-1, nx, ny, nz);
    const int e = run_step(c, cg, rho, rho, u, v, w, h, dt, interp, use_integral, corr, ng, min_area, split,
                           FIA_SPLIT_XZ_AXIS);
    copy_stats(c, stats, max_rows, rows);
    return e;
}

// fia_advect_velocity_3d (dim -1: all three) on host arrays, in place
int host_advect_velocity(void *ctx, int nx, int ny, int nz, double h, float *u, float *v, float *w, int dim, double dt,
                         int interp, int use_integral, int corr, int ng, double min_area, double *stats, int max_rows,
                         int *rows) {
    HostCtx &c = *(HostCtx *)ctx;
    c.stats.clear();
    c.counters[1] = 0;
    float *comp_ptr[3] = {u, v, w};
    const bool one_dy = interp == FIA_INTERP_1DY_LINEAR;
    for (int comp = 0; comp < 3; ++comp) {
        if (dim >= 0 && comp != dim) continue;
        CompGrid cg;
        cg.init(comp, nx, ny, nz);
        c.work[comp].ensure(cg.cells());
        const int e = run_step(c, cg, comp_ptr[comp], c.work[comp].p, u, v, w, h, dt, interp, use_integral, corr, ng,
                               min_area, velocity_split(comp, one_dy, false), velocity_split(comp, false, true));
        if (e) { copy_stats(c, stats, max_rows, rows); return e; }
    }
    for (int comp = 0; comp < 3; ++comp) {
        if (dim >= 0 && comp != dim) continue;
        CompGrid cg;
        cg.init(comp, nx, ny, nz);
        memcpy(comp_ptr[comp], c.work[comp].p, cg.cells() * sizeof(float));
    }
    copy_stats(c, stats, max_rows, rows);
    return 0;
}

// fia_advect_flux_pass on host arrays: one pass of a field q on component grid comp (-1: cells), in
// place
int host_flux_pass(void *ctx, int comp, int nx, int ny, int nz, double h, float *q, const float *u, const float *v,
                   const float *w, double dt, int split, int interp, int psi_mode, int zero_velocity, int use_integral,
                   int ng, double min_area, double *stats, int max_rows, int *rows) {
    HostCtx &c = *(HostCtx *)ctx;
    c.stats.clear();
    c.counters[1] = 0;
    CompGrid cg;
    cg.init(comp, nx, ny, nz);
    PassSpec s{split, interp, psi_mode, zero_velocity ? 1 : 0, use_integral ? 1 : 0, ng, min_area};
    do_psi(c, cg, q, s.psi_mode, s.integral, s.ng);
    if (!s.zero_velocity) do_vertices(c, cg, u, v, w, dt / h, s.interp);
    const int fm = s.zero_velocity ? 2 : 7;
    const int e = do_faces(c, cg, q, !s.zero_velocity, c.fsa, s, fm, h);
    if (e == 0) gather(cg, c.fsa, nullptr, nullptr, q, 0, fm);
    copy_stats(c, stats, max_rows, rows);
    return e;
}

}  // extern "C"

// the vertex displacements (cells) of the context's last vertex pass, component a, n values
extern "C" int host_ctx_disp(void *ctx, int a, float *out, long n) {
    HostCtx &c = *(HostCtx *)ctx;
    if (a < 0 || a > 2 || (size_t)n > c.disp[a].n) return -1;
    memcpy(out, c.disp[a].p, n * sizeof(float));
    return 0;
}

// face fluxes of the context's first flux set (family f, n values): the flux of each face for the
// cell below it, and the correction of the cell above it (nonzero for general-cutter faces)
extern "C" int host_ctx_fluxes(void *ctx, int f, double *flux, double *corr, long n) {
    HostCtx &c = *(HostCtx *)ctx;
    if (f < 0 || f > 2 || (size_t)n > c.fsa.f[f].n) return -1;
    memcpy(flux, c.fsa.f[f].p, n * sizeof(double));
    memcpy(corr, c.fsa.c[f].p, n * sizeof(double));
    return 0;
}

// the piece slots of face `id` of family f in the context's last face pass: count, and per slot
// the record (centre x, y, z in cells, y area) and value
extern "C" int host_ctx_face_pieces(void *ctx, int comp, int nx, int ny, int nz, int f, long id, double *rec,
                                    double *val, int max) {
    HostCtx &c = *(HostCtx *)ctx;
    CompGrid g;
    g.init(comp, nx, ny, nz);
    const size_t base = f == 0 ? 0 : (f == 1 ? g.faces(0) : g.faces(0) + g.faces(1));
    const unsigned n = c.face_n.p[base + id];
    const unsigned long long s0 = c.face_start.p[base + id];
    for (unsigned k = 0; k < n && (int)k < max; ++k) {
        rec[4 * k] = c.piece_c[0].p[s0 + k]; rec[4 * k + 1] = c.piece_c[1].p[s0 + k];
        rec[4 * k + 2] = c.piece_c[2].p[s0 + k]; rec[4 * k + 3] = c.piece_ay.p[s0 + k];
        val[k] = c.piece_val.p[s0 + k];
    }
    return (int)n;
}
