// Host build of the library's general face cutter and coordinate split (flux-interpolated-
// advection_amd/csrc/fia_cutter.hpp and fia_math.hpp, the same sources the GPU kernels compile),
// for tests/test_cutter_host.py.
#include "../flux-interpolated-advection_amd/csrc/fia_cutter.hpp"

extern "C" int host_general_cut(const double *col8, double min_area, int max_pieces, double *cen, double *area) {
    double col[4][2];
    for (int c = 0; c < 4; ++c) { col[c][0] = col8[2 * c]; col[c][1] = col8[2 * c + 1]; }
    int n = 0;
    bool over = false;
    auto emit = [&](const double (*pts)[2], int num, double cx, double cy) {
        if (n >= max_pieces) { over = true; return; }
        double cxx, cyy, a;
        fia::gen_center(pts, num, cxx, cyy, a);
        cen[2 * n] = cx; cen[2 * n + 1] = cy; area[n] = a;
        ++n;
    };
    const int code = fia::general_cut(col, min_area, emit);
    if (code != fia::KE_NONE) return -code;
    return over ? -1 : n;
}

// the stencil base and fraction of a sample coordinate, and whether a 6-wide stencil from base - 2
// counts as inside a field of m samples (the interior test of the quadrature's row loads)
extern "C" int host_split_coord(double s, int m, int *base, float *frac) {
    fia::split_coord(s, *base, *frac);
    return *base - 2 >= 0 && *base + 3 <= m - 1;
}

// mismatches of fia::FastDiv against hardware division: every divisor up to d_max with every
// numerator up to n_max, then `random` seeded pairs over the whole range n < 2^31
extern "C" long host_fastdiv_mismatches(unsigned d_max, unsigned n_max, long random) {
    long bad = 0;
    for (unsigned d = 1; d <= d_max; ++d) {
        const fia::FastDiv f(d);
        for (unsigned n = 0; n <= n_max; ++n) bad += f.div(n) != n / d;
        const unsigned top = 0x7fffffffu, last = d * (top / d);
        bad += f.div(top) != top / d;
        bad += f.div(last) != last / d;
        bad += f.div(last - 1) != (last - 1) / d;
    }
    unsigned long long x = 88172645463325252ull;
    for (long t = 0; t < random; ++t) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        const unsigned d = (unsigned)(x % 0x7fffffffu) + 1u, n = (unsigned)(x >> 33) & 0x7fffffffu;
        bad += fia::FastDiv(d).div(n) != n / d;
    }
    return bad;
}
