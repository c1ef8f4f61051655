// fia_smoke_simulation.hpp - header-only C++ drop-in for the reference's flux-interpolated advection
// entry points, implemented on top of the C ABI in fia.h (link with libfia.so).
//
// A caller of the reference keeps its code: the same namespace, the same Grid_3D members the flux
// path touches, and the same free-function signatures (reference include/grid_3d.h:14-88,
// include/move_substances_3d.h:48-67, include/advect_velocity_flux_advection_3d.h:14-124,
// include/calc_pressure_3d.h:7-9, include/calc_time_step_length_from_CFL_number.h:33-39).
// Strings are parsed once per call into the C option enums; the fields live in host
// std::vectors as in the reference and each call copies them to the GPU and back.  Failures
// throw std::runtime_error with the library's message (the reference calls exit(0) or assert()).
#ifndef FIA_SMOKE_SIMULATION_HPP
#define FIA_SMOKE_SIMULATION_HPP

#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fia.h"

namespace smoke_simulation {

typedef float MY_FLOAT_TYPE;   // the reference's default (include/define_float_type.h:5-8)

// The members of the reference's Grid_3D that the flux path reads or writes, same names and
// [x][y][z] z-fastest layout; the domain starts at the origin (min_pos = 0).
class Grid_3D {
public:
    const int Grid_num_x, Grid_num_y, Grid_num_z;
    const MY_FLOAT_TYPE min_pos_x, min_pos_y, min_pos_z;
    const MY_FLOAT_TYPE _cell_length, _cell_face_area, _cell_volume;
    std::vector<MY_FLOAT_TYPE> velocity_in_voxel_face_x, velocity_in_voxel_face_y, velocity_in_voxel_face_z;
    std::vector<MY_FLOAT_TYPE> pressure;
    std::vector<MY_FLOAT_TYPE> substance_density;
    // momentum taken out by the walls through the pressure, summed over projections
    // (reference include/grid_3d.h:52-54, src/calc_pressure_3d.cpp:235-266)
    MY_FLOAT_TYPE total_loss_of_momentum_x = 0, total_loss_of_momentum_y = 0, total_loss_of_momentum_z = 0;

    Grid_3D(int nx, int ny, int nz, MY_FLOAT_TYPE cell_length)
        : Grid_num_x(nx), Grid_num_y(ny), Grid_num_z(nz), min_pos_x(0), min_pos_y(0), min_pos_z(0),
          _cell_length(cell_length), _cell_face_area(cell_length * cell_length),
          _cell_volume(cell_length * cell_length * cell_length),
          velocity_in_voxel_face_x((size_t)(nx + 1) * ny * nz, 0.0f),
          velocity_in_voxel_face_y((size_t)nx * (ny + 1) * nz, 0.0f),
          velocity_in_voxel_face_z((size_t)nx * ny * (nz + 1), 0.0f),
          pressure((size_t)nx * ny * nz, 0.0f),
          substance_density((size_t)nx * ny * nz, 0.0f) {}
};

namespace fia_shim {

inline void check(fia_status s) {
    if (s != FIA_OK) throw std::runtime_error(std::string("libfia: ") + fia_last_error());
}

// One library context per process (device scratch is kept between calls of the same size).  Its
// pressure solves use the multigrid preconditioner, the counterpart of the reference's AMG
// (amgcl smoothed aggregation, src/calc_pressure_3d.cpp:141-166).  Velocity advection, projection
// and density advection share it, in that order every step.  An earlier version of the library
// once stalled on a 256^3 density step that followed a velocity step in one context, and no cause
// was found (host replays of the same sequence complete); the sequence is the library's own
// benchmark step and has run there since without a stall.
inline fia_ctx *context() {
    static fia_ctx *ctx = nullptr;
    if (!ctx) {
        check(fia_ctx_create(-1, &ctx));
        check(fia_ctx_set_pressure_preconditioner(ctx, FIA_PRECOND_MG));
    }
    return ctx;
}

inline fia_grid3d host_grid(Grid_3D &g) {
    fia_grid3d c;
    c.nx = g.Grid_num_x; c.ny = g.Grid_num_y; c.nz = g.Grid_num_z;
    c.dx = g._cell_length;
    c.density = g.substance_density.data();
    c.u = g.velocity_in_voxel_face_x.data();
    c.v = g.velocity_in_voxel_face_y.data();
    c.w = g.velocity_in_voxel_face_z.data();
    c.on_device = 0;
    return c;
}

// split strings of the reference (src/split_face_3D.cpp:26-43); anything else cuts nothing
inline fia_split split_of(const std::string &s) {
    if (s == "xz-CellFaceAligned") return FIA_SPLIT_XZ_CELLFACE;
    if (s == "xz-AxisAligned") return FIA_SPLIT_XZ_AXIS;
    if (s == "x-CellFaceAligned_z-AxisAligned") return FIA_SPLIT_X_CELLFACE_Z_AXIS;
    if (s == "x-AxisAligned_z-CellFaceAligned") return FIA_SPLIT_X_AXIS_Z_CELLFACE;
    return FIA_SPLIT_NONE;
}

inline fia_interp interp_of(const std::string &s) {
    if (s == "WENO6") return FIA_INTERP_WENO6;
    if (s == "linear") return FIA_INTERP_LINEAR;
    if (s == "1Dy_linear") return FIA_INTERP_1DY_LINEAR;
    if (s == "1Dy_WENO6") return FIA_INTERP_1DY_WENO6;
    if (s == "WENO6-optimized") return FIA_INTERP_WENO6_OPTIMIZED;   // density only (velocity throws)
    throw std::runtime_error("libfia: interpolation_method \"" + s + "\" is not implemented");
}

inline fia_advect_opts opts_of(bool use_flux_advection, bool use_MacCormack_scheme, bool use_clamping,
                               int num_gauss_quad_boundary, int num_gauss_quad_bulk, bool enable_zero_velocity,
                               const std::string &split_method, int num_gauss_quadrature_points,
                               const std::string &interpolation_method, bool use_integral,
                               int num_gauss_density, bool enable_cell_volume_correction, MY_FLOAT_TYPE minimal_area) {
    fia_advect_opts o;
    fia_default_opts(&o);
    o.use_flux_advection = use_flux_advection;
    o.use_maccormack = use_MacCormack_scheme;
    o.use_clamping_in_maccormack = use_clamping;
    o.num_gauss_quad_boundary = num_gauss_quad_boundary;
    o.num_gauss_quad_bulk = num_gauss_quad_bulk;
    o.enable_zero_velocity_correction = enable_zero_velocity;
    o.split_method = split_of(split_method);
    o.num_gauss_quadrature_points = num_gauss_quadrature_points;
    o.interpolation_method = interp_of(interpolation_method);
    o.use_integral = use_integral;
    o.num_gauss_points_integrate_density = num_gauss_density;
    o.enable_cell_volume_correction = enable_cell_volume_correction;
    o.minimal_area = minimal_area;
    return o;
}

}  // namespace fia_shim

// reference include/move_substances_3d.h:48-67
inline void move_substances_3D(Grid_3D &all_grid, int i_frame, const MY_FLOAT_TYPE time_step_length,
                               const bool use_flux_advection, const bool use_MacCormack_scheme,
                               const bool use_clamping_in_MacCormack_scheme, const int num_gauss_quad_boundary,
                               const int num_gauss_quad_bulk, const bool enable_eliminate_zero_velocity_false_diffusion,
                               const std::string split_method, const std::string integral_method,
                               const int num_gauss_quadrature_points, const std::string interpolation_method,
                               bool use_integral, const int num_gauss_quadrature_point_for_integrate_density,
                               const bool enable_cell_volume_correction, const MY_FLOAT_TYPE minimal_area) {
    (void)integral_method;   // the 3D flux path does not read it
    const fia_advect_opts o = fia_shim::opts_of(
        use_flux_advection, use_MacCormack_scheme, use_clamping_in_MacCormack_scheme, num_gauss_quad_boundary,
        num_gauss_quad_bulk, enable_eliminate_zero_velocity_false_diffusion, split_method, num_gauss_quadrature_points,
        interpolation_method, use_integral, num_gauss_quadrature_point_for_integrate_density,
        enable_cell_volume_correction, minimal_area);
    fia_grid3d g = fia_shim::host_grid(all_grid);
    fia_shim::check(fia_move_substances_3d(fia_shim::context(), &g, i_frame, time_step_length, &o));
}

// reference include/advect_velocity_flux_advection_3d.h:107-124 (all three components; the split
// of each component is fixed by the reference, so there is no split argument)
inline void advect_velocity_xyz_flux_advection_3D_std_thread(
    Grid_3D &all_grid, const MY_FLOAT_TYPE time_step_length, const bool use_MacCormack_scheme,
    const bool use_clamping_in_MacCormack_scheme, const int num_gauss_quad_boundary, const int num_gauss_quad_bulk,
    const bool enable_eliminate_zero_velocity_false_diffusion, const std::string integral_method,
    const int num_gauss_quadrature_points_on_one_triangle, const std::string interpolation_method, bool use_integral,
    const int num_gauss_quadrature_point_for_integrate_density, const bool enable_cell_volume_correction,
    const MY_FLOAT_TYPE minimal_area) {
    (void)integral_method;
    const fia_advect_opts o = fia_shim::opts_of(
        true, use_MacCormack_scheme, use_clamping_in_MacCormack_scheme, num_gauss_quad_boundary, num_gauss_quad_bulk,
        enable_eliminate_zero_velocity_false_diffusion, "", num_gauss_quadrature_points_on_one_triangle,
        interpolation_method, use_integral, num_gauss_quadrature_point_for_integrate_density,
        enable_cell_volume_correction, minimal_area);
    fia_grid3d g = fia_shim::host_grid(all_grid);
    fia_shim::check(fia_advect_velocity_3d(fia_shim::context(), &g, -1, time_step_length, &o));
}

// reference include/calc_pressure_3d.h:9: projection with the reference's Poisson system, solved by
// multigrid-preconditioned CG to a relative residual of 1e-6 (the reference's amgcl runs with its
// own defaults)
inline void calc_pressure_gradient_term_3D(Grid_3D &all_grid) {
    fia_grid3d g = fia_shim::host_grid(all_grid);
    int iters = 0;
    double res = 0.0;
    fia_shim::check(fia_pressure_project(fia_shim::context(), &g, nullptr, all_grid.pressure.data(), 0, 1e-6, 100000,
                                         &iters, &res));
    double loss[3];
    fia_shim::check(fia_wall_momentum_loss(fia_shim::context(), all_grid.Grid_num_x, all_grid.Grid_num_y,
                                           all_grid.Grid_num_z, all_grid._cell_length, all_grid.pressure.data(), 0,
                                           loss));
    all_grid.total_loss_of_momentum_x += (MY_FLOAT_TYPE)loss[0];
    all_grid.total_loss_of_momentum_y += (MY_FLOAT_TYPE)loss[1];
    all_grid.total_loss_of_momentum_z += (MY_FLOAT_TYPE)loss[2];
}

// reference include/write_substance_density_data.h (src/write_substance_density_data.cpp:11-27):
// density_data_path/substance_density_<file_number>.dat, "nx ny nz" then one value per line in
// [x][y][z] order, iostream's default formatting
inline void write_substance_density_data(Grid_3D &all_grid, int file_number, const std::string density_data_path) {
    std::ofstream f(density_data_path + "/substance_density_" + std::to_string(file_number) + ".dat", std::ios::out);
    f << all_grid.Grid_num_x << " " << all_grid.Grid_num_y << " " << all_grid.Grid_num_z << std::endl;
    for (MY_FLOAT_TYPE v : all_grid.substance_density) f << v << "\n";
    f.flush();
}

// Totals written each frame to the reference's *.dat series (src/main.cpp:518-597): total mass and
// total momenta x, y, z (sums times the cell volume), computed on the GPU in FP64.
inline void calc_total_mass_and_momentum_3D(Grid_3D &all_grid, double &total_mass, double &momentum_x,
                                            double &momentum_y, double &momentum_z) {
    fia_grid3d g = fia_shim::host_grid(all_grid);
    double t[4];
    fia_shim::check(fia_grid_totals(fia_shim::context(), &g, t));
    total_mass = t[0]; momentum_x = t[1]; momentum_y = t[2]; momentum_z = t[3];
}

// reference include/calc_time_step_length_from_CFL_number.h:33-39
inline MY_FLOAT_TYPE calc_time_step_length_from_CFL_number_3D(const Grid_3D &all_grid, const MY_FLOAT_TYPE now_time,
                                                             const int i_movie_frame, bool &write_frame_to_movie_file,
                                                             const int CFL_number, const MY_FLOAT_TYPE movie_fps) {
    fia_grid3d g = fia_shim::host_grid(const_cast<Grid_3D &>(all_grid));
    double dt = 0.0;
    int wf = 0;
    fia_shim::check(fia_cfl_time_step(fia_shim::context(), &g, now_time, i_movie_frame, CFL_number, movie_fps, &dt, &wf));
    write_frame_to_movie_file = wf != 0;
    return (MY_FLOAT_TYPE)dt;
}

}  // namespace smoke_simulation

#endif  // FIA_SMOKE_SIMULATION_HPP
