// A caller written against the reference's entry points, compiled with the header-only drop-in
// include/fia_smoke_simulation.hpp and linked with libfia.so: the calls and argument lists of the
// reference's velocity update (src/update_fluid_velocity_3d.cpp:59-76) and of its main loop's
// density step (src/main.cpp:692-711), on host std::vectors of Grid_3D.
//
//   shim_caller <dir> <n> <cell_length> <dt>
// reads <dir>/{rho,u,v,w}.bin (float32, the reference's [x][y][z] layout), advects the velocity,
// then the density with the advected velocity, writes <dir>/{u1,v1,w1,rho1}.bin, and checks that a
// bad option throws (where the reference exits) instead of returning.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "fia_smoke_simulation.hpp"

using smoke_simulation::Grid_3D;
using smoke_simulation::MY_FLOAT_TYPE;

static bool read_bin(const std::string &path, std::vector<MY_FLOAT_TYPE> &v) {
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * sizeof(MY_FLOAT_TYPE)));
    return (bool)f && f.gcount() == (std::streamsize)(v.size() * sizeof(MY_FLOAT_TYPE));
}

static bool write_bin(const std::string &path, const std::vector<MY_FLOAT_TYPE> &v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(MY_FLOAT_TYPE)));
    return (bool)f;
}

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: shim_caller <dir> <n> <cell_length> <dt>\n");
        return 2;
    }
    const std::string dir = argv[1];
    const int n = std::atoi(argv[2]);
    const MY_FLOAT_TYPE h = (MY_FLOAT_TYPE)std::atof(argv[3]);
    const MY_FLOAT_TYPE time_step_length = (MY_FLOAT_TYPE)std::atof(argv[4]);
    Grid_3D all_grid(n, n, n, h);
    if (!read_bin(dir + "/rho.bin", all_grid.substance_density) || !read_bin(dir + "/u.bin", all_grid.velocity_in_voxel_face_x) ||
        !read_bin(dir + "/v.bin", all_grid.velocity_in_voxel_face_y) || !read_bin(dir + "/w.bin", all_grid.velocity_in_voxel_face_z)) {
        std::fprintf(stderr, "cannot read the inputs\n");
        return 2;
    }
    // the reference CLI's defaults (src/main.cpp:48-83) with the cell-centred split
    const bool use_flux_advection = true, use_MacCormack_scheme = false, use_clamping_in_MacCormack_scheme = true;
    const int num_gauss_quad_boundary = 1, num_gauss_quad_bulk = 1, num_gauss_quadrature_point_for_integrate_density = 1;
    const bool enable_eliminate_zero_velocity_false_diffusion = true, use_integral = true;
    const bool enable_cell_volume_correction = false;
    const std::string split_method = "xz-CellFaceAligned", integral_method = "gauss", interpolation_method = "WENO6";
    const MY_FLOAT_TYPE minimal_area = 1.0f;
    const int i_frame = 0;
    try {
        smoke_simulation::advect_velocity_xyz_flux_advection_3D_std_thread(
            all_grid, time_step_length, use_MacCormack_scheme, use_clamping_in_MacCormack_scheme,
            num_gauss_quad_boundary, num_gauss_quad_bulk, enable_eliminate_zero_velocity_false_diffusion,
            integral_method, /*num_gauss_quadrature_points_on_one_triangle = */ num_gauss_quad_bulk,
            interpolation_method, use_integral, num_gauss_quadrature_point_for_integrate_density,
            enable_cell_volume_correction, minimal_area);
        smoke_simulation::move_substances_3D(
            all_grid, i_frame, time_step_length, use_flux_advection, use_MacCormack_scheme,
            use_clamping_in_MacCormack_scheme, num_gauss_quad_boundary, num_gauss_quad_bulk,
            enable_eliminate_zero_velocity_false_diffusion, split_method, integral_method, num_gauss_quad_bulk,
            interpolation_method, use_integral, num_gauss_quadrature_point_for_integrate_density,
            enable_cell_volume_correction, minimal_area);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "step failed: %s\n", e.what());
        return 1;
    }
    if (!write_bin(dir + "/u1.bin", all_grid.velocity_in_voxel_face_x) || !write_bin(dir + "/v1.bin", all_grid.velocity_in_voxel_face_y) ||
        !write_bin(dir + "/w1.bin", all_grid.velocity_in_voxel_face_z) || !write_bin(dir + "/rho1.bin", all_grid.substance_density)) {
        std::fprintf(stderr, "cannot write the outputs\n");
        return 2;
    }
    // failures of the library reach the caller as exceptions: MacCormack is not implemented, and
    // an unknown interpolation string is rejected while the options are parsed
    int caught = 0;
    try {
        smoke_simulation::move_substances_3D(all_grid, i_frame, time_step_length, true, /*MacCormack*/ true, true, 1, 1,
                                             true, split_method, integral_method, 1, interpolation_method, true, 1,
                                             false, 1.0f);
    } catch (const std::runtime_error &e) {
        ++caught;
        std::printf("caught: %s\n", e.what());
    }
    try {
        smoke_simulation::move_substances_3D(all_grid, i_frame, time_step_length, true, false, true, 1, 1, true,
                                             split_method, integral_method, 1, "cubic", true, 1, false, 1.0f);
    } catch (const std::runtime_error &e) {
        ++caught;
        std::printf("caught: %s\n", e.what());
    }
    std::printf("shim caller done, %d failures reported as exceptions\n", caught);
    return caught == 2 ? 0 : 3;
}
