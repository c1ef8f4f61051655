"""The C ABI library loads and exports every function include/fia.h declares (CPU only: no
compute call needs a GPU here), and the host module parses the reference's options."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "fia.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fia_\w+)\s*\(", text)))


def test_library_exports_every_declared_function(fia):
    names = declared_functions()
    assert len(names) >= 18
    L = fia.lib()
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", fia.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (fia_\w+)", out))
    assert set(names) <= exported
    assert set(fia.EXPORTS) == set(names)


def test_no_device_fails_cleanly(fia):
    if fia.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(fia.FiaError) as e:
        fia.Context(0)
    assert e.value.status == 6   # FIA_ERR_NO_DEVICE


def test_default_options_follow_reference_cli(fia):
    # reference src/main.cpp:48-83
    o = fia.AdvectOpts()
    fia.lib().fia_default_opts(ctypes.byref(o))
    assert o.use_flux_advection == 1 and o.use_maccormack == 0
    assert o.enable_zero_velocity_correction == 1 and o.use_integral == 1
    assert o.num_gauss_points_integrate_density == 1 and o.minimal_area == 1.0
    assert o.split_method == 0            # "y-CellFaceAligned" matches no split method
    assert o.enable_cell_volume_correction == 0


def test_option_strings_map_like_the_reference(fia):
    assert fia.make_opts(split_method="y-CellFaceAligned").split_method == 0
    assert fia.make_opts(split_method="xz-CellFaceAligned").split_method == 1
    assert fia.make_opts(split_method="x-AxisAligned_z-CellFaceAligned").split_method == 4
    assert fia.make_opts(interpolation_method="1Dy_linear").interpolation_method == 2
    with pytest.raises(fia.FiaError):
        fia.make_opts(interpolation_method="cubic")


def test_cpp_shim_header_compiles():
    shim = os.path.join(ROOT, "include", "fia_smoke_simulation.hpp")
    if not os.path.exists(shim):
        pytest.skip("no shim")
    src = '#include "fia_smoke_simulation.hpp"\nint main(){ smoke_simulation::Grid_3D g(4,4,4,0.25f); return (int)g.substance_density.size() - 64 + (int)g.velocity_in_voxel_face_x.size() - 80; }\n'
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), "-x", "c++", "-"],
                       input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_struct_layouts_match_c(fia, tmp_path):
    # every struct the ctypes mirror declares has the C header's size and field offsets
    import ctypes
    pairs = [("fia_grid3d", fia.GridC), ("fia_advect_opts", fia.AdvectOpts), ("fia_kernel_stats", fia.KernelStats),
             ("fia_learned_info", fia.LearnedInfo), ("fia_scene", fia.SceneC), ("fia_tuning", fia.TuningC)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fia.h"', 'int main(void) {']
    for cname, py in pairs:
        lines.append('printf("%%zu\\n", sizeof(%s));' % cname)
        for f, _ in py._fields_:
            lines.append('printf("%%zu\\n", offsetof(%s, %s));' % (cname, f))
    lines.append('return 0; }')
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = []
    for _, py in pairs:
        want.append(ctypes.sizeof(py))
        want += [getattr(py, f).offset for f, _ in py._fields_]
    assert got == want
