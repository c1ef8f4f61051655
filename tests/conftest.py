"""Shared fixtures.  `gpu` tests need an MI355X; everything else runs on the CPU."""
import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

PKG_NAME = "flux-interpolated-advection_amd"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X) and the built libfia.so")


def load_pkg():
    return importlib.import_module(PKG_NAME)


def load_scenes():
    return importlib.import_module(PKG_NAME + ".scenes")


@pytest.fixture(scope="session")
def fia():
    return load_pkg()


@pytest.fixture(scope="session")
def scenes():
    return load_scenes()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx(fia):
    if fia.device_count() < 1:
        pytest.skip("no GPU visible")
    c = fia.Context(0)
    yield c
    c.close()


class HostKernels:
    """The library's __host__ __device__ cutter and interpolation code (csrc/fia_cutter.hpp,
    csrc/fia_math.hpp) built for the host from tests/cutter_host.cpp."""

    def __init__(self, path):
        import ctypes
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        L.host_general_cut.argtypes = [dp, ctypes.c_double, ctypes.c_int, dp, dp]
        L.host_general_cut.restype = ctypes.c_int
        L.host_split_coord.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_float)]
        L.host_split_coord.restype = ctypes.c_int
        L.host_fastdiv_mismatches.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_long]
        L.host_fastdiv_mismatches.restype = ctypes.c_long
        L.host_face_pieces.argtypes = [dp, dp, ctypes.c_int, dp, dp]
        L.host_face_pieces.restype = ctypes.c_int
        L.host_weno6.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.host_weno6.restype = ctypes.c_double
        L.host_weno6_opt.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.host_weno6_opt.restype = ctypes.c_double
        L.host_split_clamped.argtypes = [ctypes.c_double, ctypes.c_int, dp]
        L.host_split_clamped.restype = ctypes.c_int
        L.host_pipeline_pieces.argtypes = [dp, dp, dp, ctypes.c_int]
        L.host_pipeline_pieces.restype = ctypes.c_int
        L.host_pipeline_records.argtypes = [dp, dp, dp, ctypes.c_int]
        L.host_pipeline_records.restype = ctypes.c_int
        L.host_segment_records.argtypes = [dp, dp, dp, ctypes.c_int]
        L.host_segment_records.restype = ctypes.c_int
        L.host_face_class.argtypes = [dp]
        L.host_face_class.restype = ctypes.c_int
        self.lib = L
        self._dp = dp

    def pipeline_pieces(self, col, max_pieces=64):
        """Clipping-pipeline pieces of a face with corner colours col (4 x 2): (centroid, area)
        per (strip, band) slot, in xi-eta space."""
        col = np.ascontiguousarray(col, dtype=np.float64).reshape(8)
        cen = np.zeros(2 * max_pieces); area = np.zeros(max_pieces)
        k = self.lib.host_pipeline_pieces(col.ctypes.data_as(self._dp), cen.ctypes.data_as(self._dp),
                                          area.ctypes.data_as(self._dp), max_pieces)
        if k < 0:
            raise RuntimeError("pipeline: %d" % k)
        return cen[:2 * k].reshape(-1, 2), area[:k]

    def face_class(self, col):
        """(twisted, degenerate) as the face kernels decide them from the corner colours."""
        col = np.ascontiguousarray(col, dtype=np.float64).reshape(8)
        k = self.lib.host_face_class(col.ctypes.data_as(self._dp))
        return bool(k & 1), bool(k & 2)

    def weno6(self, f, t, kappa=1.0, dbl=False, opt=False):
        f = np.ascontiguousarray(f, dtype=np.float64)
        fn = self.lib.host_weno6_opt if opt else self.lib.host_weno6
        return fn(f.ctypes.data_as(self._dp), float(t), float(kappa), int(dbl))

    def split_clamped(self, s, m):
        """The "WENO6-optimized" split of sample coordinate s over m samples: (base, fraction)."""
        import ctypes
        t = ctypes.c_double(0.0)
        b = self.lib.host_split_clamped(float(s), int(m), ctypes.byref(t))
        return b, t.value


@pytest.fixture(scope="session")
def cutter_host(tmp_path_factory):
    import subprocess
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    out = str(tmp_path_factory.mktemp("cutter") / "cutter_host.so")
    cmd = [hipcc, "-x", "c++", "-O2", "-std=c++17", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", os.path.join(ROOT, "tests", "cutter_host.cpp"), "-o", out]
    subprocess.run(cmd, check=True, capture_output=True)
    return HostKernels(out)


def rel_linf(a, b):
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
