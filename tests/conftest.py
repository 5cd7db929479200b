"""Shared fixtures: the CPU oracle, the reference harness (when built), the
snapshot generator and the HIP library."""
import ctypes
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ahf_ref")
SNAPGEN_SO = os.path.join(ROOT, "ahf_amd", "lib", "libahf_snapgen.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _load_oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)
    lib = ctypes.CDLL(ORACLE_SO)
    lib.oracle_amr_dump.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    lib.oracle_unbind_dump.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    lib.oracle_construct_dump.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    lib.oracle_rem_outsideRvir.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.oracle_rem_outsideRvir.restype = ctypes.c_uint64
    lib.oracle_sort_halo.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_int32, ctypes.c_int]
    lib.oracle_sort_halo.restype = None
    lib.oracle_hilbert_c2i.argtypes = [ctypes.c_void_p, ctypes.c_uint]
    lib.oracle_hilbert_c2i.restype = ctypes.c_uint64
    lib.oracle_sfc_key.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint]
    lib.oracle_sfc_key.restype = ctypes.c_uint64
    lib.oracle_gather_dump.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p]
    lib.oracle_tsc_weights.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    lib.oracle_tsc_weights.restype = None
    lib.oracle_indexx.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]
    lib.oracle_indexx.restype = None
    lib.oracle_profiles_dump.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    lib.oracle_prof_nbins.argtypes = [ctypes.c_uint64]
    lib.oracle_prof_get_axes.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.oracle_prof_get_axes.restype = None
    lib.oracle_prof_cnfw.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.oracle_prof_cnfw.restype = ctypes.c_double
    lib.oracle_prof_cr1.argtypes = [ctypes.c_double]
    lib.oracle_prof_cr1.restype = ctypes.c_double
    lib.oracle_prof_find_max.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.oracle_prof_find_max.restype = None
    lib.oracle_gridinfo_dump.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double]
    return lib


@pytest.fixture(scope="session")
def oracle():
    return _load_oracle()


@pytest.fixture(scope="session")
def ref_bin():
    if not os.path.exists(REF_BIN):
        pytest.skip("reference harness oracle/_ref/ahf_ref not built (reference tree absent)")
    return REF_BIN


@pytest.fixture(scope="session")
def snapgen():
    if not os.path.exists(SNAPGEN_SO):
        from ahf_amd import build
        build.build_snapgen()
    lib = ctypes.CDLL(SNAPGEN_SO)
    lib.ahf_snapgen_write.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_char_p]
    return lib
