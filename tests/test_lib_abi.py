"""The C-ABI library loads on the CPU and exports every symbol the header declares.

No compute calls here (there is no GPU in the CPU suite).
"""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER_DIR = os.path.join(ROOT, "include")


def declared_functions():
    names = []
    for fname in sorted(os.listdir(HEADER_DIR)):
        if not fname.endswith(".h"):
            continue
        text = open(os.path.join(HEADER_DIR, fname)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(vp_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def _lib_path():
    from vlasov_amd.build import LIB_PATH, build
    if not os.path.exists(LIB_PATH):
        build()
    return LIB_PATH


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("vp_advect_f64", "vp_step_2d2v_f64", "vp_step_1d1v_f64", "vp_poisson_2d_f64"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  (bind to torch's HIP runtime, as the package does)
    lib = ctypes.CDLL(_lib_path())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_error_strings_and_version():
    from vlasov_amd import _lib
    L = _lib.lib()
    assert L.vp_abi_version() >= 1
    assert L.vp_error_string(0) == b"ok"
    assert b"workspace" in L.vp_error_string(-3)


def test_argument_validation_without_gpu():
    """Rejected arguments return before any HIP call, so they work on the CPU."""
    from vlasov_amd import _lib
    L = _lib.lib()
    shape = _lib.i64x4((4, 4, 4, 4))
    vs = _lib.i64x4((0, 0, 0, 0))
    assert L.vp_advect_f64(None, None, shape, 0, None, vs, 0.1, 0.1, 1, None) == -1
    assert L.vp_advect_f64(8, 16, shape, 7, 24, vs, 0.1, 0.1, 1, None) == -1
    assert L.vp_advect_f64(8, 16, _lib.i64x4((4, 4, 0, 4)), 0, 24, vs, 0.1, 0.1, 1, None) == -2
    assert L.vp_advect_f64(8, 16, _lib.i64x4((1, 1, 1, 9000)), 3, 24, vs, 0.1, 0.1, 1, None) == -4
    assert L.vp_poisson_2d_f64(8, 16, 16, 8, 8, 8, None, None, None, None, 8, 1, None) == -3
    assert L.vp_poisson_2d_f64(8, 1024, 16, 8, 8, 8, None, None, None, None, 8, 1 << 30, None) == -2


def test_noalloc_step_rejects_oversized_problems_without_gpu():
    """The allocation-free step checks the size against its static scratch
    before any HIP call."""
    from vlasov_amd import _lib
    L = _lib.lib()
    rc = L.vp_step_1d1v_f64_noalloc(8, 16, 1024, 2048, 0.1, 8, 8, 8, 0.1, 0.1, 1, 8, 8, 8, None)
    assert rc == -5
    assert b"scratch" in L.vp_error_string(rc)
    assert L.vp_step_1d1v_f64_noalloc(8, 16, 0, 128, 0.1, 8, 8, 8, 0.1, 0.1, 1, 8, 8, 8, None) == -2


def test_stream_argument_checks_without_gpu():
    """vp_stream_f64 rejects bad kinds, odd sizes, misaligned pointers and
    unroll factors before launching anything."""
    from vlasov_amd import _lib
    L = _lib.lib()
    assert L.vp_stream_f64(3, 16, 32, 64, 1, 1, 1, None) == -1
    assert L.vp_stream_f64(0, 16, 32, 63, 1, 1, 1, None) == -1
    assert L.vp_stream_f64(0, 24, 32, 64, 1, 1, 1, None) == -1
    assert L.vp_stream_f64(0, 16, 32, 64, 3, 1, 1, None) == -1
    assert L.vp_stream_f64(0, 16, 32, 64, 1, 1, 0, None) == -1
    assert L.vp_stream_f64(0, None, 32, 64, 1, 1, 1, None) == -1
