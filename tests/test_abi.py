"""The C-ABI library loads and exports every entry point include/qpslice.h
declares; calls report errors by status code (no GPU is needed for these)."""
import ctypes
import os

import pytest

qps = pytest.importorskip("qpslice")


def test_library_exports_every_declared_symbol():
    L = qps.lib()
    declared = qps.exported_symbols()
    assert len(declared) >= 60
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_null_arguments_are_errors():
    L = qps.lib()
    assert L.qp_ctx_sync(None) == -1
    assert L.qp_field2d_fftrk(None, 1) == -1
    assert L.qp_species_push(None, None, None, None, ctypes.c_double(1.0)) == -1
    assert L.qp_slice_run(None, 1, 2, ctypes.c_double(0.0), None, None) == -1
    assert b"null" in L.qp_last_error()


def test_bad_grid_is_rejected_before_touching_a_device():
    L = qps.lib()
    h = ctypes.c_void_p()
    box = (ctypes.c_double * 6)(-10, 10, -10, 10, 0, 10)
    assert L.qp_ctx_create(3, 3, 3, box, 0, 0, ctypes.byref(h)) == -6  # 8x8 unsupported
    assert L.qp_ctx_create(6, 6, 6, box, 7, 0, ctypes.byref(h)) == -1  # bad precision


def test_python_mirror_names_match_the_reference_methods():
    # field2d_class.f03:38-62, species2d_class.f03:37-51, beam3d_class.f03:35-43
    for m in ("as_", "fftrk", "fftkr", "div", "grad", "pot", "smooth", "elf", "bf", "bfqp",
              "add", "sub", "mult", "ag", "cg", "cb", "ca"):
        assert hasattr(qps.Field2d, m), m
    for m in ("qdp", "extpsi", "amjdp", "push", "cbq", "renew"):
        assert hasattr(qps.Species, m), m
    for m in ("qdp", "qdpcopy", "push"):
        assert hasattr(qps.Beam, m), m


def test_library_exports_the_network_entry_points():
    header = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "include", "qpnet.h")
    declared = qps.exported_symbols(header)
    assert len(declared) >= 13
    missing = [s for s in declared if not hasattr(qps.lib(), s)]
    assert not missing, missing
