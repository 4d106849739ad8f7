"""CPU checks of the C boundary: the library loads and exports every symbol
the header declares; the Python layer refuses to run without a GPU."""
import ctypes as C
import os

import pytest

from spider_amd import _lib


def test_header_declares_the_boundary():
    syms = _lib.declared_symbols()
    for s in ("spbp_create", "spbp_destroy", "spbp_insert", "spbp_insert_dm", "spbp_insert_host",
              "spbp_finalize", "spbp_finalize_host", "spbp_export_partial", "spbp_import_partial",
              "spbp_last_error", "spbp_reset"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libspbp.so not built (run __graft_entry__.build())")
    L = _lib.lib()
    missing = [s for s in _lib.declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert L.spbp_version().startswith(b"spbp")


def test_create_rejects_bad_sizes_without_touching_a_device():
    L = _lib.lib()
    ctx = C.c_void_p()
    for nx in (0, 3, 2000):
        rc = L.spbp_create(C.byref(_lib.Opts(nx, 0, -1, 0, 0, 0)), C.byref(ctx))
        assert rc == _lib.ERR_SIZE and not ctx.value
    assert b"outside" in L.spbp_last_error(None)
    assert L.spbp_create(None, C.byref(ctx)) == _lib.ERR_ARG
    # NULL context everywhere is an argument error, never a crash
    assert L.spbp_insert(None, None, 0, 1, None, None, None) == _lib.ERR_ARG
    assert L.spbp_finalize(None, None, None, None, None) == _lib.ERR_ARG
    assert L.spbp_export_partial(None, 0, None, None, None) == _lib.ERR_ARG
    assert L.spbp_kernel_stats(None, 0, None, None, None) == _lib.ERR_ARG
    r = C.c_double()
    assert L.spbp_calibrate(7, 0, C.byref(r), None) == _lib.ERR_ARG
    assert L.spbp_calibrate(_lib.CAL_LDS_U64, 0, None, None) == _lib.ERR_ARG
    assert L.spbp_scatter_images(None, None, 1, None, None, None, None) == _lib.ERR_ARG
    assert L.spbp_gather_images(None, None, 1, None, None, None, None) == _lib.ERR_ARG
    assert L.spbp_destroy(None) == 0


def test_python_layer_needs_a_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from spider_amd.bp import Backprojector
    with pytest.raises(RuntimeError):
        Backprojector(16)


def test_random_halves_is_a_balanced_split():
    from spider_amd.bp import random_halves
    for n in (1, 2, 7, 12, 1001):
        h = random_halves(n, seed=n)
        assert h.dtype.name == "int8" and set(h.tolist()) <= {0, 1}
        assert (h == 0).sum() == n // 2
