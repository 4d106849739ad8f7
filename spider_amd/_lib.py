"""ctypes binding of libspbp.so (include/spbp.h).

There is no CPU fallback: if the library is missing or no GPU is present the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspbp.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "spbp.h")

OK = 0
ERR_ALLOC = 46
ERR_SIZE = 101
ERR_ANGLE = 102
ERR_FFT = 201
ERR_DEVICE = 202
ERR_ARG = 203
ERR_NONFINITE = 205

F32 = 0
BF16 = 1

TWO_HALVES = 0x01
EXACT_PI = 0x02
TWOSUM_W = 0x04
CHECK_FINITE = 0x08
TIMING = 0x10
GEOM_F64 = 0x20
PLAIN_THRESHOLD = 0x40
HIPFFT = 0x80
FBS = 0x100

K_INSERT = 0
K_PAD = 1
K_FFT2 = 2
K_FINALIZE = 3
K_BIN = 4
K_GATHER = 5
K_RESOLVE = 6

CAL_LDS_U64 = 0
CAL_GLOBAL_F32 = 1

STRAT_AUTO = 0
STRAT_DIRECT = 1
STRAT_BRICK = 2
STRAT_GATHER = 3


class SpbpError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"spbp error {code}: {msg}")
        self.code = code


class Job(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("imgs", C.c_void_p), ("dtype", C.c_int), ("nimg", C.c_int),
                ("angles_deg", C.c_void_p), ("dm", C.c_void_p), ("half", C.c_void_p)]


class FinalJob(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("out_full", C.c_void_p), ("out_h1", C.c_void_p), ("out_h2", C.c_void_p)]


SHAPES = [(16, 16, 8), (16, 8, 8), (8, 8, 8), (16, 12, 8), (16, 16, 10), (16, 16, 12), (16, 16, 13), (16, 16, 13)]
# owner-computes bricks (SHAPES gives the LDS region; each sample is evaluated in one brick only)
SHAPE_OWN = [False] * 7 + [True]
TUNE_C = [16, 32, 64, 128]


class Opts(C.Structure):
    _fields_ = [("nx", C.c_int), ("flags", C.c_int), ("device", C.c_int), ("batch", C.c_int),
                ("strategy", C.c_int), ("brick_proj", C.c_int)]


_LIB = None
vp = C.c_void_p


def lib():
    """Load libspbp.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        L.spbp_version.restype = C.c_char_p
        L.spbp_last_error.restype = C.c_char_p
        L.spbp_last_error.argtypes = [vp]
        L.spbp_grid_elems.restype = C.c_size_t
        L.spbp_grid_elems.argtypes = [vp]
        L.spbp_create.argtypes = [C.POINTER(Opts), C.POINTER(vp)]
        L.spbp_destroy.argtypes = [vp]
        L.spbp_reset.argtypes = [vp, vp]
        L.spbp_insert.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp, vp]
        L.spbp_insert_dm.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp, vp]
        L.spbp_insert_host.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp, vp]
        L.spbp_finalize.argtypes = [vp, vp, vp, vp, vp]
        L.spbp_finalize_host.argtypes = [vp, vp, vp, vp, vp]
        L.spbp_finalize_batch.argtypes = [vp, C.c_int, vp]
        L.spbp_export_partial.argtypes = [vp, C.c_int, vp, vp, vp]
        L.spbp_import_partial.argtypes = [vp, C.c_int, vp, vp, C.c_int, vp]
        L.spbp_export_raw.argtypes = [vp, C.c_int, vp, vp, vp]
        L.spbp_kernel_stats.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.spbp_reset_stats.argtypes = [vp]
        L.spbp_gather.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp]
        L.spbp_scatter.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp, vp]
        L.spbp_kb_table.argtypes = [C.c_int, C.c_int, vp, C.POINTER(C.c_float)]
        L.spbp_tune.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp, vp]
        L.spbp_set_symmetries.argtypes = [vp, C.c_int, vp]
        L.spbp_insert_rb.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp, vp, vp]
        L.spbp_rtsq.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp]
        L.spbp_insert_rb_host.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp, vp, vp]
        L.spbp_rtsq_host.argtypes = [vp, C.c_int, C.c_int, vp, vp]
        L.spbp_rtsf.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp]
        L.spbp_rtsf_host.argtypes = [vp, C.c_int, C.c_int, vp, vp]
        L.spbp_get_plan.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), vp]
        L.spbp_set_plan.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.spbp_insert_batch.argtypes = [vp, C.c_int, vp]
        L.spbp_graph_create.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp)]
        L.spbp_graph_launch.argtypes = [vp, vp]
        L.spbp_graph_destroy.argtypes = [vp]
        L.spbp_rotations.argtypes = [C.c_int, vp, C.c_int, vp]
        L.spbp_threshold_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.spbp_reference_weights.argtypes = [vp, C.c_int, vp, C.c_int, vp, vp]
        L.spbp_calibrate.argtypes = [C.c_int, C.c_size_t, C.POINTER(C.c_double), vp]
        L.spbp_scatter_images.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp]
        L.spbp_gather_images.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp]
        L.spbp_debug_fail_alloc.argtypes = [C.c_long]
        L.spbp_scatter_noalloc.argtypes = [C.c_int, C.c_int, vp, C.c_int, vp, vp, vp, vp, vp]
        for name in ("spbp_create", "spbp_destroy", "spbp_reset", "spbp_insert", "spbp_insert_dm",
                     "spbp_insert_host", "spbp_finalize", "spbp_finalize_host", "spbp_finalize_batch", "spbp_export_partial",
                     "spbp_import_partial", "spbp_export_raw", "spbp_kernel_stats", "spbp_reset_stats",
                     "spbp_gather", "spbp_scatter", "spbp_kb_table", "spbp_rotations", "spbp_tune",
                     "spbp_get_plan", "spbp_set_plan", "spbp_insert_batch", "spbp_graph_create",
                     "spbp_graph_launch", "spbp_graph_destroy", "spbp_set_symmetries", "spbp_insert_rb",
                     "spbp_rtsq", "spbp_rtsq_host", "spbp_rtsf", "spbp_rtsf_host", "spbp_insert_rb_host", "spbp_threshold_stats", "spbp_reference_weights",
                     "spbp_calibrate", "spbp_scatter_images", "spbp_gather_images", "spbp_debug_fail_alloc", "spbp_scatter_noalloc"):
            getattr(L, name).restype = C.c_int
        _LIB = L
    return _LIB


def declared_symbols() -> list[str]:
    """Every function the C header declares."""
    with open(HEADER) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spbp_[a-z0-9_]+)\s*\(", text)))


def check(rc: int, ctx=None) -> None:
    if rc != OK:
        msg = lib().spbp_last_error(ctx)
        raise SpbpError(rc, msg.decode() if msg else "")


def kb_table(n2: int, geom_f64: bool = False):
    """(TABI(0:4999) float32, FLTB) for padded size n2, from the library: the
    reference's REAL evaluation, or the FP64 one that FP64-geometry contexts use."""
    import numpy as np
    tab = np.zeros(5000, np.float32)
    f = C.c_float()
    check(lib().spbp_kb_table(int(n2), GEOM_F64 if geom_f64 else 0, tab.ctypes.data_as(C.c_void_p), C.byref(f)))
    return tab, np.float32(f.value)


def rotations(angles, exact_pi: bool = False):
    """CANG matrices (n, 9) float32 of doc registers (PSI, THETA, PHI), from the library."""
    import numpy as np
    a = np.ascontiguousarray(np.asarray(angles, np.float64).reshape(-1, 3))
    dm = np.zeros((a.shape[0], 9), np.float32)
    check(lib().spbp_rotations(a.shape[0], a.ctypes.data_as(C.c_void_p), EXACT_PI if exact_pi else 0,
                               dm.ctypes.data_as(C.c_void_p)))
    return dm
