"""CPU oracle for SPIDER's `BP 3F` / `BP 32F` Fourier-gridding backprojection.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
checker (or the timed CPU baseline).  The product (``spider_amd`` and
``libspbp.so``) never imports it and has no CPU fallback.

Two libraries back it:

* ``liboracle.so`` (``spbp_oracle.cpp``): a restatement of the reference's
  algorithm with an FP32-faithful instantiation (reference loop order and
  rounding) and an FP64 instantiation (the "FP64 oracle").
* ``_ref/libspider_ref.so``: the reference's own hot-path Fortran routines
  (``ONELINE``, ``CANG``, ``ASTA_D``, ``FILLBESSIL``, ``BESI1``, ``RIBSL``,
  ``RJBSL``) compiled from ``src/*.f`` of the reference tree by
  ``oracle/Makefile``; present only where that tree was available at build
  time.  Used to pin the restatement.

Function-by-function provenance (reference file:line) is in the C++ source.
Every function here is pinned by a test in ``tests/test_oracle_*.py``; none is
"parity unpinned".
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and the reference checker when the tree exists)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "all"])
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build(ref=False)
        L = C.CDLL(path)
        vp = C.c_void_p
        L.or_bessel_i1.restype = C.c_double
        L.or_bessel_i1.argtypes = [C.c_double]
        L.or_bessel_i15.restype = C.c_double
        L.or_bessel_i15.argtypes = [C.c_double]
        L.or_bessel_j15.restype = C.c_double
        L.or_bessel_j15.argtypes = [C.c_double]
        L.or_cang_ref.argtypes = [C.c_float, C.c_float, C.c_float, _f32p]
        L.or_cang_f64.argtypes = [C.c_double, C.c_double, C.c_double, _f64p, C.c_int]
        L.or_kb_table_f32.argtypes = [C.c_int, _f32p, C.POINTER(C.c_float)]
        L.or_kb_table_f64.argtypes = [C.c_int, _f64p, C.POINTER(C.c_double)]
        L.or_asta_mean.restype = C.c_double
        L.or_asta_mean.argtypes = [_f32p, C.c_int]
        L.or_pad_f32.argtypes = [_f32p, C.c_int, C.c_int, _f32p]
        L.or_pad_f64.argtypes = [_f32p, C.c_int, C.c_int, _f64p]
        L.or_fft1.argtypes = [_f64p, C.c_int, C.c_int]
        L.or_fft2_r2c.argtypes = [_f64p, C.c_int, _f64p]
        L.or_fft3_c2r.argtypes = [_f64p, C.c_int, _f64p]
        L.or_spectrum_f32.argtypes = [_f32p, C.c_int, _f32p]
        L.or_spectrum_f64.argtypes = [_f32p, C.c_int, _f64p]
        L.or_oneline_f32.argtypes = [C.c_int, C.c_int, _f32p, _f32p, _f32p, _f32p, _f32p, C.c_float]
        L.or_insert_f32.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p, _f32p, C.c_float]
        L.or_insert_f64.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_double]
        L.or_insert_f32_acc64.argtypes = [C.c_int, _f64p, _f64p, _f32p, _f32p, _f32p, C.c_float]
        L.or_gather.argtypes = [C.c_int, _f64p, C.c_int, _f32p, _f32p, C.c_float, _f64p]
        L.or_bp_sym_f32.argtypes = [C.c_int, C.c_int, _f32p, _f64p, vp, vp, vp, vp, vp, C.c_int, _f64p]
        L.or_bp_sym_f32.restype = C.c_int
        L.or_sym_matmul_f32.argtypes = [_f32p, _f32p, _f32p]
        L.or_triquad.argtypes = [C.c_double, C.c_double, C.c_double, _f32p]
        L.or_triquad.restype = C.c_float
        L.or_symvol.argtypes = [C.c_int, _f32p, _f32p, _f32p, C.c_int]
        L.or_quadri.argtypes = [C.c_float, C.c_float, C.c_int, C.c_int, _f32p]
        L.or_quadri.restype = C.c_float
        L.or_rtsq.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]
        L.or_fbs2.argtypes = [C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, _f32p, C.c_int, _f32p, _f32p, _f32p,
                              C.c_int]
        L.or_fbs2.restype = C.c_float
        L.or_fbs2_prep.argtypes = [_f32p, C.c_int, C.c_int, _f32p, _f32p, _f32p]
        L.or_rtsf.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float]
        L.or_scatter.argtypes = [C.c_int, _f64p, C.c_int, _f32p, _f32p, C.c_float, _f64p, _f64p]
        L.or_symplane0_f32.argtypes = [C.c_int, _f32p, _f32p]
        L.or_symplane0_f64.argtypes = [C.c_int, _f64p, _f64p]
        L.or_normalize_f32.argtypes = [C.c_int, _f32p, _f32p]
        L.or_normalize_f64.argtypes = [C.c_int, _f64p, _f64p]
        L.or_inverse3_f32.argtypes = [C.c_int, _f32p, _f32p]
        L.or_inverse3_f64.argtypes = [C.c_int, _f64p, _f64p]
        L.or_window_f32.argtypes = [C.c_int, _f32p, _f32p]
        L.or_window_f64.argtypes = [C.c_int, _f64p, _f64p]
        L.or_finalize_f32.argtypes = [C.c_int, _f32p, _f32p, _f32p]
        L.or_finalize_f64.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.or_bp_f32.restype = C.c_int
        L.or_bp_f32.argtypes = [C.c_int, C.c_int, _f32p, _f64p, vp, vp, vp, vp, vp]
        L.or_bp_f64.restype = C.c_int
        L.or_bp_f64.argtypes = [C.c_int, C.c_int, _f32p, _f64p, vp, vp, vp, vp, vp, C.c_int]
        L.or_count_distinct.restype = C.c_long
        L.or_count_distinct.argtypes = [C.c_int, _f64p]
        L.or_count_updates.restype = C.c_long
        L.or_count_updates.argtypes = [C.c_int, _f64p]
        L.or_insert_omp_f32.restype = C.c_int
        L.or_insert_omp_f32.argtypes = [C.c_int, C.c_int, _f32p, _f32p, _f32p, _f32p, C.c_int]
        L.or_max_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _dt(prec):
    if prec not in ("f32", "f64"):
        raise ValueError("prec must be 'f32' or 'f64'")
    return np.float32 if prec == "f32" else np.float64


# ------------------------------------------------------------------ small pieces
def bessel_i1(x: float) -> float:
    return lib().or_bessel_i1(float(x))


def bessel_i15(x: float) -> float:
    return lib().or_bessel_i15(float(x))


def cang(phi, theta, psi, prec="f32", exact_pi=True) -> np.ndarray:
    """CANG(PHI, THETA, PSI) rotation, DM(1..9) as a flat array (cang.f:44-82).
    f32: the reference bit for bit (float32(pi)/180, REAL result).  f64: FP64,
    with exact pi or (exact_pi=False) the reference's degree conversion."""
    if prec == "f32":
        dm = np.zeros(9, np.float32)
        lib().or_cang_ref(float(phi), float(theta), float(psi), dm)
    else:
        dm = np.zeros(9, np.float64)
        lib().or_cang_f64(float(phi), float(theta), float(psi), dm, int(bool(exact_pi)))
    return dm


def rot_from_doc(angles, prec="f32", exact_pi=True) -> np.ndarray:
    """Doc-file registers (PSI, THETA, PHI) per image -> DM rows (bp3f.f:1290)."""
    a = np.asarray(angles, np.float64).reshape(-1, 3)
    return np.stack([cang(r[2], r[1], r[0], prec, exact_pi) for r in a])


def kb_table(n2: int, prec="f32"):
    dt = _dt(prec)
    tab = np.zeros(5000, dt)
    if prec == "f32":
        f = C.c_float()
        lib().or_kb_table_f32(n2, tab, C.byref(f))
    else:
        f = C.c_double()
        lib().or_kb_table_f64(n2, tab, C.byref(f))
    return tab, dt(f.value)


def asta_mean(img) -> float:
    img = np.ascontiguousarray(img, np.float32)
    return lib().or_asta_mean(img, img.shape[0])


def pad(img, n: int, prec="f32") -> np.ndarray:
    img = np.ascontiguousarray(img, np.float32)
    out = np.zeros((n, n), _dt(prec))
    getattr(lib(), "or_pad_" + prec)(img, img.shape[0], n, out)
    return out


def fft1(x, sign=-1) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, np.complex128)).view(np.float64).copy()
    lib().or_fft1(a, a.size // 2, sign)
    return a.view(np.complex128)


def fft2_r2c(x) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    n = x.shape[0]
    out = np.zeros((n, n // 2 + 1), np.complex128)
    lib().or_fft2_r2c(x, n, out.view(np.float64))
    return out


def fft3_c2r(X) -> np.ndarray:
    X = np.ascontiguousarray(X, np.complex128)
    n = X.shape[0]
    out = np.zeros((n, n, n), np.float64)
    lib().or_fft3_c2r(X.view(np.float64), n, out)
    return out


def spectrum(img, prec="f32") -> np.ndarray:
    """PADD2 + forward FFT + SPIDER sign + checkerboard: BI(0:NX, N) as (N, NX+1)."""
    img = np.ascontiguousarray(img, np.float32)
    nx = img.shape[0]
    ct = np.complex64 if prec == "f32" else np.complex128
    bi = np.zeros((2 * nx, nx + 1), ct)
    getattr(lib(), "or_spectrum_" + prec)(img, nx, bi.view(_dt(prec)))
    return bi


def empty_grid(nx: int, prec="f32"):
    """Half-grid X(0:NX, N, N) complex and W real, stored [kz][ky][kx]."""
    n = 2 * nx
    ct = np.complex64 if prec == "f32" else np.complex128
    return np.zeros((n, n, nx + 1), ct), np.zeros((n, n, nx + 1), _dt(prec))


def insert(X, W, bi, dm, tab, fltb, prec="f32") -> None:
    nx = X.shape[2] - 1
    dt = _dt(prec)
    getattr(lib(), "or_insert_" + prec)(nx, X.view(dt), W, bi.view(dt),
                                        np.ascontiguousarray(dm, dt), tab, dt(fltb))


def insert_f32_acc64(X, W, bi, dm, tab, fltb) -> None:
    """Reference geometry and fp32 per-tap terms accumulated in FP64 (X complex128, W float64)."""
    nx = X.shape[2] - 1
    lib().or_insert_f32_acc64(nx, X.view(np.float64), W, bi.view(np.float32),
                              np.ascontiguousarray(dm, np.float32), tab, np.float32(fltb))


def lib_table(n2: int, geom_f64: bool = False):
    """The table libspbp's kernels weight with: TABI in REAL arithmetic as the
    reference evaluates it (== kb_table(n2, "f32")), or with FP64 geometry the
    FP64 evaluation rounded to fp32; FLTB = 4999/3 in fp32.  For checks of
    functions that have no reference counterpart (the projector)."""
    f = np.float32(np.float32(4999.0) / np.float32(3.0))
    if geom_f64:
        return kb_table(n2, "f64")[0].astype(np.float32), f
    return kb_table(n2, "f32")[0], f


def gather(X, dm, table=None):
    """Central-slice KB gather (transpose of ONELINE) of the FP64 half grid
    X (complex128, N x N x (NX+1)) for rotations dm (nproj x 9 fp32):
    nproj x N x (NX+1) complex128 spectra in the BI layout.  table = (tab,
    fltb), default lib_table(N)."""
    X = np.ascontiguousarray(X, np.complex128)
    nx = X.shape[2] - 1
    N = 2 * nx
    tab, fltb = table if table is not None else lib_table(N)
    dm = np.ascontiguousarray(np.reshape(dm, (-1, 9)), np.float32)
    out = np.zeros((dm.shape[0], N, nx + 1), np.complex128)
    lib().or_gather(nx, X.view(np.float64), dm.shape[0], dm, np.ascontiguousarray(tab, np.float32), float(fltb),
                    out.view(np.float64))
    return out


def scatter(bi, dm, nx, table=None):
    """ONELINE scatter of FP64 BI spectra (nproj x N x (NX+1) complex128) with
    fp32 weights, FP64 terms: returns (X, W).  table as for gather()."""
    bi = np.ascontiguousarray(bi, np.complex128)
    dm = np.ascontiguousarray(np.reshape(dm, (-1, 9)), np.float32)
    N = 2 * nx
    tab, fltb = table if table is not None else lib_table(N)
    X = np.zeros((N, N, nx + 1), np.complex128)
    W = np.zeros((N, N, nx + 1), np.float64)
    lib().or_scatter(nx, bi.view(np.float64), dm.shape[0], dm, np.ascontiguousarray(tab, np.float32), float(fltb),
                     X.view(np.float64), W)
    return X, W


def oneline_f32(J, X, W, bi, dm, tab, fltb) -> None:
    nx = X.shape[2] - 1
    lib().or_oneline_f32(J, nx, X.view(np.float32), W, bi.view(np.float32),
                         np.ascontiguousarray(dm, np.float32), tab, np.float32(fltb))


def symplane0(X, W, prec="f32") -> None:
    dt = _dt(prec)
    getattr(lib(), "or_symplane0_" + prec)(X.shape[2] - 1, X.view(dt), W)


def normalize(X, W, prec="f32") -> None:
    dt = _dt(prec)
    getattr(lib(), "or_normalize_" + prec)(X.shape[2] - 1, X.view(dt), W)


def inverse3(X, prec="f32") -> np.ndarray:
    dt = _dt(prec)
    nx = X.shape[2] - 1
    n = 2 * nx
    out = np.zeros((n, n, n), dt)
    getattr(lib(), "or_inverse3_" + prec)(nx, np.ascontiguousarray(X).view(dt), out)
    return out


def window(real, prec="f32") -> np.ndarray:
    dt = _dt(prec)
    n = real.shape[0]
    nx = n // 2
    out = np.zeros((nx, nx, nx), dt)
    getattr(lib(), "or_window_" + prec)(nx, np.ascontiguousarray(real, dt), out)
    return out


def finalize(X, W, prec="f32") -> np.ndarray:
    """NRMW2 + WINDKB2A on a copy of (X, W): NX^3 volume."""
    dt = _dt(prec)
    nx = X.shape[2] - 1
    Xc = np.array(X, copy=True)
    out = np.zeros((nx, nx, nx), dt)
    getattr(lib(), "or_finalize_" + prec)(nx, Xc.view(dt), np.ascontiguousarray(W, dt), out)
    return out


def sym_matrices(sym_angles):
    """BUILDS (builds.f:35-63): fp32 CANG matrices of the symmetry doc rows
    (PSI, THETA, PHI), (nsym, 9)."""
    a = np.asarray(sym_angles, np.float64).reshape(-1, 3)
    return np.stack([cang(r[2], r[1], r[0], prec="f32") for r in a]).astype(np.float32).reshape(-1, 9)


def sym_matmul(sm, dm):
    """DMS = MATMUL(SM, DM) in fp32 (Fortran (3,3) storage, 9 floats each)."""
    out = np.zeros(9, np.float32)
    lib().or_sym_matmul_f32(np.ascontiguousarray(sm, np.float32).ravel(),
                            np.ascontiguousarray(dm, np.float32).ravel(), out)
    return out


def triquad(r, s, t, f):
    return np.float32(lib().or_triquad(float(r), float(s), float(t), np.ascontiguousarray(f, np.float32).ravel()))


def symvol(q1, sm):
    """SYMVOL of an NX^3 fp32 volume ([z][y][x]) with matrices sm (nsym, 9)."""
    q1 = np.ascontiguousarray(q1, np.float32)
    sm = np.ascontiguousarray(sm, np.float32).reshape(-1, 9)
    q2 = np.zeros_like(q1)
    lib().or_symvol(q1.shape[0], q1, q2, sm.ravel(), sm.shape[0])
    return q2


def quadri(x, y, img):
    img = np.ascontiguousarray(img, np.float32)
    return np.float32(lib().or_quadri(float(x), float(y), img.shape[1], img.shape[0], img))


def rtsq(img, theta, shx, shy, scale=1.0):
    """RTSQ of one image ([y][x] fp32): rotate, scale and shift with QUADRI."""
    img = np.ascontiguousarray(img, np.float32)
    out = np.zeros_like(img)
    lib().or_rtsq(img, out, img.shape[1], img.shape[0], float(theta), float(scale), float(shx), float(shy))
    return out


def fbs2(x, y, img, x1, y1, xy2, chk=True) -> np.float32:
    """FBS2 (fbs2.f:95-160) at 1-based (x, y) of img ([y][x] fp32) with its
    derivative images x1 = dF/dx, y1 = dF/dy, xy2 = d2F/dxdy."""
    a = [np.ascontiguousarray(v, np.float32) for v in (img, x1, y1, xy2)]
    ny, nx = a[0].shape
    return np.float32(lib().or_fbs2(float(x), float(y), nx, nx, ny, a[0], nx, a[1], a[2], a[3], int(bool(chk))))


def fbs2_prep(img):
    """FBS2_PREP (fbs2.f:180-245): (dF/dx, dF/dy, d2F/dxdy) of img ([y][x]) by
    the Fourier method, FP64 transforms rounded to fp32."""
    img = np.ascontiguousarray(img, np.float32)
    out = [np.zeros_like(img) for _ in range(3)]
    lib().or_fbs2_prep(img, img.shape[1], img.shape[0], *out)
    return tuple(out)


def rtsf(img, theta, shx, shy):
    """RTSF (rtsf.f:106-205, scale 1): rotation by theta degrees and shift
    (shx, shy) of one image ([y][x] fp32) with FBS2 interpolation."""
    img = np.ascontiguousarray(img, np.float32)
    out = np.zeros_like(img)
    lib().or_rtsf(img, out, img.shape[1], img.shape[0], float(theta), float(shx), float(shy))
    return out


def bp_sym(imgs, angles, sym_angles, half=None, want_xw=False):
    """FP32-faithful BP 3F / BP 32F with a symmetry document (sym_angles:
    (nsym, 3) PSI, THETA, PHI rows): every projection inserted once per
    DMS = SM DM, and SYMVOL applied to every output volume ('X', 'W' as in
    bp() when want_xw)."""
    imgs = np.ascontiguousarray(imgs, np.float32)
    nimg, nx, _ = imgs.shape
    ang = np.ascontiguousarray(np.asarray(angles, np.float64).reshape(nimg, 3))
    sym = np.ascontiguousarray(np.asarray(sym_angles, np.float64).reshape(-1, 3))
    out = {"full": np.zeros((nx, nx, nx), np.float32)}
    hp = None
    if half is not None:
        hp = np.ascontiguousarray(np.asarray(half, np.int8).reshape(nimg))
        out["h1"] = np.zeros((nx, nx, nx), np.float32)
        out["h2"] = np.zeros((nx, nx, nx), np.float32)
    ptr = lambda a: None if a is None else a.ctypes.data
    n = 2 * nx
    xw = np.zeros(3 * (nx + 1) * n * n, np.float32) if want_xw else None
    rc = lib().or_bp_sym_f32(nx, nimg, imgs, ang, ptr(hp), ptr(out["full"]), ptr(out.get("h1")),
                             ptr(out.get("h2")), ptr(xw), sym.shape[0], sym)
    if rc != 0:
        raise RuntimeError(f"oracle bp failed: {rc}")
    if want_xw:
        nh = (nx + 1) * n * n
        out["X"] = xw[: 2 * nh].view(np.complex64).reshape(n, n, nx + 1)
        out["W"] = xw[2 * nh:].reshape(n, n, nx + 1)
    return out


def bp(imgs, angles, half=None, prec="f32", want_xw=False, exact_pi=False):
    """Whole operation: BP 3F (half is None) or BP 32F (half[i] = 0 -> first
    sample volume, 1 -> second).  angles: (nimg, 3) doc registers (PSI, THETA,
    PHI) in degrees.  Returns dict with 'full' (and 'h1', 'h2' for BP 32F;
    'X', 'W' = accumulated post-SYMPLANE0 half grid when want_xw).
    exact_pi (f64 only) replaces the reference's float32(pi)/180 degree
    conversion by the exact one."""
    imgs = np.ascontiguousarray(imgs, np.float32)
    nimg, nx, ny = imgs.shape
    if nx != ny:
        raise ValueError("ONLY WORKS ON SQUARE IMAGES")
    ang = np.ascontiguousarray(np.asarray(angles, np.float64).reshape(nimg, 3))
    dt = _dt(prec)
    out = {"full": np.zeros((nx, nx, nx), dt)}
    hp = None
    if half is not None:
        hp = np.ascontiguousarray(np.asarray(half, np.int8).reshape(nimg))
        out["h1"] = np.zeros((nx, nx, nx), dt)
        out["h2"] = np.zeros((nx, nx, nx), dt)
    n = 2 * nx
    xw = np.zeros(3 * (nx + 1) * n * n, dt) if want_xw else None
    ptr = lambda a: None if a is None else a.ctypes.data
    extra = () if prec == "f32" else (int(bool(exact_pi)),)
    rc = getattr(lib(), "or_bp_" + prec)(nx, nimg, imgs, ang, ptr(hp), ptr(out["full"]),
                                         ptr(out.get("h1")), ptr(out.get("h2")), ptr(xw), *extra)
    if rc != 0:
        raise RuntimeError(f"oracle bp failed: {rc}")
    if want_xw:
        nh = (nx + 1) * n * n
        ct = np.complex64 if prec == "f32" else np.complex128
        out["X"] = xw[: 2 * nh].view(ct).reshape(n, n, nx + 1)
        out["W"] = xw[2 * nh:].reshape(n, n, nx + 1)
    return out


def count_distinct(nx: int, dm) -> int:
    return lib().or_count_distinct(nx, np.ascontiguousarray(dm, np.float64))


def count_updates(nx: int, dm) -> int:
    return lib().or_count_updates(nx, np.ascontiguousarray(dm, np.float64))


def insert_omp_f32(imgs, dms, X, W, nthreads=0) -> None:
    """CPU baseline: reference production scheme (OpenMP stride-6 rows)."""
    imgs = np.ascontiguousarray(imgs, np.float32)
    nimg, nx, _ = imgs.shape
    lib().or_insert_omp_f32(nx, nimg, imgs, np.ascontiguousarray(dms, np.float32),
                            X.view(np.float32), W, int(nthreads))


def max_threads() -> int:
    return lib().or_max_threads()


# ------------------------------------------------------------------ reference build
class _Tabs(C.Structure):
    # COMMON /TABS/ LN2, FLTB, TABI(0:4999)  (oneline.f:44-45)
    _fields_ = [("ln2", C.c_int32), ("fltb", C.c_float), ("tabi", C.c_float * 5000)]


class RefLib:
    """The reference's own Fortran routines, compiled from its sources."""

    def __init__(self, path):
        self.L = C.CDLL(path)
        self.tabs = _Tabs.in_dll(self.L, "tabs_")
        for name in ("besi1_", "ribsl_", "rjbsl_"):
            getattr(self.L, name).restype = C.c_float

    # CANG(PHI,THETA,PSI,DOSS,SS,DM)
    def cang(self, phi, theta, psi) -> np.ndarray:
        dm = np.zeros(9, np.float32)
        ss = np.zeros(6, np.float32)
        self.L.cang_(C.byref(C.c_float(phi)), C.byref(C.c_float(theta)), C.byref(C.c_float(psi)),
                     C.byref(C.c_int32(0)), ss.ctypes.data_as(C.c_void_p), dm.ctypes.data_as(C.c_void_p))
        return dm

    # TRIQUAD(R,S,T,F): REAL function, DOUBLE PRECISION offsets
    def triquad(self, r, s, t, f) -> np.float32:
        self.L.triquad_.restype = C.c_float
        f = np.ascontiguousarray(f, np.float32).ravel()
        return np.float32(self.L.triquad_(C.byref(C.c_double(r)), C.byref(C.c_double(s)), C.byref(C.c_double(t)),
                                          f.ctypes.data_as(C.c_void_p)))

    # QUADRI(XX,YY,NX,NY,FDATA): REAL function
    def quadri(self, x, y, img) -> np.float32:
        self.L.quadri_.restype = C.c_float
        img = np.ascontiguousarray(img, np.float32)
        return np.float32(self.L.quadri_(C.byref(C.c_float(x)), C.byref(C.c_float(y)),
                                         C.byref(C.c_int32(img.shape[1])), C.byref(C.c_int32(img.shape[0])),
                                         img.ctypes.data_as(C.c_void_p)))

    # RTSQ(XIMG,BUFOUT,NX,NY,NXP,NYP,THETA,SCLI,SHXI,SHYI,IRTFLG)
    def rtsq(self, img, theta, shx, shy, scale=1.0) -> np.ndarray:
        img = np.ascontiguousarray(img, np.float32)
        ny, nx = img.shape
        out = np.zeros_like(img)
        irt = C.c_int32()
        i = lambda v: C.byref(C.c_int32(v))
        f = lambda v: C.byref(C.c_float(v))
        self.L.rtsq_(img.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), i(nx), i(ny), i(nx), i(ny),
                     f(theta), f(scale), f(shx), f(shy), C.byref(irt))
        return out

    # FBS2(X,Y,NXLD,NX,NY,FDATA,NXP,X1,Y1,XY2,CHKBOUND): REAL function
    def fbs2(self, x, y, img, x1, y1, xy2, chk=True) -> np.float32:
        self.L.fbs2_.restype = C.c_float
        a = [np.ascontiguousarray(v, np.float32) for v in (img, x1, y1, xy2)]
        ny, nx = a[0].shape
        i = lambda v: C.byref(C.c_int32(v))
        return np.float32(self.L.fbs2_(C.byref(C.c_float(x)), C.byref(C.c_float(y)), i(nx), i(nx), i(ny),
                                       a[0].ctypes.data_as(C.c_void_p), i(nx), a[1].ctypes.data_as(C.c_void_p),
                                       a[2].ctypes.data_as(C.c_void_p), a[3].ctypes.data_as(C.c_void_p),
                                       C.byref(C.c_int32(1 if chk else 0))))

    # SYMVOL(Q1,Q2,KLX,KNX,KLY,KNY,KLZ,KNZ,RM,NSYM)
    def symvol(self, q1, sm) -> np.ndarray:
        q1 = np.ascontiguousarray(q1, np.float32)
        nx = q1.shape[0]
        sm = np.ascontiguousarray(sm, np.float32).reshape(-1, 9)
        q2 = np.zeros_like(q1)
        klx = -(nx // 2)
        knx = nx // 2 - 1 if nx % 2 == 0 else nx // 2
        i = lambda v: C.byref(C.c_int32(v))
        self.L.symvol_(q1.ctypes.data_as(C.c_void_p), q2.ctypes.data_as(C.c_void_p), i(klx), i(knx), i(klx), i(knx),
                       i(klx), i(knx), sm.ctypes.data_as(C.c_void_p), i(sm.shape[0]))
        return q2

    # FILLBESSIL(N2,LN2,LTAB,FLTB,TABI,ALPHA,AAAA,NNN) -> into COMMON /TABS/
    def fillbessil(self, n2: int):
        ln2 = C.c_int32()
        fltb = C.c_float()
        alpha = C.c_float()
        aaaa = C.c_float()
        nnn = C.c_int32()
        tab = np.zeros(5000, np.float32)
        self.L.fillbessil_(C.byref(C.c_int32(n2)), C.byref(ln2), C.byref(C.c_int32(4999)), C.byref(fltb),
                           tab.ctypes.data_as(C.c_void_p), C.byref(alpha), C.byref(aaaa), C.byref(nnn))
        self.tabs.ln2 = ln2.value
        self.tabs.fltb = fltb.value
        C.memmove(self.tabs.tabi, tab.ctypes.data, 5000 * 4)
        return tab, np.float32(fltb.value), ln2.value, alpha.value, aaaa.value, nnn.value

    # ONELINE(J,N,N2,X,W,BI,DM) with N = padded size, N2 = NX
    def oneline(self, J: int, X, W, bi, dm) -> None:
        nx = X.shape[2] - 1
        assert X.dtype == np.complex64 and W.dtype == np.float32 and bi.dtype == np.complex64
        assert X.flags.c_contiguous and W.flags.c_contiguous and bi.flags.c_contiguous
        dm = np.ascontiguousarray(dm, np.float32)
        self.L.oneline_(C.byref(C.c_int32(J)), C.byref(C.c_int32(2 * nx)), C.byref(C.c_int32(nx)),
                        X.ctypes.data_as(C.c_void_p), W.ctypes.data_as(C.c_void_p),
                        bi.ctypes.data_as(C.c_void_p), dm.ctypes.data_as(C.c_void_p))

    # ASTA_D(X,N,RI,ABA,DKLP)
    def asta_d(self, img, ri: float):
        img = np.ascontiguousarray(img, np.float32)
        aba = C.c_double(0.0)
        dklp = C.c_double(0.0)
        self.L.asta_d_(img.ctypes.data_as(C.c_void_p), C.byref(C.c_int32(img.shape[0])),
                       C.byref(C.c_float(ri)), C.byref(aba), C.byref(dklp))
        return aba.value, dklp.value

    # PR3DB(A,B,PR,AMP,CSUM1,LR,CSUM,CSUM2,AVSUM,LSD,NX,NY,NZ,DSCALE,NSCALE,
    #       SCALE1,SSANG,INC,Y1,WI,SER): shell sums of the FSC (pr3db.f:34)
    def pr3db(self, A, B, nx, wi=0.5, ser=b"C", ssang=90.0):
        """A, B: (NZ, NY, LSD) float32 transforms in the SPIDER LSD layout."""
        nz, ny, lsd = A.shape
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        inc = int(float(max(nx, ny, nz)) / wi) // 2 + 1
        nscale = 20
        pr = np.zeros((inc, nscale), np.float32)
        amp = np.zeros((inc, nscale), np.float32)
        avsum = np.zeros((inc, nscale), np.float32)
        csum1 = np.zeros(inc, np.float32)
        csum = np.zeros(inc, np.float32)
        csum2 = np.zeros(inc, np.float32)
        lr = np.zeros(inc, np.int32)
        i32 = lambda v: C.byref(C.c_int32(v))
        f32 = lambda v: C.byref(C.c_float(v))
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        scale1, scale2 = 0.2, 2.0
        self.L.pr3db_(vp(A), vp(B), vp(pr), vp(amp), vp(csum1), vp(lr), vp(csum), vp(csum2), vp(avsum),
                      i32(lsd), i32(nx), i32(ny), i32(nz), f32((scale2 - scale1) / (nscale - 1)), i32(nscale),
                      f32(scale1), f32(ssang), i32(inc), f32(float(max(nx, ny, nz))), f32(wi),
                      C.c_char_p(ser), C.c_size_t(1))
        return {"csum": csum, "csum1": csum1, "csum2": csum2, "lr": lr, "inc": inc}

    def besi1(self, x: float) -> float:
        return self.L.besi1_(C.byref(C.c_float(x)))

    def ribsl(self, x: float, nu: float) -> float:
        return self.L.ribsl_(C.byref(C.c_float(x)), C.byref(C.c_float(nu)))

    def rjbsl(self, x: float, nu: float) -> float:
        return self.L.rjbsl_(C.byref(C.c_float(x)), C.byref(C.c_float(nu)))


def ref():
    """The reference checker library, or None where it was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libspider_ref.so")
        if not os.path.exists(path):
            return None
        _REF = RefLib(path)
    return _REF
