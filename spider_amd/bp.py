"""Operation-level API of the backprojection, mirroring SPIDER's `BP 3F` and
`BP 32F` (bp3f.f:79, bp32f.f:57): image series + angles in, volume(s) out.

Images are torch tensors resident on the GPU (n, NX, NX), float32 or
bfloat16, x fastest; angles are the doc-file registers (PSI, THETA, PHI) in
degrees per image.  Work is queued on torch's current HIP stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import SpbpError  # noqa: F401  (re-export)


def _torch():
    import torch
    return torch


def _stream(device=None):
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def random_halves(nang: int, seed=None) -> np.ndarray:
    """Half-set split as RANDLIST builds it (bp32f.f:544-553): the first
    NANG/2 flags true, then NANG random swaps.  Returns int8 0 (first sample
    volume, RANDLIST true) / 1 (second)."""
    rng = np.random.default_rng(seed)
    flags = np.zeros(nang, bool)
    flags[: nang // 2] = True
    for k in range(nang):
        iord = min(nang, max(1, int(rng.random() * nang + 0.5))) - 1
        flags[iord], flags[k] = flags[k], flags[iord]
    return np.where(flags, 0, 1).astype(np.int8)


class Backprojector:
    """One libspbp context: half-grid accumulators for an NX^3 reconstruction."""

    def __init__(self, nx: int, two_halves: bool = False, exact_pi: bool = False,
                 twosum: bool = False, check_finite: bool = False, device=None, batch: int = 0,
                 timing: bool = False, strategy: str = "auto", brick_proj: int = 0, symmetries=None,
                 geom_f64: bool = False, plain_threshold: bool = False, hipfft: bool = False, fbs: bool = False):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("spider_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else int(device))
        flags = (_lib.TWO_HALVES if two_halves else 0) | (_lib.EXACT_PI if exact_pi else 0) | \
                (_lib.TWOSUM_W if twosum else 0) | (_lib.CHECK_FINITE if check_finite else 0) | \
                (_lib.TIMING if timing else 0) | (_lib.GEOM_F64 if geom_f64 else 0) | \
                (_lib.PLAIN_THRESHOLD if plain_threshold else 0) | (_lib.HIPFFT if hipfft else 0) | \
                (_lib.FBS if fbs else 0)
        strat = {"auto": _lib.STRAT_AUTO, "direct": _lib.STRAT_DIRECT, "brick": _lib.STRAT_BRICK,
                 "gather": _lib.STRAT_GATHER}[strategy]
        opts = _lib.Opts(int(nx), flags, self.device.index, int(batch), strat, int(brick_proj))
        ctx = C.c_void_p()
        L = _lib.lib()
        _lib.check(L.spbp_create(C.byref(opts), C.byref(ctx)), None)
        self._ctx = ctx
        self.nx = int(nx)
        self.two_halves = bool(two_halves)
        if symmetries is not None:
            self.set_symmetries(symmetries)

    # -------------------------------------------------------------- lifecycle
    def close(self) -> None:
        """Release the context (refused, with SpbpError, while a BatchGraph
        that replays work on it exists)."""
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            self._check(_lib.lib().spbp_destroy(self._ctx))
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        _lib.check(rc, self._ctx)

    @property
    def grid_shape(self):
        n = 2 * self.nx
        return (n, n, self.nx + 1)

    def reset(self) -> None:
        self._check(_lib.lib().spbp_reset(self._ctx, _stream(self.device)))

    def threshold_stats(self) -> dict:
        """Last finalisation's NRMW2 re-summation: voxels re-summed in the
        reference's order, decisions that differ from the GPU sums' own, and
        re-sums that disagreed with the GPU weight (left to the GPU)."""
        v = [C.c_longlong() for _ in range(3)]
        self._check(_lib.lib().spbp_threshold_stats(self._ctx, *[C.byref(x) for x in v]))
        return {"listed": v[0].value, "changed": v[1].value, "inconsistent": v[2].value}

    def reference_weights(self, voxels, which: int = -1) -> np.ndarray:
        """W as NRMW2 reads it at half-grid voxels (flat indices into
        grid_shape, or (kz, ky, kx) index triples), summed in the reference's
        order: fp32 values (FP64 with geom_f64) as float64."""
        v = np.asarray(voxels, np.int64)
        if v.ndim == 2:
            n = 2 * self.nx
            v = (v[:, 0] * n + v[:, 1]) * (self.nx + 1) + v[:, 2]
        v = np.ascontiguousarray(v.ravel())
        out = np.zeros(v.size, np.float64)
        self._check(_lib.lib().spbp_reference_weights(self._ctx, int(which), v.ctypes.data_as(C.c_void_p), int(v.size),
                                                      out.ctypes.data_as(C.c_void_p), _stream(self.device)))
        return out

    def set_symmetries(self, angles) -> None:
        """Symmetry operations as doc rows (PSI, THETA, PHI), (nsym, 3); one
        row or None switches symmetrisation off (bp32f.f:133-172)."""
        if angles is None:
            self._check(_lib.lib().spbp_set_symmetries(self._ctx, 0, None))
            return
        a = np.ascontiguousarray(np.asarray(angles, np.float64).reshape(-1, 3))
        self._check(_lib.lib().spbp_set_symmetries(self._ctx, a.shape[0], a.ctypes.data_as(C.c_void_p)))

    # -------------------------------------------------------------- insertion
    def _images(self, imgs):
        torch = _torch()
        if not isinstance(imgs, torch.Tensor):
            raise TypeError("images must be a torch tensor on the GPU")
        if imgs.dim() == 2:
            imgs = imgs[None]
        if imgs.dim() != 3:
            raise ValueError("images must be (n, NX, NX)")
        if imgs.shape[1] != imgs.shape[2]:
            raise SpbpError(_lib.ERR_SIZE, "ONLY WORKS ON SQUARE IMAGES")
        if imgs.shape[1] != self.nx:
            raise SpbpError(_lib.ERR_SIZE, f"image size {imgs.shape[1]} != context NX {self.nx}")
        if imgs.device != self.device:
            raise ValueError(f"images on {imgs.device}, context on {self.device}")
        if imgs.dtype == torch.float32:
            dt = _lib.F32
        elif imgs.dtype == torch.bfloat16:
            dt = _lib.BF16
        else:
            raise TypeError(f"unsupported image dtype {imgs.dtype}")
        return imgs.contiguous(), dt

    def insert(self, imgs, angles, half=None, inplane=None) -> None:
        """Insert projections with doc-file angles (n, 3): PSI, THETA, PHI.
        inplane (n, 3) = (angle, x shift, y shift), doc registers 6-8: `RB 3F`
        pre-rotation of each image by RTSQ before the insertion."""
        imgs, dt = self._images(imgs)
        n = imgs.shape[0]
        ang = np.ascontiguousarray(np.asarray(angles, np.float64).reshape(-1, 3))
        if ang.shape[0] != n:
            raise SpbpError(_lib.ERR_ANGLE, f"MISSING ANGLE FOR IMAGE: {ang.shape[0]} angles for {n} images")
        hp = None
        if half is not None:
            hp = np.ascontiguousarray(np.asarray(half).astype(np.int8).reshape(n))
        hptr = None if hp is None else hp.ctypes.data_as(C.c_void_p)
        if inplane is None:
            self._check(_lib.lib().spbp_insert(self._ctx, C.c_void_p(imgs.data_ptr()), dt, n,
                                               ang.ctypes.data_as(C.c_void_p), hptr, _stream(self.device)))
            self._keep = (imgs, ang, hp)   # alive until the next call orders after this one
            return
        ip = np.ascontiguousarray(np.asarray(inplane, np.float32).reshape(-1, 3))
        if ip.shape[0] != n:
            raise SpbpError(_lib.ERR_ANGLE, f"MISSING ANGLE FOR IMAGE: {ip.shape[0]} in-plane entries for {n} images")
        self._check(_lib.lib().spbp_insert_rb(self._ctx, C.c_void_p(imgs.data_ptr()), dt, n,
                                              ang.ctypes.data_as(C.c_void_p), ip.ctypes.data_as(C.c_void_p),
                                              hptr, _stream(self.device)))
        self._keep = (imgs, ang, hp, ip)

    def insert_host(self, imgs, angles, half=None, inplane=None) -> None:
        """insert() for images in host memory: a numpy float32 array or a CPU
        torch tensor (float32 or bfloat16) of shape (n, NX, NX), through
        spbp_insert_host / spbp_insert_rb_host.  The library uploads them in
        sub-chunks that overlap the insertion; the call returns when they
        are inserted."""
        torch = _torch()
        if isinstance(imgs, np.ndarray):
            imgs = torch.from_numpy(np.ascontiguousarray(imgs, dtype=np.float32))
        if not isinstance(imgs, torch.Tensor) or imgs.device.type != "cpu":
            raise TypeError("host images must be a numpy array or a CPU torch tensor")
        if imgs.dim() == 2:
            imgs = imgs[None]
        if imgs.dim() != 3 or imgs.shape[1] != imgs.shape[2]:
            raise SpbpError(_lib.ERR_SIZE, "ONLY WORKS ON SQUARE IMAGES")
        if imgs.shape[1] != self.nx:
            raise SpbpError(_lib.ERR_SIZE, f"image size {imgs.shape[1]} != context NX {self.nx}")
        if imgs.dtype == torch.float32:
            dt = _lib.F32
        elif imgs.dtype == torch.bfloat16:
            dt = _lib.BF16
        else:
            raise TypeError(f"unsupported image dtype {imgs.dtype}")
        imgs = imgs.contiguous()
        n = imgs.shape[0]
        ang = np.ascontiguousarray(np.asarray(angles, np.float64).reshape(-1, 3))
        if ang.shape[0] != n:
            raise SpbpError(_lib.ERR_ANGLE, f"MISSING ANGLE FOR IMAGE: {ang.shape[0]} angles for {n} images")
        hp = None if half is None else np.ascontiguousarray(np.asarray(half).astype(np.int8).reshape(n))
        hptr = None if hp is None else hp.ctypes.data_as(C.c_void_p)
        if inplane is None:
            self._check(_lib.lib().spbp_insert_host(self._ctx, C.c_void_p(imgs.data_ptr()), dt, n,
                                                    ang.ctypes.data_as(C.c_void_p), hptr, _stream(self.device)))
            return
        ip = np.ascontiguousarray(np.asarray(inplane, np.float32).reshape(-1, 3))
        if ip.shape[0] != n:
            raise SpbpError(_lib.ERR_ANGLE, f"MISSING ANGLE FOR IMAGE: {ip.shape[0]} in-plane entries for {n} images")
        self._check(_lib.lib().spbp_insert_rb_host(self._ctx, C.c_void_p(imgs.data_ptr()), dt, n,
                                                   ang.ctypes.data_as(C.c_void_p), ip.ctypes.data_as(C.c_void_p),
                                                   hptr, _stream(self.device)))

    def insert_rotations(self, imgs, dm, half=None) -> None:
        """Insert projections with precomputed CANG matrices (n, 9) fp32."""
        imgs, dt = self._images(imgs)
        n = imgs.shape[0]
        dm = np.ascontiguousarray(np.asarray(dm, np.float32).reshape(-1, 9))
        if dm.shape[0] != n:
            raise SpbpError(_lib.ERR_ANGLE, "MISSING ANGLE FOR IMAGE")
        hp = None if half is None else np.ascontiguousarray(np.asarray(half).astype(np.int8).reshape(n))
        self._check(_lib.lib().spbp_insert_dm(self._ctx, C.c_void_p(imgs.data_ptr()), dt, n,
                                              dm.ctypes.data_as(C.c_void_p),
                                              None if hp is None else hp.ctypes.data_as(C.c_void_p),
                                              _stream(self.device)))
        self._keep = (imgs, dm, hp)

    # -------------------------------------------------------------- gather / scatter of spectra
    def _rotation_args(self, n, angles, dm):
        if dm is not None:
            dm = np.ascontiguousarray(np.asarray(dm, np.float32).reshape(-1, 9))
            if dm.shape[0] != n:
                raise SpbpError(_lib.ERR_ANGLE, f"{dm.shape[0]} rotations for {n} projections")
            return None, dm
        if angles is None:
            raise SpbpError(_lib.ERR_ANGLE, "MISSING ANGLE FOR IMAGE")
        ang = np.ascontiguousarray(np.asarray(angles, np.float64).reshape(-1, 3))
        if ang.shape[0] != n:
            raise SpbpError(_lib.ERR_ANGLE, f"MISSING ANGLE FOR IMAGE: {ang.shape[0]} angles for {n} projections")
        return ang, None

    def _grid_tensor(self, X, name):
        torch = _torch()
        if not isinstance(X, torch.Tensor) or X.dtype != torch.complex64 or tuple(X.shape) != self.grid_shape:
            raise ValueError(f"{name} must be a complex64 tensor of shape {self.grid_shape}")
        if X.device != self.device or not X.is_contiguous():
            raise ValueError(f"{name} must be contiguous on {self.device}")
        if X.is_conj() or X.is_neg():
            raise ValueError(f"{name} is a lazily conjugated / negated view; resolve it first")
        return X

    def gather(self, X, angles=None, dm=None):
        """Central-slice KB gather (the adjoint of the insertion) of the half
        grid X (N, N, NX+1) complex64 for n projections (angles (n, 3) or dm
        (n, 9)): (n, N, NX+1) complex64 spectra in the BI layout."""
        torch = _torch()
        if isinstance(X, torch.Tensor):
            X = X.resolve_conj().resolve_neg().contiguous()
        X = self._grid_tensor(X, "X")
        n = (np.asarray(dm).reshape(-1, 9) if dm is not None else np.asarray(angles).reshape(-1, 3)).shape[0]
        ang, dmc = self._rotation_args(n, angles, dm)
        N, _, H = self.grid_shape
        out = torch.empty((n, N, H), dtype=torch.complex64, device=self.device)
        self._check(_lib.lib().spbp_gather(self._ctx, C.c_void_p(X.data_ptr()), n,
                                           None if ang is None else ang.ctypes.data_as(C.c_void_p),
                                           None if dmc is None else dmc.ctypes.data_as(C.c_void_p),
                                           C.c_void_p(out.data_ptr()), _stream(self.device)))
        return out

    def gather_images(self, X, angles=None, dm=None, out=None):
        """The central-slice projector applied to the half grid X: the gather,
        completed by Hermitian symmetry, C2R-transformed and cropped to (n, NX,
        NX) float32 images (into `out` when given)."""
        torch = _torch()
        if isinstance(X, torch.Tensor):
            X = X.resolve_conj().resolve_neg().contiguous()
        X = self._grid_tensor(X, "X")
        n = (np.asarray(dm).reshape(-1, 9) if dm is not None else np.asarray(angles).reshape(-1, 3)).shape[0]
        ang, dmc = self._rotation_args(n, angles, dm)
        if out is None:
            out = torch.empty((n, self.nx, self.nx), dtype=torch.float32, device=self.device)
        elif out.dtype != torch.float32 or tuple(out.shape) != (n, self.nx, self.nx) or not out.is_contiguous() \
                or out.device != self.device:
            raise ValueError(f"out must be contiguous float32 ({n}, {self.nx}, {self.nx}) on {self.device}")
        self._check(_lib.lib().spbp_gather_images(self._ctx, C.c_void_p(X.data_ptr()), n,
                                                  None if ang is None else ang.ctypes.data_as(C.c_void_p),
                                                  None if dmc is None else dmc.ctypes.data_as(C.c_void_p),
                                                  C.c_void_p(out.data_ptr()), _stream(self.device)))
        return out

    def scatter_images(self, imgs, angles=None, dm=None, X=None):
        """The central-slice projector's adjoint applied to real images (n, NX,
        NX) float32 on the context's device, as spectra added into X (a new
        zero grid when None); power-of-two padded sizes only.  Equals
        scatter() of the BI spectra the adjoint's chain builds from imgs."""
        torch = _torch()
        if not isinstance(imgs, torch.Tensor) or imgs.dtype != torch.float32 or imgs.dim() != 3 or \
                tuple(imgs.shape[1:]) != (self.nx, self.nx):
            raise ValueError(f"imgs must be float32 (n, {self.nx}, {self.nx})")
        if imgs.device != self.device:
            raise ValueError(f"imgs on {imgs.device}, context on {self.device}")
        imgs = imgs.contiguous()
        n = imgs.shape[0]
        ang, dmc = self._rotation_args(n, angles, dm)
        if X is None:
            X = torch.zeros(self.grid_shape, dtype=torch.complex64, device=self.device)
        X = self._grid_tensor(X, "X")
        self._check(_lib.lib().spbp_scatter_images(self._ctx, C.c_void_p(imgs.data_ptr()), n,
                                                   None if ang is None else ang.ctypes.data_as(C.c_void_p),
                                                   None if dmc is None else dmc.ctypes.data_as(C.c_void_p),
                                                   C.c_void_p(X.data_ptr()), _stream(self.device)))
        self._keep = (imgs,)
        return X

    def scatter(self, B, angles=None, dm=None, X=None, W=None):
        """ONELINE scatter of BI-layout spectra B (n, N, NX+1) complex64, added
        into X (a new zero grid when None) and W when given; returns X."""
        torch = _torch()
        N, _, H = self.grid_shape
        if not isinstance(B, torch.Tensor) or B.dtype != torch.complex64 or B.dim() != 3 or tuple(B.shape[1:]) != (N, H):
            raise ValueError(f"B must be complex64 (n, {N}, {H})")
        if B.device != self.device:
            raise ValueError(f"B on {B.device}, context on {self.device}")
        B = B.resolve_conj().resolve_neg().contiguous()
        n = B.shape[0]
        ang, dmc = self._rotation_args(n, angles, dm)
        if X is None:
            X = torch.zeros(self.grid_shape, dtype=torch.complex64, device=self.device)
        X = self._grid_tensor(X, "X")
        wp = None
        if W is not None:
            if W.dtype != torch.float32 or tuple(W.shape) != self.grid_shape or not W.is_contiguous() or W.device != self.device:
                raise ValueError(f"W must be contiguous float32 {self.grid_shape} on {self.device}")
            wp = C.c_void_p(W.data_ptr())
        self._check(_lib.lib().spbp_scatter(self._ctx, C.c_void_p(B.data_ptr()), n,
                                            None if ang is None else ang.ctypes.data_as(C.c_void_p),
                                            None if dmc is None else dmc.ctypes.data_as(C.c_void_p),
                                            C.c_void_p(X.data_ptr()), wp, _stream(self.device)))
        self._keep = (B,)
        return X

    # -------------------------------------------------------------- plans
    def tune(self, imgs, angles=None, dm=None) -> dict:
        """Time every insertion plan on a sample of these images (up to 1024;
        the accumulators are untouched) and keep the fastest.  Returns plan()."""
        imgs, dt = self._images(imgs)
        n = imgs.shape[0]
        ang, dmc = self._rotation_args(n, angles, dm)
        self._check(_lib.lib().spbp_tune(self._ctx, C.c_void_p(imgs.data_ptr()), dt, n,
                                         None if ang is None else ang.ctypes.data_as(C.c_void_p),
                                         None if dmc is None else dmc.ctypes.data_as(C.c_void_p),
                                         _stream(self.device)))
        return self.plan()

    def plan(self) -> dict:
        st, sh, cp = C.c_int(), C.c_int(), C.c_int()
        nb = len(_lib.SHAPES) * len(_lib.TUNE_C)
        ms = (C.c_float * (nb + 2))()
        self._check(_lib.lib().spbp_get_plan(self._ctx, C.byref(st), C.byref(sh), C.byref(cp), ms))
        table = np.array(ms[:nb], np.float64).reshape(len(_lib.SHAPES), len(_lib.TUNE_C))
        return {"strategy": {_lib.STRAT_DIRECT: "direct", _lib.STRAT_BRICK: "brick",
                             _lib.STRAT_GATHER: "gather"}[st.value],
                "brick": _lib.SHAPES[sh.value], "owner_computes": _lib.SHAPE_OWN[sh.value], "shape": sh.value,
                "brick_proj": cp.value,
                "ms_per_proj": table, "direct_ms_per_proj": float(ms[nb]), "gather_ms_per_proj": float(ms[nb + 1])}

    def set_plan(self, strategy="brick", shape: int = 0, brick_proj: int = 0) -> None:
        """strategy: "brick" (LDS bins of brick shape `shape`), "direct"
        (fp32 atomics per tap), "gather" (voxel-driven, the serial
        reference's summation order) or "auto"."""
        strat = {"auto": _lib.STRAT_AUTO, "direct": _lib.STRAT_DIRECT, "brick": _lib.STRAT_BRICK,
                 "gather": _lib.STRAT_GATHER}[strategy]
        self._check(_lib.lib().spbp_set_plan(self._ctx, strat, int(shape), int(brick_proj)))

    def _job(self, imgs, angles=None, dm=None, half=None):
        imgs, dt = self._images(imgs)
        n = imgs.shape[0]
        ang, dmc = self._rotation_args(n, angles, dm)
        hp = None if half is None else np.ascontiguousarray(np.asarray(half).astype(np.int8).reshape(n))
        job = _lib.Job(self._ctx.value, imgs.data_ptr(), dt, n,
                       None if ang is None else ang.ctypes.data, None if dmc is None else dmc.ctypes.data,
                       None if hp is None else hp.ctypes.data)
        return job, (imgs, ang, dmc, hp)

    # -------------------------------------------------------------- finalisation
    def _final_outputs(self, full, halves) -> dict:
        torch = _torch()
        nx = self.nx
        out = {}
        if full:
            out["full"] = torch.empty((nx, nx, nx), dtype=torch.float32, device=self.device)
        if halves:
            if not self.two_halves:
                raise SpbpError(_lib.ERR_ARG, "sample volumes need two_halves=True")
            out["h1"] = torch.empty((nx, nx, nx), dtype=torch.float32, device=self.device)
            out["h2"] = torch.empty((nx, nx, nx), dtype=torch.float32, device=self.device)
        return out

    def finalize(self, full=True, halves=False) -> dict:
        """NX^3 volumes: 'full' (BP 3F volume / BP 32F overall volume) and, for
        two-halves contexts, 'h1', 'h2' (first / second sample volume)."""
        out = self._final_outputs(full, halves)
        ptr = lambda k: C.c_void_p(out[k].data_ptr()) if k in out else None
        self._check(_lib.lib().spbp_finalize(self._ctx, ptr("full"), ptr("h1"), ptr("h2"), _stream(self.device)))
        return out

    # -------------------------------------------------------------- partial volumes
    def export_partial(self, half: int = -1):
        """Post-SYMPLANE0 half grid (X complex64, W float32), shape (N, N, NX+1):
        the content of `BP 3F P`'s volume and _W files (bp3f.f:236-248)."""
        torch = _torch()
        X = torch.empty(self.grid_shape, dtype=torch.complex64, device=self.device)
        W = torch.empty(self.grid_shape, dtype=torch.float32, device=self.device)
        self._check(_lib.lib().spbp_export_partial(self._ctx, int(half), C.c_void_p(X.data_ptr()),
                                                   C.c_void_p(W.data_ptr()), _stream(self.device)))
        return X, W

    def export_raw(self, half: int = 0):
        """Raw accumulators (before SYMPLANE0) of one half, shape (N, N, NX+1)."""
        torch = _torch()
        X = torch.empty(self.grid_shape, dtype=torch.complex64, device=self.device)
        W = torch.empty(self.grid_shape, dtype=torch.float32, device=self.device)
        self._check(_lib.lib().spbp_export_raw(self._ctx, int(half), C.c_void_p(X.data_ptr()),
                                               C.c_void_p(W.data_ptr()), _stream(self.device)))
        return X, W

    def kernel_stats(self, kernel: int = _lib.K_INSERT):
        """(total device ms, launches, units) recorded with timing=True."""
        ms, n, u = C.c_double(), C.c_long(), C.c_long()
        self._check(_lib.lib().spbp_kernel_stats(self._ctx, int(kernel), C.byref(ms), C.byref(n), C.byref(u)))
        return ms.value, n.value, u.value

    def reset_stats(self) -> None:
        self._check(_lib.lib().spbp_reset_stats(self._ctx))

    def import_partial(self, X, W, half: int = 0, accumulate: bool = True) -> None:
        """Merge a post-SYMPLANE0 half grid (`BP 3F M`, bp3f.f:646-685)."""
        torch = _torch()
        X = torch.as_tensor(X, device=self.device).to(torch.complex64).contiguous()
        W = torch.as_tensor(W, device=self.device).to(torch.float32).contiguous()
        if tuple(X.shape) != self.grid_shape or tuple(W.shape) != self.grid_shape:
            raise SpbpError(_lib.ERR_SIZE, f"partial volume shape {tuple(X.shape)} != {self.grid_shape}")
        self._check(_lib.lib().spbp_import_partial(self._ctx, int(half), C.c_void_p(X.data_ptr()),
                                                   C.c_void_p(W.data_ptr()), int(bool(accumulate)),
                                                   _stream(self.device)))
        self._keep = (X, W)


def rtsq(imgs, params, fbs: bool = False):
    """RTSQ (rtsq.f:56, scale 1) of device images (n, NX, NX) fp32/bf16 by
    params (n, 3) = (angle in degrees, x shift, y shift): fp32 (n, NX, NX).
    fbs: RTSF instead (FBS ON, rtsf.f:106: Fourier-based bicubic
    interpolation)."""
    torch = _torch()
    if imgs.dim() != 3 or imgs.shape[1] != imgs.shape[2]:
        raise SpbpError(_lib.ERR_SIZE, "ONLY WORKS ON SQUARE IMAGES")
    imgs = imgs.contiguous()
    dt = _lib.BF16 if imgs.dtype == torch.bfloat16 else _lib.F32
    if dt == _lib.F32 and imgs.dtype != torch.float32:
        raise TypeError(f"unsupported image dtype {imgs.dtype}")
    p = np.ascontiguousarray(np.asarray(params, np.float32).reshape(-1, 3))
    out = torch.empty(imgs.shape, dtype=torch.float32, device=imgs.device)
    fn = _lib.lib().spbp_rtsf if fbs else _lib.lib().spbp_rtsq
    _lib.check(fn(C.c_void_p(imgs.data_ptr()), dt, imgs.shape[0], imgs.shape[1],
                  p.ctypes.data_as(C.c_void_p), C.c_void_p(out.data_ptr()), _stream(imgs.device)))
    return out


def bp3f(imgs, angles, exact_pi: bool = False, inplane=None, **kw):
    """`BP 3F` (`RB 3F` with inplane): one volume from all projections
    (symmetries=(nsym, 3) doc rows optional)."""
    with Backprojector(imgs.shape[-1], exact_pi=exact_pi, device=imgs.device.index, **kw) as bp:
        bp.insert(imgs, angles, inplane=inplane)
        return bp.finalize()["full"]


def bp32f(imgs, angles, half=None, seed=None, exact_pi: bool = False, inplane=None, **kw):
    """`BP 32F`: overall volume plus the two sample volumes.  `half` gives the
    split (0 = first sample volume); default is a RANDLIST-style random split."""
    if half is None:
        half = random_halves(imgs.shape[0], seed)
    with Backprojector(imgs.shape[-1], two_halves=True, exact_pi=exact_pi, device=imgs.device.index, **kw) as bp:
        bp.insert(imgs, angles, half=half, inplane=inplane)
        out = bp.finalize(full=True, halves=True)
    return out["full"], out["h1"], out["h2"]


# ------------------------------------------------------------------ file-level operations
def read_symmetries(symdoc, ext: str = ""):
    """Symmetry doc rows (PSI, THETA, PHI) for keys 1..max key, or None for
    '*' (GETDOCDAT + BUILDS, bp32f.f:133-172, builds.f:44-52)."""
    from . import spiderio
    if symdoc in (None, "*", ""):
        return None
    path = symdoc if (not ext or symdoc.endswith(ext)) else symdoc + ext
    doc = spiderio.read_doc(path)
    if not doc:
        return None
    rows = []
    for k in range(1, max(doc) + 1):
        if k not in doc or len(doc[k]) < 3:
            raise SpbpError(_lib.ERR_ANGLE, f"MISSING SYMMETRY {k}")
        rows.append(doc[k][:3])
    return np.asarray(rows, np.float64)


def _load_series(template, numbers, angdoc, symdoc, ext, device, inplane=False):
    from . import spiderio
    torch = _torch()
    imgs = spiderio.read_images(template, numbers, ext)
    if imgs.shape[1] != imgs.shape[2]:
        raise SpbpError(_lib.ERR_SIZE, "ONLY WORKS ON SQUARE IMAGES")
    try:
        doc = spiderio.read_doc(angdoc)
        ang = spiderio.angles_for(doc, numbers)
        rot = spiderio.inplane_for(doc, numbers) if inplane else None
    except KeyError as e:
        raise SpbpError(_lib.ERR_ANGLE, str(e)) from None
    dev = torch.device("cuda", torch.cuda.current_device() if device is None else int(device))
    if inplane:
        return torch.from_numpy(imgs).to(dev), ang, rot
    return torch.from_numpy(imgs).to(dev), ang


def bp3f_files(template, numbers, angdoc, symdoc, out, ext: str = "", device=None, **kw):
    """`BP 3F`: template for 2D images, file numbers, angles doc, symmetries
    doc ('*'), reconstructed volume (the answers of bp3f.f:127-283)."""
    from . import spiderio
    imgs, ang = _load_series(template, numbers, angdoc, symdoc, ext, device)
    vol = bp3f(imgs, ang, symmetries=read_symmetries(symdoc, ext), **kw).cpu().numpy()
    spiderio.write_image(out, vol)
    return vol


def bp32f_files(template, numbers, angdoc, symdoc, out_full, out_h1, out_h2, ext: str = "", seed=None,
                half=None, device=None, **kw):
    """`BP 32F`: as BP 3F plus the two sample volumes (bp32f.f:102-410)."""
    from . import spiderio
    imgs, ang = _load_series(template, numbers, angdoc, symdoc, ext, device)
    full, h1, h2 = bp32f(imgs, ang, half=half, seed=seed, symmetries=read_symmetries(symdoc, ext), **kw)
    vols = [v.cpu().numpy() for v in (full, h1, h2)]
    for path, v in zip((out_full, out_h1, out_h2), vols):
        spiderio.write_image(path, v)
    return vols


def _write_transformed(rot_out, numbers, imgs, rot, ext, fbs=False):
    """The transformed-images output of `RB 3F` / `RB 32F` (bp32f.f:517-527,
    826-827): every image after its RTSQ (RTSF with fbs), under the template
    rot_out ('*' or None: not written), with the input's file numbers."""
    from . import spiderio
    if rot_out in (None, "", "*"):
        return
    spiderio.write_images(rot_out, numbers, rtsq(imgs, rot, fbs=fbs).cpu().numpy(), ext)


def rb3f_files(template, numbers, angdoc, symdoc, out, ext: str = "", device=None, rot_out=None, **kw):
    """`RB 3F`: as `BP 3F`, each image first rotated by its in-plane angle and
    shifted (doc registers 6, 7, 8; RTSQ, bp32f.f:805-837, or RTSF with
    fbs=True, FBS ON); rot_out is the template for the transformed images
    (None or '*': not written)."""
    from . import spiderio
    imgs, ang, rot = _load_series(template, numbers, angdoc, symdoc, ext, device, inplane=True)
    vol = bp3f(imgs, ang, inplane=rot, symmetries=read_symmetries(symdoc, ext), **kw).cpu().numpy()
    spiderio.write_image(out, vol)
    _write_transformed(rot_out, numbers, imgs, rot, ext, kw.get("fbs", False))
    return vol


def rb32f_files(template, numbers, angdoc, symdoc, out_full, out_h1, out_h2, ext: str = "", seed=None,
                half=None, device=None, rot_out=None, **kw):
    """`RB 32F`: `BP 32F` with the RTSQ pre-rotation of `RB 3F` and its
    transformed-images output."""
    from . import spiderio
    imgs, ang, rot = _load_series(template, numbers, angdoc, symdoc, ext, device, inplane=True)
    _write_transformed(rot_out, numbers, imgs, rot, ext, kw.get("fbs", False))
    full, h1, h2 = bp32f(imgs, ang, half=half, seed=seed, inplane=rot, symmetries=read_symmetries(symdoc, ext), **kw)
    vols = [v.cpu().numpy() for v in (full, h1, h2)]
    for path, v in zip((out_full, out_h1, out_h2), vols):
        spiderio.write_image(path, v)
    return vols


def bp3f_partial_files(template, numbers, angdoc, symdoc, out, ext: str = "", device=None, **kw):
    """`BP 3F P`: write the post-SYMPLANE0 Fourier half grid and its weights
    (bp3f.f:191-248) instead of a volume."""
    from . import spiderio
    imgs, ang = _load_series(template, numbers, angdoc, symdoc, ext, device)
    with Backprojector(imgs.shape[-1], device=imgs.device.index, symmetries=read_symmetries(symdoc, ext), **kw) as b:
        b.insert(imgs, ang)
        X, W = b.export_partial()
        return spiderio.write_partial(out, X.cpu().numpy(), W.cpu().numpy())


def bp3f_merge_files(partials, out, device=None, symdoc=None, ext: str = "", **kw):
    """`BP 3F M`: sum partial volumes and finalise (bp3f.f:531-739); with a
    symmetries doc the merged volume is symmetrised by SYMVOL (bp3f.f:700-722)."""
    from . import spiderio
    torch = _torch()
    # every shard is read and checked before any device work, so a lost,
    # truncated or mismatched one stops the merge with nothing written (the
    # reference stops at the file it cannot open, bp3f.f:646-685)
    pairs = [spiderio.read_partial(p) for p in partials]
    if not pairs:
        raise SpbpError(_lib.ERR_ARG, "no partial volumes")
    nx = pairs[0][1].shape[2] - 1
    for p, (X, W) in zip(partials, pairs):
        if W.shape != pairs[0][1].shape:
            raise spiderio.SpiderFormatError(f"{p}: partial volume size {W.shape} differs from {pairs[0][1].shape}")
    dev = torch.device("cuda", torch.cuda.current_device() if device is None else int(device))
    with Backprojector(nx, device=dev.index, symmetries=read_symmetries(symdoc, ext), **kw) as b:
        for X, W in pairs:
            b.import_partial(torch.from_numpy(X).to(dev), torch.from_numpy(W).to(dev), accumulate=True)
        vol = b.finalize()["full"].cpu().numpy()
    spiderio.write_image(out, vol)
    return vol


def _jobs(items):
    """items: (Backprojector, imgs, angles) or (Backprojector, imgs, angles, half)
    tuples, or dicts with keys ctx, imgs, angles / dm, half."""
    jobs, keep = [], []
    for it in items:
        if isinstance(it, dict):
            j, k = it["ctx"]._job(it["imgs"], it.get("angles"), it.get("dm"), it.get("half"))
        else:
            ctx, imgs, angles = it[0], it[1], it[2]
            j, k = ctx._job(imgs, angles, None, it[3] if len(it) > 3 else None)
        jobs.append(j)
        keep.append((it["ctx"] if isinstance(it, dict) else it[0], k))   # contexts outlive the jobs' users
    arr = (_lib.Job * len(jobs))(*jobs)
    return arr, keep


def insert_batch(items, device=None) -> None:
    """Insert several reconstructions (any box sizes) with one persistent
    scatter launch per round of chunks (spbp_insert_batch)."""
    arr, keep = _jobs(items)
    _lib.check(_lib.lib().spbp_insert_batch(arr, len(arr), _stream(device)))
    return keep


def finalize_batch(items, full=True, halves=False, device=None) -> list:
    """Finalise several contexts (any box sizes) in one pass
    (spbp_finalize_batch): the threshold re-summation of all volumes shares
    its host synchronisations, and each wave of volumes -- every context's
    first requested volume, then the second, then the third -- takes one
    normalisation and one window launch.  items: Backprojectors (finalised
    with `full` / `halves`) or (Backprojector, full, halves) tuples.
    Returns one finalize()-style dict per context."""
    ctxs, outs = [], []
    for it in items:
        c, f, h = it if isinstance(it, tuple) else (it, full, halves)
        ctxs.append(c)
        outs.append(c._final_outputs(f, h))
    ptr = lambda o, k: o[k].data_ptr() if k in o else None
    arr = (_lib.FinalJob * len(ctxs))(*[_lib.FinalJob(c._ctx.value, ptr(o, "full"), ptr(o, "h1"), ptr(o, "h2"))
                                         for c, o in zip(ctxs, outs)])
    dev = ctxs[0].device if device is None and ctxs else device
    _lib.check(_lib.lib().spbp_finalize_batch(arr, len(arr), _stream(dev)))
    return outs


class BatchGraph:
    """A batch of insertions captured once as a HIP graph and replayed:
    `launch()` re-runs the whole batch (zeroing the accumulators first when
    reset_first) on torch's current stream.  Images are re-read from the same
    device tensors at every launch; the geometry is fixed at capture."""

    def __init__(self, items, reset_first: bool = True):
        arr, self._keep = _jobs(items)
        g = C.c_void_p()
        _lib.check(_lib.lib().spbp_graph_create(arr, len(arr), 1 if reset_first else 0, C.byref(g)))
        self._g = g
        self.device = _torch().device("cuda", _torch().cuda.current_device())

    def launch(self) -> None:
        _lib.check(_lib.lib().spbp_graph_launch(self._g, _stream(self.device)))

    def close(self) -> None:
        """Destroy the graph; its contexts become free to reallocate again.
        The graph holds references to its Backprojectors until then."""
        if getattr(self, "_g", None) is not None and self._g.value:
            _lib.lib().spbp_graph_destroy(self._g)
            self._g = None
        self._keep = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
