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
                 timing: bool = False, strategy: str = "auto", brick_proj: int = 0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("spider_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else int(device))
        flags = (_lib.TWO_HALVES if two_halves else 0) | (_lib.EXACT_PI if exact_pi else 0) | \
                (_lib.TWOSUM_W if twosum else 0) | (_lib.CHECK_FINITE if check_finite else 0) | \
                (_lib.TIMING if timing else 0)
        strat = {"auto": _lib.STRAT_AUTO, "direct": _lib.STRAT_DIRECT, "brick": _lib.STRAT_BRICK}[strategy]
        opts = _lib.Opts(int(nx), flags, self.device.index, int(batch), strat, int(brick_proj))
        ctx = C.c_void_p()
        L = _lib.lib()
        _lib.check(L.spbp_create(C.byref(opts), C.byref(ctx)), None)
        self._ctx = ctx
        self.nx = int(nx)
        self.two_halves = bool(two_halves)

    # -------------------------------------------------------------- lifecycle
    def close(self) -> None:
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            _lib.lib().spbp_destroy(self._ctx)
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

    def insert(self, imgs, angles, half=None) -> None:
        """Insert projections with doc-file angles (n, 3): PSI, THETA, PHI."""
        imgs, dt = self._images(imgs)
        n = imgs.shape[0]
        ang = np.ascontiguousarray(np.asarray(angles, np.float64).reshape(-1, 3))
        if ang.shape[0] != n:
            raise SpbpError(_lib.ERR_ANGLE, f"MISSING ANGLE FOR IMAGE: {ang.shape[0]} angles for {n} images")
        hp = None
        if half is not None:
            hp = np.ascontiguousarray(np.asarray(half).astype(np.int8).reshape(n))
        self._check(_lib.lib().spbp_insert(self._ctx, C.c_void_p(imgs.data_ptr()), dt, n,
                                           ang.ctypes.data_as(C.c_void_p),
                                           None if hp is None else hp.ctypes.data_as(C.c_void_p),
                                           _stream(self.device)))
        self._keep = (imgs, ang, hp)   # alive until the next call orders after this one

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

    # -------------------------------------------------------------- finalisation
    def finalize(self, full=True, halves=False) -> dict:
        """NX^3 volumes: 'full' (BP 3F volume / BP 32F overall volume) and, for
        two-halves contexts, 'h1', 'h2' (first / second sample volume)."""
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


def bp3f(imgs, angles, exact_pi: bool = False, **kw):
    """`BP 3F`: one volume from all projections."""
    with Backprojector(imgs.shape[-1], exact_pi=exact_pi, device=imgs.device.index, **kw) as bp:
        bp.insert(imgs, angles)
        return bp.finalize()["full"]


def bp32f(imgs, angles, half=None, seed=None, exact_pi: bool = False, **kw):
    """`BP 32F`: overall volume plus the two sample volumes.  `half` gives the
    split (0 = first sample volume); default is a RANDLIST-style random split."""
    if half is None:
        half = random_halves(imgs.shape[0], seed)
    with Backprojector(imgs.shape[-1], two_halves=True, exact_pi=exact_pi, device=imgs.device.index, **kw) as bp:
        bp.insert(imgs, angles, half=half)
        out = bp.finalize(full=True, halves=True)
    return out["full"], out["h1"], out["h2"]


# ------------------------------------------------------------------ file-level operations
def _load_series(template, numbers, angdoc, symdoc, ext, device):
    from . import spiderio
    torch = _torch()
    if symdoc not in (None, "*", ""):
        raise NotImplementedError("symmetry documents are not supported (answer '*')")
    imgs = spiderio.read_images(template, numbers, ext)
    if imgs.shape[1] != imgs.shape[2]:
        raise SpbpError(_lib.ERR_SIZE, "ONLY WORKS ON SQUARE IMAGES")
    try:
        ang = spiderio.angles_for(spiderio.read_doc(angdoc), numbers)
    except KeyError as e:
        raise SpbpError(_lib.ERR_ANGLE, str(e)) from None
    dev = torch.device("cuda", torch.cuda.current_device() if device is None else int(device))
    return torch.from_numpy(imgs).to(dev), ang


def bp3f_files(template, numbers, angdoc, symdoc, out, ext: str = "", device=None, **kw):
    """`BP 3F`: template for 2D images, file numbers, angles doc, symmetries
    doc ('*'), reconstructed volume (the answers of bp3f.f:127-283)."""
    from . import spiderio
    imgs, ang = _load_series(template, numbers, angdoc, symdoc, ext, device)
    vol = bp3f(imgs, ang, **kw).cpu().numpy()
    spiderio.write_image(out, vol)
    return vol


def bp32f_files(template, numbers, angdoc, symdoc, out_full, out_h1, out_h2, ext: str = "", seed=None,
                half=None, device=None, **kw):
    """`BP 32F`: as BP 3F plus the two sample volumes (bp32f.f:102-410)."""
    from . import spiderio
    imgs, ang = _load_series(template, numbers, angdoc, symdoc, ext, device)
    full, h1, h2 = bp32f(imgs, ang, half=half, seed=seed, **kw)
    vols = [v.cpu().numpy() for v in (full, h1, h2)]
    for path, v in zip((out_full, out_h1, out_h2), vols):
        spiderio.write_image(path, v)
    return vols


def bp3f_partial_files(template, numbers, angdoc, symdoc, out, ext: str = "", device=None, **kw):
    """`BP 3F P`: write the post-SYMPLANE0 Fourier half grid and its weights
    (bp3f.f:191-248) instead of a volume."""
    from . import spiderio
    imgs, ang = _load_series(template, numbers, angdoc, symdoc, ext, device)
    with Backprojector(imgs.shape[-1], device=imgs.device.index, **kw) as b:
        b.insert(imgs, ang)
        X, W = b.export_partial()
        return spiderio.write_partial(out, X.cpu().numpy(), W.cpu().numpy())


def bp3f_merge_files(partials, out, device=None, **kw):
    """`BP 3F M`: sum partial volumes and finalise (bp3f.f:531-739)."""
    from . import spiderio
    torch = _torch()
    pairs = [spiderio.read_partial(p) for p in partials]
    if not pairs:
        raise SpbpError(_lib.ERR_ARG, "no partial volumes")
    nx = pairs[0][1].shape[2] - 1
    dev = torch.device("cuda", torch.cuda.current_device() if device is None else int(device))
    with Backprojector(nx, device=dev.index, **kw) as b:
        for X, W in pairs:
            b.import_partial(torch.from_numpy(X).to(dev), torch.from_numpy(W).to(dev), accumulate=True)
        vol = b.finalize()["full"].cpu().numpy()
    spiderio.write_image(out, vol)
    return vol
