"""Differentiable central-slice projector whose adjoint is the BP 3F insertion.

The projector A maps an NX^3 volume to NX x NX projections by the
Fourier-slice theorem on the grid the backprojection uses:

    v --(divide by the KB kernel's transform D)--> u
      --(zero-pad to N^3 at IP = (N-NX)/2 + mod(NX,2), as WINDKB2A crops,
         bp32f.f:945-960)--> u_pad
      --(3D transform with SPIDER's sign and centring: X = (-1)^(i+j+k)
         conj(rfftn(u_pad)), the inverse of NRMW2's c2r, bp3f.f:1355-1366)--> X
      --(KB gather of each projection's central slice: the exact adjoint of
         ONELINE, libspbp spbp_gather)--> BI
      --(inverse of the BI spectrum: conj((-1)^(I+J) BI), the kx = 0 column
         completed by Hermitian symmetry (ONELINE never samples I = 0, J < 0,
         oneline.f:53-55), irfft2, crop at the PADD2 offset,
         bp3f.f:1336-1341)--> images.

`SliceGather` is the torch.autograd.Function around the gather; its
hand-written backward is the insertion kernel (spbp_scatter).  All other
steps are torch operations, so autograd differentiates the whole chain.
`Projector.adjoint` restates A^T explicitly for CG (no autograd graph).

D(x, y, z) = k^(x) k^(y) k^(z) with k^ the Fourier transform of the 1-D
table kernel TABI(NINT(|s| FLTB)) (bp32f.f:193-214) at the voxel's distance
from the centre: gathering with the kernel multiplies the volume by D, which
the division undoes.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import _lib
from . import bp as _bp


def _torch():
    import torch
    return torch


def kernel_transform(nx: int) -> np.ndarray:
    """k^(d) for d = i - NX//2, i = 0..NX-1 (FP64): the continuous Fourier
    transform of the piecewise-constant table kernel at frequency d / N."""
    N = 2 * nx
    tab, fltb = _lib.kb_table(N)
    tab = tab.astype(np.float64)
    fltb = float(fltb)
    i = np.arange(tab.size, dtype=np.float64)
    lo = np.maximum(i - 0.5, 0.0) / fltb
    hi = (i + 0.5) / fltb
    d = np.arange(nx, dtype=np.float64) - nx // 2
    w = 2.0 * np.pi * d / N
    out = np.empty(nx)
    for q, om in enumerate(w):
        if om == 0.0:
            seg = hi - lo
        else:
            seg = (np.sin(om * hi) - np.sin(om * lo)) / om
        out[q] = 2.0 * np.dot(tab, seg)
    return out


class SliceOp:
    """Gather / scatter of central slices for a fixed set of rotations on one
    libspbp context."""

    def __init__(self, ctx: "_bp.Backprojector", dm: np.ndarray):
        self.ctx = ctx
        self.dm = np.ascontiguousarray(np.asarray(dm, np.float32).reshape(-1, 9))

    def gather(self, X):
        return self.ctx.gather(X, dm=self.dm)

    def scatter(self, B):
        return self.ctx.scatter(B, dm=self.dm)


def _slice_gather_fn():
    torch = _torch()

    class SliceGather(torch.autograd.Function):
        """B = G X (central-slice KB gather); backward: dX = G^T dB = ONELINE
        scatter of dB.  Complex tensors; the gradient convention is torch's
        (d/dRe + i d/dIm), for which the transpose over pairs of reals is
        the adjoint."""

        @staticmethod
        def forward(ctx, X, op):
            ctx.op = op
            return op.gather(X.contiguous())

        @staticmethod
        def backward(ctx, gB):
            return ctx.op.scatter(gB.contiguous()), None

    return SliceGather


_SLICE_GATHER = None


def slice_gather(X, op: SliceOp):
    global _SLICE_GATHER
    if _SLICE_GATHER is None:
        _SLICE_GATHER = _slice_gather_fn()
    return _SLICE_GATHER.apply(X, op)


class Projector:
    """A: NX^3 volume -> projections for the given rotations (CANG matrices
    (n, 9), or doc angles (n, 3) PSI, THETA, PHI)."""

    def __init__(self, nx: int, angles=None, dm=None, exact_pi: bool = False, device=None,
                 strategy: str = "auto", batch: int = 2048, timing: bool = False):
        torch = _torch()
        self.nx = int(nx)
        self.N = 2 * self.nx
        self.H = self.nx + 1
        self.ip = (self.N - self.nx) // 2 + self.nx % 2
        self.ctx = _bp.Backprojector(self.nx, exact_pi=exact_pi, device=device, strategy=strategy, timing=timing)
        self.device = self.ctx.device
        if dm is None:
            if angles is None:
                raise ValueError("Projector needs angles or rotations")
            dm = _lib.rotations(angles, exact_pi=exact_pi)
        self.dm = np.ascontiguousarray(np.asarray(dm, np.float32).reshape(-1, 9))
        self.batch = int(batch)
        # apply() and adjoint(): the library synthesises the images from the
        # gather (completion, C2R, crop) and pads, transforms and folds the
        # residual images before the scatter (the adjoint for power-of-two N
        # only); False runs those steps as torch operations around
        # spbp_gather / spbp_scatter (the chain the library's versions are
        # checked against)
        self.library_chain = True
        k = torch.as_tensor(kernel_transform(self.nx), dtype=torch.float64)
        self.D = (k[:, None, None] * k[None, :, None] * k[None, None, :]).to(torch.float32).to(self.device)
        N, H = self.N, self.H
        ii = torch.arange(N, device=self.device)
        hh = torch.arange(H, device=self.device)
        self.chk2 = (1 - 2 * ((ii[:, None] + hh[None, :]) % 2)).to(torch.float32)          # (N, H)
        self.chk3 = (1 - 2 * ((ii[:, None, None] + ii[None, :, None] + hh[None, None, :]) % 2)).to(torch.float32)
        # irfft along x counts columns 1..N/2-1 twice
        c = torch.full((H,), 2.0, device=self.device)
        c[0] = 1.0
        c[-1] = 1.0
        self.cx = c
        jp = torch.arange(N, device=self.device)
        self.neg_rows = jp > self.nx                  # rows J < 0 of the BI layout
        self.mirror = (N - jp) % N

    @property
    def nproj(self) -> int:
        return self.dm.shape[0]

    def close(self):
        self.ctx.close()

    # ---------------------------------------------------------------- pieces
    def grid(self, v):
        """X = (-1)^(i+j+k) conj(rfftn(pad(v / D)))."""
        torch = _torch()
        u = v / self.D
        p0, p1 = self.ip, self.N - self.ip - self.nx
        up = torch.nn.functional.pad(u, (p0, p1, p0, p1, p0, p1))
        return self.chk3 * torch.conj(torch.fft.rfftn(up))

    def images_from_bi(self, B):
        torch = _torch()
        Bc = torch.conj(B * self.chk2)
        col = Bc[:, :, 0]
        col = torch.where(self.neg_rows[None, :], torch.conj(col[:, self.mirror]), col)
        Bc = torch.cat([col[:, :, None], Bc[:, :, 1:]], dim=2)
        img = torch.fft.irfft2(Bc, s=(self.N, self.N))
        return img[:, self.ip:self.ip + self.nx, self.ip:self.ip + self.nx]

    def forward(self, v, sel=None):
        """Projections (n, NX, NX) of v (NX^3 float32) for the rotations
        selected by `sel` (slice / index array; all when None).
        Differentiable in v."""
        X = self.grid(v)
        dm = self.dm if sel is None else self.dm[sel]
        B = slice_gather(X, SliceOp(self.ctx, dm))
        return self.images_from_bi(B)

    __call__ = forward

    # ---------------------------------------------------------------- explicit A and A^T
    def apply(self, v, out=None):
        """A v over all rotations, in batches, without an autograd graph."""
        torch = _torch()
        with torch.no_grad():
            X = self.grid(v)
            n = self.nproj
            if out is None:
                out = torch.empty((n, self.nx, self.nx), dtype=torch.float32, device=self.device)
            for i0 in range(0, n, self.batch):
                i1 = min(n, i0 + self.batch)
                if self.library_chain:
                    self.ctx.gather_images(X, dm=self.dm[i0:i1], out=out[i0:i1])
                else:
                    out[i0:i1] = self.images_from_bi(self.ctx.gather(X, dm=self.dm[i0:i1]))
        return out

    def adjoint(self, r):
        """A^T r for projections r (n, NX, NX): the transpose of every step of
        forward(), with the gather's transpose the ONELINE scatter."""
        torch = _torch()
        N, nx, ip = self.N, self.nx, self.ip
        with torch.no_grad():
            X = torch.zeros((N, N, self.H), dtype=torch.complex64, device=self.device)
            if self.library_chain and N & (N - 1) == 0:
                # the library's own padded transforms and spectrum folding (power-of-two N)
                rr = r.to(torch.float32).contiguous()
                for i0 in range(0, rr.shape[0], self.batch):
                    i1 = min(rr.shape[0], i0 + self.batch)
                    self.ctx.scatter_images(rr[i0:i1], dm=self.dm[i0:i1], X=X)
                F = torch.conj(X * self.chk3) / self.cx
                up = torch.fft.irfftn(F, s=(N, N, N)) * float(N) ** 3
                return up[ip:ip + nx, ip:ip + nx, ip:ip + nx] / self.D
            for i0 in range(0, r.shape[0], self.batch):
                i1 = min(r.shape[0], i0 + self.batch)
                g = torch.zeros((i1 - i0, N, N), dtype=torch.float32, device=self.device)
                g[:, ip:ip + nx, ip:ip + nx] = r[i0:i1]
                # transpose of irfft2(s=(N, N)): rfft2 * c / N^2
                G = torch.fft.rfft2(g) * (self.cx / float(N * N))
                # transpose of the Hermitian completion of column 0
                col = G[:, :, 0]
                add = torch.conj(col[:, self.mirror])
                lo = (torch.arange(N, device=self.device) >= 1) & (torch.arange(N, device=self.device) < nx)
                col = torch.where(lo[None, :], col + add, torch.where(self.neg_rows[None, :], torch.zeros_like(col), col))
                G = torch.cat([col[:, :, None], G[:, :, 1:]], dim=2)
                # transpose of conj(B chk2)
                Bg = torch.conj(G) * self.chk2
                self.ctx.scatter(Bg.contiguous(), dm=self.dm[i0:i1], X=X)
            # transpose of X = chk3 conj(rfftn(up)): up = N^3 irfftn(conj(chk3 X) / c)
            F = torch.conj(X * self.chk3) / self.cx
            up = torch.fft.irfftn(F, s=(N, N, N)) * float(N) ** 3
            return up[ip:ip + nx, ip:ip + nx, ip:ip + nx] / self.D


def cgls(proj: Projector, images, iters: int = 30, x0=None, callback=None):
    """Conjugate gradients on the normal equations A^T A x = A^T y (CGLS):
    min ||A x - y||^2.  Returns (x, history) with the residual norm and the
    seconds per iteration."""
    torch = _torch()
    with torch.no_grad():
        y = images.to(torch.float32)
        x = torch.zeros((proj.nx,) * 3, dtype=torch.float32, device=proj.device) if x0 is None else x0.clone()
        r = y - proj.apply(x) if x0 is not None else y.clone()
        s = proj.adjoint(r)
        p = s.clone()
        gamma = float(torch.sum(s.double() ** 2))
        hist = []
        for it in range(iters):
            t0 = time.perf_counter()
            q = proj.apply(p)
            qq = float(torch.sum(q.double() ** 2))
            if qq == 0.0 or gamma == 0.0:
                break
            alpha = gamma / qq
            x.add_(p, alpha=alpha)
            r.sub_(q, alpha=alpha)
            s = proj.adjoint(r)
            gnew = float(torch.sum(s.double() ** 2))
            p.mul_(gnew / gamma).add_(s)
            gamma = gnew
            torch.cuda.synchronize(proj.device)
            hist.append({"iter": it + 1, "residual": math.sqrt(float(torch.sum(r.double() ** 2))),
                         "seconds": time.perf_counter() - t0})
            if callback is not None:
                callback(it + 1, x)
        return x, hist


class DensityPreconditioner:
    """M ~ (A^T A)^-1 built from the sampling density of the projector's
    rotations: M r = D crop(irfftn(rfftn(pad(D r)) / max(W, floor))), W the
    post-SYMPLANE0 weights the insertion accumulates for those rotations
    (their value does not depend on the images), D the kernel's transform.
    A^T A is close to (1/D) crop(F^-1 rho F pad(v / D)) with rho the density,
    so one application of M to A^T y is a gridding reconstruction; M is
    symmetric positive definite (pad is crop's transpose, the multiplier is
    real, non-negative and Hermitian-symmetric).  As NRMW2 does, voxels whose
    W does not exceed `threshold` (0.1, bp3f.f:1359) carry no estimate: there
    X / W is dominated by kernel tails, and M is semi-definite, so CG works in
    the well-sampled subspace.  threshold=None keeps every voxel, with W
    clamped below at `floor` times max W (M then positive definite)."""

    def __init__(self, proj: Projector, floor: float = 1e-3, batch: int = 4096, threshold=0.1):
        torch = _torch()
        nx, n = proj.nx, proj.nproj
        with _bp.Backprojector(nx, device=proj.device.index) as b:
            zeros = torch.zeros((min(batch, n), nx, nx), dtype=torch.float32, device=proj.device)
            for i0 in range(0, n, batch):
                i1 = min(n, i0 + batch)
                b.insert_rotations(zeros[: i1 - i0], proj.dm[i0:i1])
            _, W = b.export_partial()
        W = W.to(torch.float32)
        self.inv = 1.0 / torch.clamp(W, min=float(floor) * float(W.max()))
        if threshold is not None:
            # NRMW2's rule (bp3f.f:1359): voxels with W <= threshold carry no estimate;
            # M is then positive semi-definite and CG stays in the well-sampled subspace
            self.inv = torch.where(W > float(threshold), self.inv, torch.zeros_like(self.inv))
        self.proj = proj

    def __call__(self, r):
        torch = _torch()
        P = self.proj
        p0, p1 = P.ip, P.N - P.ip - P.nx
        u = torch.nn.functional.pad(r * P.D, (p0, p1, p0, p1, p0, p1))
        g = torch.fft.irfftn(torch.fft.rfftn(u) * self.inv, s=(P.N,) * 3)
        return g[P.ip:P.ip + P.nx, P.ip:P.ip + P.nx, P.ip:P.ip + P.nx] * P.D


def pcg(proj: Projector, images, iters: int = 30, precond=None, damp: float = 0.0, x0=None, callback=None):
    """Preconditioned conjugate gradients on the (damped) normal equations
    (A^T A + damp^2 I) x = A^T y: min ||A x - y||^2 + damp^2 ||x||^2, with
    an SPD preconditioner M ~ (A^T A)^-1 (DensityPreconditioner; None: the
    identity, i.e. plain CG on the normal equations).  Returns (x, history)
    with the data residual norm and the seconds per iteration."""
    torch = _torch()
    d2 = float(damp) ** 2
    M = precond if precond is not None else (lambda r: r)

    def dot(a, b):
        return float(torch.sum(a.double() * b.double()))

    with torch.no_grad():
        y = images.to(torch.float32)
        x = torch.zeros((proj.nx,) * 3, dtype=torch.float32, device=proj.device) if x0 is None else x0.clone()
        r = proj.adjoint(y - proj.apply(x) if x0 is not None else y)
        if x0 is not None and d2:
            r.sub_(x, alpha=d2)
        z = M(r)
        p = z.clone()
        rz = dot(r, z)
        hist = []
        for it in range(iters):
            t0 = time.perf_counter()
            Ap = proj.apply(p)
            q = proj.adjoint(Ap)
            if d2:
                q.add_(p, alpha=d2)
            pq = dot(p, q)
            if pq <= 0.0 or rz == 0.0:
                break
            alpha = rz / pq
            x.add_(p, alpha=alpha)
            r.sub_(q, alpha=alpha)
            z = M(r)
            rz_new = dot(r, z)
            p.mul_(rz_new / rz).add_(z)
            rz = rz_new
            torch.cuda.synchronize(proj.device)
            hist.append({"iter": it + 1, "normal_residual": math.sqrt(max(dot(r, r), 0.0)),
                         "seconds": time.perf_counter() - t0})
            if callback is not None:
                callback(it + 1, x)
        return x, hist

