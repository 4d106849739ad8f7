"""FP64 CPU restatement of the central-slice projector and its adjoint.

TEST INFRASTRUCTURE ONLY (see the package docstring): the checker for the
GPU projector's forward values and for the finite-difference test of its
gradient.  The chain is

    A v = crop( irfft2( H( conj((-1)^(I+J) G( (-1)^(i+j+k) conj(rfftn(pad(v / D))) )) ) ) )

with G the KB gather of liboracle's ``or_gather`` (the transpose of the
ONELINE restatement, oneline.f:38-158), H the Hermitian completion of the
kx = 0 column (ONELINE never samples I = 0, J < 0, oneline.f:53-55), pad /
crop at IP = (N-NX)/2 + mod(NX,2) (bp3f.f:1336-1341, bp32f.f:945-960), and
D the deapodisation volume handed in by the caller.  ``adjoint64`` is the
transpose of each step, with G^T the FP64 ONELINE scatter ``or_scatter``.
Pinned by: the adjoint identity of or_gather / or_scatter, or_scatter
against the reference-pinned ONELINE restatement, <A v, r> = <v, A^T r>,
and A on a Gaussian-blob phantom against its analytic projections
(tests/test_projector_oracle.py).
"""
from __future__ import annotations

import numpy as np

from . import gather, scatter


def _layout(nx):
    N = 2 * nx
    H = nx + 1
    ip = (N - nx) // 2 + nx % 2
    i = np.arange(N)
    h = np.arange(H)
    chk2 = np.where((i[:, None] + h[None, :]) % 2, -1.0, 1.0)
    chk3 = np.where((i[:, None, None] + i[None, :, None] + h[None, None, :]) % 2, -1.0, 1.0)
    return N, H, ip, chk2, chk3


def forward64(v, dm, D):
    """Projections (n, NX, NX) float64 of the volume v (NX^3, [z][y][x])."""
    v = np.asarray(v, np.float64)
    nx = v.shape[0]
    N, H, ip, chk2, chk3 = _layout(nx)
    up = np.zeros((N, N, N))
    up[ip:ip + nx, ip:ip + nx, ip:ip + nx] = v / np.asarray(D, np.float64)
    X = chk3 * np.conj(np.fft.rfftn(up))
    B = gather(X, dm)
    Bc = np.conj(B * chk2)
    neg = np.arange(N) > nx
    mirror = (N - np.arange(N)) % N
    Bc[:, neg, 0] = np.conj(Bc[:, mirror[neg], 0])
    img = np.fft.irfft2(Bc, s=(N, N))
    return img[:, ip:ip + nx, ip:ip + nx]


def adjoint64(r, dm, D):
    """A^T r (NX^3 float64) for projections r (n, NX, NX)."""
    r = np.asarray(r, np.float64)
    nx = r.shape[1]
    N, H, ip, chk2, chk3 = _layout(nx)
    g = np.zeros((r.shape[0], N, N))
    g[:, ip:ip + nx, ip:ip + nx] = r
    c = np.full(H, 2.0)
    c[0] = c[-1] = 1.0
    G = np.fft.rfft2(g) * c / float(N * N)
    col = G[:, :, 0].copy()
    mirror = (N - np.arange(N)) % N
    out = col.copy()
    for q in range(N):
        if 1 <= q < nx:
            out[:, q] = col[:, q] + np.conj(col[:, mirror[q]])
        elif q > nx:
            out[:, q] = 0.0
    G[:, :, 0] = out
    Bg = np.conj(G) * chk2
    X, _ = scatter(Bg, dm, nx)
    F = np.conj(X * chk3) / c
    up = np.fft.irfftn(F, s=(N, N, N), axes=(0, 1, 2)) * float(N) ** 3
    return up[ip:ip + nx, ip:ip + nx, ip:ip + nx] / np.asarray(D, np.float64)
