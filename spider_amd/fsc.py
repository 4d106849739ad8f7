"""Fourier shell correlation between two volumes, as SPIDER's `FSC`
operation defines it (PR3D / PR3DB / RFACTSD2: pr3d.f:44, pr3db.f:34,
rfactsd2.f:245).

Half-space coefficients (a, b) of the two transforms are binned by the
normalised radius R = 0.5 sqrt((i/(NX/2))^2 + (j/(NY/2))^2 + (k/(NZ/2))^2)
into shell L = NINT(R 2 (INC-1)) + 1, INC = INT(max(NX,NY,NZ)/WI)/2 + 1 for
shell width WI (pr3d.f:194-195); the kx = 0 plane counts each Hermitian pair
once (pr3db.f:63-64).  FSC(L) = sum Re(a conj b) / sqrt(sum|a|^2 sum|b|^2).
Runs on whatever device the volumes are on (torch FFT).
"""
from __future__ import annotations

import numpy as np


def fsc(v1, v2, shell_width: float = 0.5):
    """Returns dict: freq (normalised, 0.5 = Nyquist), fsc, voxels, csum,
    csum1, csum2 per shell L = 1..INC (numpy arrays)."""
    import torch
    a = torch.as_tensor(v1)
    b = torch.as_tensor(v2, device=a.device)
    if a.shape != b.shape or a.dim() != 3:
        raise ValueError("FSC needs two volumes of the same size")
    nz, ny, nx = a.shape
    A = torch.fft.rfftn(a.to(torch.float64))
    B = torch.fft.rfftn(b.to(torch.float64))
    inc = int(max(nx, ny, nz) / shell_width) // 2 + 1
    dev = a.device
    iii = torch.arange(nx // 2 + 1, device=dev, dtype=torch.float64)
    iij = torch.arange(ny, device=dev, dtype=torch.float64)
    iij = torch.where(iij > ny // 2, iij - ny, iij)
    iik = torch.arange(nz, device=dev, dtype=torch.float64)
    iik = torch.where(iik > max(1, nz // 2), iik - nz, iik)
    K, J, I = torch.meshgrid(iik, iij, iii, indexing="ij")
    r = 0.5 * torch.sqrt((I / (nx // 2)) ** 2 + (J / (ny // 2)) ** 2 + (K / max(1, nz // 2)) ** 2)
    keep = (I > 0) | ((K >= 0) & ((J >= 0) | (K != 0)))
    keep &= r > 0
    L = torch.floor(r * 2 * (inc - 1) + 0.5).to(torch.int64)      # NINT of a non-negative value
    keep &= L <= inc - 1
    L = L[keep]
    a_ = A[keep]
    b_ = B[keep]
    csum = torch.bincount(L, weights=(a_.real * b_.real + a_.imag * b_.imag), minlength=inc)
    csum1 = torch.bincount(L, weights=a_.abs() ** 2, minlength=inc)
    csum2 = torch.bincount(L, weights=b_.abs() ** 2, minlength=inc)
    vox = 2 * torch.bincount(L, minlength=inc)
    den = torch.sqrt(csum1 * csum2)
    f = torch.where(den > 0, csum / torch.where(den > 0, den, torch.ones_like(den)), torch.zeros_like(den))
    out = {"freq": np.arange(inc) / (2.0 * (inc - 1)), "fsc": f.cpu().numpy(), "voxels": vox.cpu().numpy(),
           "csum": csum.cpu().numpy(), "csum1": csum1.cpu().numpy(), "csum2": csum2.cpu().numpy()}
    return out


def resolution(curve: dict, cutoff: float = 0.5) -> float:
    """Normalised frequency where the FSC first falls below `cutoff`
    (linear interpolation); 0.5 when it never does."""
    f, c = curve["freq"][1:], curve["fsc"][1:]
    for k in range(1, len(c)):
        if c[k] < cutoff <= c[k - 1]:
            t = (c[k - 1] - cutoff) / (c[k - 1] - c[k])
            return float(f[k - 1] + t * (f[k] - f[k - 1]))
    return 0.5
