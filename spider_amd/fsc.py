"""Fourier shell correlation between two volumes, as SPIDER's `FSC`
operation defines it (PR3D / PR3DB / RFACTSD2: pr3d.f:44, pr3db.f:34,
rfactsd2.f:45).

* Masking (pr3d.f:209-257): with a masking radius R > 0 both volumes are
  multiplied by the super-Gaussian exp(-0.5 (2E)^2), E = 0.5 r^2 / R^2, r the
  distance from the centre voxel (NX/2+1, NY/2+1, NZ/2+1) (1-based), and set to
  zero where E >= ln(1/TINY) -- in REAL arithmetic, as the reference.
* Shells (pr3db.f:34-102): half-space coefficients (a, b) of the two
  transforms are binned by the normalised radius R = 0.5 sqrt((i/(NX/2))^2 +
  (j/(NY/2))^2 + (k/(NZ/2))^2) into shell L = NINT(R 2 (INC-1)) + 1, INC =
  INT(max(NX,NY,NZ)/WI)/2 + 1 for shell width WI (pr3d.f:194-195); the kx = 0
  plane counts each Hermitian pair once (pr3db.f:63-64).
  FSC(L) = sum Re(a conj b) / sqrt(sum|a|^2 sum|b|^2) (rfactsd2.f:240-246).
* Resolution (rfactsd2.f:282-325): the last downward crossing of the cutoff
  at shell L >= 3, linearly interpolated in normalised frequency; in
  Angstrom, pixel size / frequency (FMAXSPFREQ = 0.5 / pixel size).
Runs on whatever device the volumes are on (torch FFT, FP64 sums).
"""
from __future__ import annotations

import numpy as np

# ln(1 / TINY(1.0)) in REAL, the reference's cut of the mask's exponent
# (pr3d.f:216: TNM = ALOG(1.0 / TINY(TNM)))
_TNM = np.float32(np.log(np.float32(1.0) / np.finfo(np.float32).tiny))


def supergauss_mask(shape, radius: float, device=None):
    """The masking volume of `FSC` for masking radius `radius` (pr3d.f:213-248),
    float32 [z][y][x], evaluated operation by operation in fp32."""
    import torch
    nz, ny, nx = shape
    f32 = torch.float32
    xc, yc, zc = float(nx // 2 + 1), float(ny // 2 + 1), float(nz // 2 + 1)
    ri = torch.tensor(1.0, dtype=f32) / torch.tensor(radius, dtype=f32) ** 2   # RADMASKSQI
    ri = ri.to(device)
    i = torch.arange(1, nx + 1, device=device, dtype=f32)
    j = torch.arange(1, ny + 1, device=device, dtype=f32)
    k = torch.arange(1, nz + 1, device=device, dtype=f32)
    dx = (i - xc) ** 2 * ri
    dy = (j - yc) ** 2 * ri
    dz = (k - zc) ** 2 * ri
    # EEE = 0.5 * ((I-XCEN)**2 * RI + (J-YCEN)**2 * RI + (K-ZCEN)**2 * RI)
    e = 0.5 * ((dx[None, None, :] + dy[None, :, None]) + dz[:, None, None])
    m = torch.exp(-0.5 * (2 * e) ** 2)
    return torch.where(e >= float(_TNM), torch.zeros_like(m), m)


def fsc(v1, v2, shell_width: float = 0.5, mask_radius: float = 0.0):
    """Returns dict: freq (normalised, 0.5 = Nyquist), fsc, voxels, csum,
    csum1, csum2 per shell L = 1..INC (numpy arrays).  mask_radius > 0 masks
    both volumes first (the `FSC` operation's MASKING RADIUS)."""
    import torch
    a = torch.as_tensor(v1)
    b = torch.as_tensor(v2, device=a.device)
    if a.shape != b.shape or a.dim() != 3:
        raise ValueError("FSC needs two volumes of the same size")
    nz, ny, nx = a.shape
    if mask_radius > 0:
        m = supergauss_mask(a.shape, mask_radius, a.device)
        a = a.to(torch.float32) * m
        b = b.to(torch.float32) * m
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


def fsc_resolution(curve: dict, cutoff: float = 0.5, pixel_size: float = 1.0) -> dict:
    """The `FSC` operation's resolution (rfactsd2.f:282-325): over the shells
    with voxels, the last one L >= 3 where the curve falls below `cutoff`
    from a value >= cutoff at the shell before; the normalised frequency is
    interpolated linearly between the two shells and the resolution is
    pixel_size / frequency.  Without a crossing the reference reports the
    last shell and takes that shell's FSC value as the frequency
    (rfactsd2.f:317-318); that is returned as it is, with crossed = False.
    Returns shell (1-based, fractional), freq, angstrom, crossed."""
    lr = np.asarray(curve["voxels"])
    fscv = np.asarray(curve["fsc"], np.float64)
    inc = len(fscv)
    xprev = 0
    fsc_last = np.inf
    spf_last = np.inf
    shell_last, fsc_cur = 0, 0.0
    prev = now = None
    for L in range(1, inc + 1):
        if lr[L - 1] == 0:
            continue
        spf = (L - 1) / (inc - 1) * 0.5
        f = fscv[L - 1]
        if L >= 3 and fsc_last >= cutoff and f < cutoff:
            xprev = L - 1
            prev, now = (spf_last, fsc_last), (spf, f)
        fsc_last, spf_last = f, spf
        shell_last, fsc_cur = L, f
    if xprev > 0:
        t = (cutoff - prev[1]) / (now[1] - prev[1])
        freq = prev[0] + t * (now[0] - prev[0])
        shell = xprev + t
        crossed = True
    else:
        shell, freq, crossed = float(shell_last), float(fsc_cur), False
    ang = pixel_size / freq if freq != 0 else float("inf")
    return {"shell": float(shell), "freq": float(freq), "angstrom": float(ang), "crossed": crossed}


def resolution(curve: dict, cutoff: float = 0.5) -> float:
    """Normalised frequency where the FSC falls below `cutoff` by the `FSC`
    operation's rule (fsc_resolution); 0.5 when it never does."""
    r = fsc_resolution(curve, cutoff)
    return r["freq"] if r["crossed"] else 0.5


def fsc_doc(curve: dict, pixel_size: float = 1.0, sqrts: bool = True) -> dict:
    """The `FSC` document's rows (rfactsd2.f:180-270): key L for each shell
    with voxels -> [NORM-FREQ, ANGST, FSC] and, with sqrts (the operation's
    default), |SQRT(FSC)| and |SQRT(2 FSC / (FSC + 1))| of the FSC clipped at
    zero.  Write it with spiderio.write_doc."""
    lr = np.asarray(curve["voxels"])
    fscv = np.asarray(curve["fsc"], np.float64)
    inc = len(fscv)
    fmax = 0.5 / pixel_size
    rows = {}
    for L in range(1, inc + 1):
        if lr[L - 1] == 0:
            continue
        spf = (L - 1) / (inc - 1) * 0.5
        row = [spf, 0.5 / (fmax * spf) if spf > 0 else float("inf"), fscv[L - 1]]
        if sqrts:
            z = max(fscv[L - 1], 0.0)
            row += [0.0, 0.0] if z == 0.0 else [float(np.sqrt(z)), float(np.sqrt(2 * z / (z + 1)))]
        rows[L] = row
    return rows


def fsc_files(vol1: str, vol2: str, doc_out: str | None = None, shell_width: float = 0.5,
              mask_radius: float = 0.0, pixel_size: float = 1.0, cutoff: float = 0.5, device=None) -> dict:
    """The `FSC` operation on SPIDER volume files (pr3d.f:44): the two
    volumes, shell width and masking radius, voxel size and resolution
    cutoff; writes the FSC document (fsc_doc) when doc_out is given and
    returns fsc_resolution's dict with the curve under 'curve'."""
    import torch
    from . import spiderio
    a = torch.from_numpy(np.ascontiguousarray(spiderio.read_image(vol1), np.float32))
    b = torch.from_numpy(np.ascontiguousarray(spiderio.read_image(vol2), np.float32))
    if a.dim() != 3 or b.dim() != 3:
        raise ValueError("OPERATION ONLY FOR VOLUMES")
    if device is not None:
        a, b = a.to(device), b.to(device)
    curve = fsc(a, b, shell_width, mask_radius)
    if doc_out:
        spiderio.write_doc(doc_out, fsc_doc(curve, pixel_size))
    out = fsc_resolution(curve, cutoff, pixel_size)
    out["curve"] = curve
    return out
