"""Seeded synthetic inputs for the backprojection: orientations, Gaussian-blob
phantoms with exact analytic projections, and white noise at a given SNR.

Geometry follows the reference's central-section mapping: the 2D frequency
(kx, ky) of a projection lands at kx*r1 + ky*r2 of the 3D transform, with
r1 = DM(1:3) and r2 = DM(4:6) of CANG (oneline.f:56-58, cang.f:72-80).  By the
Fourier-slice theorem the image pixel (u, v) therefore integrates the volume
along u*r1 + v*r2 + t*r3, so a 3D isotropic Gaussian A exp(-|r-c|^2 / 2 s^2)
projects to A sqrt(2 pi) s exp(-((u - r1.c)^2 + (v - r2.c)^2) / 2 s^2).
Pixel (i, j) (x = i fastest) sits at u = i - NX//2, v = j - NX//2, the same
centre the reconstruction uses (PADD2 offset, bp3f.f:1336-1341).

Default seeds: PCG64(20261202 + 1000*config + volume_index).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

SEED_BASE = 20261202


def config_seed(config: int, volume: int = 0) -> int:
    return SEED_BASE + 1000 * config + volume


def uniform_angles(n: int, rng: np.random.Generator) -> np.ndarray:
    """(n, 3) doc-file registers (PSI, THETA, PHI) in degrees:
    phi ~ U[0,360), cos(theta) ~ U[-1,1], psi ~ U[0,360)."""
    phi = rng.uniform(0.0, 360.0, n)
    theta = np.degrees(np.arccos(rng.uniform(-1.0, 1.0, n)))
    psi = rng.uniform(0.0, 360.0, n)
    return np.stack([psi, theta, phi], axis=1)


def rotations(angles) -> np.ndarray:
    """FP64 rotation rows DM(1..9) from doc registers (PSI, THETA, PHI),
    CANG(PHI, THETA, PSI) with exact pi (cang.f:72-80)."""
    a = np.radians(np.asarray(angles, np.float64).reshape(-1, 3))
    cps, sps = np.cos(a[:, 0]), np.sin(a[:, 0])
    cth, sth = np.cos(a[:, 1]), np.sin(a[:, 1])
    cph, sph = np.cos(a[:, 2]), np.sin(a[:, 2])
    dm = np.empty((a.shape[0], 9))
    dm[:, 0] = cph * cth * cps - sph * sps
    dm[:, 1] = sph * cth * cps + cph * sps
    dm[:, 2] = -sth * cps
    dm[:, 3] = -cph * cth * sps - sph * cps
    dm[:, 4] = -sph * cth * sps + cph * cps
    dm[:, 5] = sth * sps
    dm[:, 6] = sth * cph
    dm[:, 7] = sth * sph
    dm[:, 8] = cth
    return dm


@dataclass
class Phantom:
    nx: int
    centres: np.ndarray   # (K, 3) relative to the box centre, voxels (x, y, z)
    sigmas: np.ndarray    # (K,)
    amps: np.ndarray      # (K,)

    def volume(self) -> np.ndarray:
        """Point-sampled NX^3 phantom, [z][y][x] order."""
        c = np.arange(self.nx) - self.nx // 2
        z, y, x = np.meshgrid(c, c, c, indexing="ij")
        v = np.zeros((self.nx,) * 3)
        for (cx, cy, cz), s, a in zip(self.centres, self.sigmas, self.amps):
            v += a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (2 * s * s))
        return v


def blob_phantom(nx: int, k: int, radius: float, rng: np.random.Generator) -> Phantom:
    """K isotropic Gaussians, centres uniform in a ball of `radius`,
    sigma ~ U[1.5, 4] px, amplitude ~ U[0.5, 1]."""
    d = rng.standard_normal((k, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius * rng.uniform(0.0, 1.0, k) ** (1.0 / 3.0)
    return Phantom(nx, d * r[:, None], rng.uniform(1.5, 4.0, k), rng.uniform(0.5, 1.0, k))


def project(ph: Phantom, dm: np.ndarray, device="cpu", dtype=None, chunk: int = 2048):
    """Exact projections of the phantom for rotation rows `dm` (n, 9):
    torch tensor (n, NX, NX) float32, [v][u] order (x fastest)."""
    import torch
    dtype = dtype or torch.float32
    dmt = torch.as_tensor(np.asarray(dm, np.float64), device=device)
    c = torch.as_tensor(ph.centres, dtype=torch.float64, device=device)
    s = torch.as_tensor(ph.sigmas, dtype=torch.float64, device=device)
    a = torch.as_tensor(ph.amps, dtype=torch.float64, device=device) * math.sqrt(2 * math.pi) * s
    u = torch.arange(ph.nx, device=device, dtype=torch.float64) - ph.nx // 2
    out = torch.empty((dmt.shape[0], ph.nx, ph.nx), dtype=dtype, device=device)
    for i0 in range(0, dmt.shape[0], chunk):
        d = dmt[i0:i0 + chunk]
        cu = c @ d[:, 0:3].T          # (K, n): r1 . c
        cv = c @ d[:, 3:6].T          # (K, n): r2 . c
        gu = torch.exp(-(u[None, None, :] - cu.T[:, :, None]) ** 2 / (2 * s[None, :, None] ** 2))  # (n,K,NX)
        gv = torch.exp(-(u[None, None, :] - cv.T[:, :, None]) ** 2 / (2 * s[None, :, None] ** 2))
        img = torch.bmm((gv * a[None, :, None]).transpose(1, 2).contiguous(), gu.contiguous())  # (n, NX_v, NX_u)
        out[i0:i0 + chunk] = img.to(dtype)
    return out


def add_noise(imgs, snr: float, seed: int):
    """White Gaussian noise with per-image variance var(clean)/SNR (in place)."""
    import torch
    if snr is None or snr <= 0:
        return imgs
    g = torch.Generator(device=imgs.device)
    g.manual_seed(int(seed))
    var = imgs.float().reshape(imgs.shape[0], -1).var(dim=1)
    sd = torch.sqrt(var / snr).to(imgs.dtype)
    noise = torch.randn(imgs.shape, generator=g, device=imgs.device, dtype=torch.float32).to(imgs.dtype)
    imgs += noise * sd[:, None, None]
    return imgs


# configuration recipes (images NX^2, projection count, blobs, radius, SNR)
CONFIGS = {
    1: dict(nx=64, nproj=2000, blobs=30, radius=28.0, snr=0.5),
    2: dict(nx=128, nproj=20000, blobs=60, radius=58.0, snr=0.1),
    4: dict(nx=128, nproj=20000, blobs=60, radius=58.0, snr=0.1),
    5: dict(nx=256, nproj=100000, blobs=120, radius=122.0, snr=0.05),
}


def make_dataset(config: int, nproj: int | None = None, device="cpu", volume: int = 0, nx: int | None = None,
                 snr: float | None = None):
    """Angles (nproj, 3), phantom and noisy images (torch, on `device`) for a
    configuration recipe; `nproj` / `nx` override the recipe's sizes."""
    rec = dict(CONFIGS[config])
    if nx is not None:
        rec["nx"] = nx
        rec["radius"] = min(rec["radius"], nx / 2 - 6)
    n = rec["nproj"] if nproj is None else nproj
    rng = np.random.Generator(np.random.PCG64(config_seed(config, volume)))
    ph = blob_phantom(rec["nx"], rec["blobs"], rec["radius"], rng)
    ang = uniform_angles(n, rng)
    imgs = project(ph, rotations(ang), device=device)
    add_noise(imgs, rec["snr"] if snr is None else snr, config_seed(config, volume) + 7)
    return ang, ph, imgs


def mo_t_image(nx: int) -> np.ndarray:
    """SPIDER `MO` ... `T` test image (one sine wave, 2 cycles in x and y),
    evaluated in single precision as model.f:283-296 does:
    img(x, y) = sin(2 pi 2 (x-1)/NX + 2 pi 2 (y-1)/NY)."""
    f = np.float32
    twopi = f(2) * f(3.1415926535897932)
    img = np.zeros((nx, nx), np.float32)
    for i in range(1, nx + 1):
        phase = f(i - 1) * twopi / f(nx)
        j = np.arange(nx, dtype=np.float32)
        t = j * twopi * f(2) / f(nx) + phase * f(2) + f(0)
        img[i - 1] = np.sin(t).astype(np.float32)
    return img


def read_angle_doc(path: str) -> dict:
    """Minimal SPIDER document reader: {key: [registers...]} (docs/docfile.html)."""
    out = {}
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if not s or s.startswith(";"):
                continue
            parts = s.split()
            key, n = int(parts[0]), int(parts[1])
            out[key] = [float(v) for v in parts[2:2 + n]]
    return out
