"""GPU checks of the library's own forward transforms (PADD2 + 2D R2C in two
passes, power-of-two N) against hipFFT on the fully padded images and against
the oracle's FP64 spectrum.

Bars: weights with identical support and within 1e-6 relative per voxel
(they do not depend on the spectra; only the order of the fp32 atomic adds
that flush overlapping brick regions differs); accumulated X within rel-L2
1e-6 of the hipFFT path (two fp32 FFTs, different rounding);
the raw spectrum of a power-of-two size within 2e-7 relative (RMS) of an FP64
FFT of the same padded image.
"""
import numpy as np
import pytest

import oracle
from spider_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def bpmod():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    from spider_amd import bp
    return bp


def rel(a, b):
    a = np.asarray(a).astype(np.complex128).ravel()
    b = np.asarray(b).astype(np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("nx", [4, 16, 32, 64, 128, 256, 25])
@pytest.mark.parametrize("bf16", [False, True])
def test_own_transforms_match_hipfft(bpmod, nx, bf16):
    rng = np.random.default_rng(nx + 7 * bf16)
    m = 40
    imgs = torch.from_numpy(rng.standard_normal((m, nx, nx)).astype(np.float32)).cuda()
    if bf16:
        imgs = imgs.to(torch.bfloat16)
    ang = synth.uniform_angles(m, rng)
    out = []
    for hipfft in (False, True):
        with bpmod.Backprojector(nx, hipfft=hipfft, batch=16) as b:
            b.insert(imgs, ang)
            X, W = b.export_raw(0)
            out.append((X.cpu().numpy(), W.cpu().numpy()))
    W0, W1 = out[0][1], out[1][1]
    np.testing.assert_array_equal(W0 != 0, W1 != 0)
    nz = W1 != 0
    assert np.max(np.abs(W0[nz] - W1[nz]) / W1[nz]) <= 1e-6
    assert rel(out[0][0], out[1][0]) <= 1e-6


@pytest.mark.parametrize("nx", [16, 64, 128, 256])
def test_own_spectrum_vs_fp64(bpmod, nx):
    """Insertion of projections whose spectra come from the own transforms and
    from hipFFT, against the same insertion (FP64 sums) of the oracle's FP64
    spectra rounded once to fp32: the own transforms are at least as close."""
    rng = np.random.default_rng(nx)
    m = 6
    imgs = rng.standard_normal((m, nx, nx)).astype(np.float32)
    ang = synth.uniform_angles(m, rng)
    tab, fltb = oracle.kb_table(2 * nx, "f32")
    X64, W64 = oracle.empty_grid(nx, "f64")
    for p in range(m):
        dm = oracle.cang(ang[p][2], ang[p][1], ang[p][0], prec="f32")
        oracle.insert_f32_acc64(X64, W64, oracle.spectrum(imgs[p], "f64").astype(np.complex64), dm, tab, fltb)
    err = {}
    for hipfft in (False, True):
        with bpmod.Backprojector(nx, hipfft=hipfft) as b:
            b.insert(torch.from_numpy(imgs).cuda(), ang)
            X, _ = b.export_raw(0)
        err[hipfft] = rel(X.cpu().numpy(), X64)
    assert err[False] <= 4e-7, err
    assert err[False] <= 1.5 * err[True] + 1e-8, err
