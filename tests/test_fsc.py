"""FSC (spider_amd.fsc) against the reference PR3DB shell sums (fixture written
from the reference routine pr3db.f:34 by tests/golden/make_golden.py) and
against what the definition fixes (self-correlation, sign, scaling, Parseval)."""
import os

import numpy as np
import pytest
import torch

from spider_amd import fsc as F

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fsc.npz")


@pytest.mark.parametrize("tag,wi", [("w05", 0.5), ("w10", 1.0)])
def test_shell_sums_match_reference(tag, wi):
    d = np.load(GOLD)
    out = F.fsc(torch.from_numpy(d["v1"]), torch.from_numpy(d["v2"]), shell_width=wi)
    np.testing.assert_array_equal(out["voxels"], d[f"{tag}_lr"])
    for k in ("csum", "csum1", "csum2"):
        ref = d[f"{tag}_{k}"].astype(np.float64)
        # reference sums in fp32; ours in fp64
        np.testing.assert_allclose(out[k], ref, rtol=2e-5, atol=1e-3 * np.abs(ref).max())
    ref_fsc = d[f"{tag}_csum"] / np.sqrt(np.maximum(d[f"{tag}_csum1"].astype(np.float64) * d[f"{tag}_csum2"], 1e-30))
    ok = d[f"{tag}_lr"] > 0
    np.testing.assert_allclose(out["fsc"][ok], ref_fsc[ok], atol=2e-5)


def test_identities():
    rng = np.random.default_rng(5)
    v = torch.from_numpy(rng.standard_normal((20, 20, 20)))
    f = F.fsc(v, v)
    ok = f["voxels"] > 0
    np.testing.assert_allclose(f["fsc"][ok], 1.0, atol=1e-12)
    np.testing.assert_allclose(F.fsc(v, -v)["fsc"][ok], -1.0, atol=1e-12)
    np.testing.assert_allclose(F.fsc(v, 3.0 * v)["fsc"][ok], 1.0, atol=1e-12)
    # every half-space coefficient inside the outer shell is counted once
    # with its Hermitian partner: for a volume with no Nyquist components,
    # 2 * (sum of the shell sums of |a|^2) is Parseval's total over the
    # full-space coefficients inside the outer shell, DC excluded
    n = 20
    k = np.fft.fftfreq(n) * n
    K, J, I = np.meshgrid(k, k, k, indexing="ij")
    A = np.fft.fftn(v.numpy())
    A[(np.abs(I) == n // 2) | (np.abs(J) == n // 2) | (np.abs(K) == n // 2)] = 0
    w = np.real(np.fft.ifftn(A))
    A = np.fft.fftn(w)
    f = F.fsc(torch.from_numpy(w), torch.from_numpy(w))
    r = 0.5 * np.sqrt((I / (n // 2)) ** 2 + (J / (n // 2)) ** 2 + (K / (n // 2)) ** 2)
    inc = len(f["fsc"])
    inside = (r > 0) & (np.floor(r * 2 * (inc - 1) + 0.5) <= inc - 1)
    np.testing.assert_allclose(2 * f["csum1"].sum(), np.sum(np.abs(A[inside]) ** 2), rtol=1e-10)



def test_independent_noise_near_zero_and_resolution():
    rng = np.random.default_rng(9)
    n = 32
    a = torch.from_numpy(rng.standard_normal((n, n, n)))
    b = torch.from_numpy(rng.standard_normal((n, n, n)))
    f = F.fsc(a, b)
    ok = f["voxels"] > 200
    assert np.abs(f["fsc"][ok]).max() < 0.35
    # low-pass common signal plus independent noise: FSC falls through 0.5
    k = np.fft.fftfreq(n)
    K, J, I = np.meshgrid(k, k, k, indexing="ij")
    lp = np.exp(-(I ** 2 + J ** 2 + K ** 2) / (2 * 0.08 ** 2))
    s = np.real(np.fft.ifftn(np.fft.fftn(rng.standard_normal((n, n, n))) * lp)) * 40
    f2 = F.fsc(torch.from_numpy(s) + 0.3 * a, torch.from_numpy(s) + 0.3 * b)
    res = F.resolution(f2)
    assert 0.05 < res < 0.45
    assert f2["fsc"][2] > 0.9
