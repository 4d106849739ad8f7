"""FSC (spider_amd.fsc) against the reference PR3DB shell sums (fixture written
from the reference routine pr3db.f:34 by tests/golden/make_golden.py) and
against what the definition fixes (self-correlation, sign, scaling, Parseval);
the super-Gaussian mask of pr3d.f:209-257 against its closed form; the
resolution rule of rfactsd2.f:282-325 on hand-built curves."""
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


def test_supergauss_mask_closed_form():
    """mask = exp(-0.5 (2E)^2), E = 0.5 r^2 / R^2 from the 1-based centre
    (NX/2+1, ...): exp(-r^4 / (2 R^4)); zero where E >= ln(1/TINY)."""
    for shape, R in (((24, 24, 24), 7.5), ((20, 22, 19), 4.0), ((16, 16, 16), 1.0)):
        m = F.supergauss_mask(shape, R).numpy().astype(np.float64)
        nz, ny, nx = shape
        k, j, i = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
        r2 = (i - (nx // 2 + 1)) ** 2 + (j - (ny // 2 + 1)) ** 2 + (k - (nz // 2 + 1)) ** 2
        want = np.exp(-(r2 / R ** 2) ** 2 / 2.0)
        cut = 0.5 * r2 / R ** 2 >= np.log(1.0 / np.finfo(np.float32).tiny)
        assert m[nz // 2, ny // 2, nx // 2] == 1.0
        np.testing.assert_array_equal(m[cut], 0.0)
        np.testing.assert_allclose(m[~cut], want[~cut], rtol=5e-6, atol=1e-30)
        if R == 1.0:
            assert cut.any()        # the corners of a 16^3 box lie beyond the cut


def test_masked_fsc_is_fsc_of_masked_volumes():
    rng = np.random.default_rng(12)
    a = torch.from_numpy(rng.standard_normal((20, 20, 20)).astype(np.float32))
    b = torch.from_numpy(rng.standard_normal((20, 20, 20)).astype(np.float32)) + a
    m = F.supergauss_mask(a.shape, 6.0)
    f1 = F.fsc(a, b, mask_radius=6.0)
    f2 = F.fsc(a * m, b * m)
    np.testing.assert_array_equal(f1["fsc"], f2["fsc"])
    assert not np.allclose(f1["fsc"], F.fsc(a, b)["fsc"])


def _curve(values, voxels=None):
    v = np.asarray(values, np.float64)
    return {"fsc": v, "voxels": np.full(len(v), 10) if voxels is None else np.asarray(voxels),
            "freq": np.arange(len(v)) / (2.0 * (len(v) - 1))}


def test_resolution_rule():
    """rfactsd2.f:282-325: last downward crossing at shell L >= 3, linear in
    normalised frequency (L-1)/(INC-1)/2; Angstrom = pixel size / frequency."""
    inc = 11                      # shells 1..11, frequency (L-1)/20
    c = [1.0, 0.99, 0.95, 0.9, 0.7, 0.6, 0.4, 0.3, 0.2, 0.1, 0.05]
    r = F.fsc_resolution(_curve(c), 0.5, pixel_size=2.0)
    # crossing between shell 6 (0.6, freq 0.25) and shell 7 (0.4, freq 0.30): halfway
    assert r["crossed"] and r["shell"] == pytest.approx(6.5)
    assert r["freq"] == pytest.approx(0.275) and r["angstrom"] == pytest.approx(2.0 / 0.275)
    assert F.resolution(_curve(c)) == pytest.approx(0.275)
    # a fall at shell 2 does not count (L >= 3); a later crossing does
    c2 = [1.0, 0.3, 0.8, 0.8, 0.45, 0.2, 0.1, 0.1, 0.1, 0.1, 0.1]
    assert F.fsc_resolution(_curve(c2))["shell"] == pytest.approx(4 + (0.5 - 0.8) / (0.45 - 0.8))
    # down, up, down again: the last crossing
    c3 = [1.0, 0.9, 0.4, 0.6, 0.7, 0.55, 0.3, 0.2, 0.1, 0.1, 0.1]
    r3 = F.fsc_resolution(_curve(c3))
    assert r3["shell"] == pytest.approx(6 + (0.5 - 0.55) / (0.3 - 0.55))
    # shells without voxels are skipped: the crossing pairs neighbouring shells with voxels
    vox = [10, 10, 10, 10, 0, 10, 10, 10, 10, 10, 10]
    c4 = [1.0, 0.9, 0.8, 0.7, 0.0, 0.3, 0.2, 0.1, 0.1, 0.1, 0.1]
    r4 = F.fsc_resolution(_curve(c4, vox))
    assert r4["freq"] == pytest.approx(0.15 + (0.5 - 0.7) / (0.3 - 0.7) * (0.25 - 0.15))
    # no crossing: the reference reports the last shell and its FSC value as frequency
    c5 = [1.0] * 6 + [0.9] * 5
    r5 = F.fsc_resolution(_curve(c5), pixel_size=1.5)
    assert not r5["crossed"] and r5["shell"] == inc and r5["freq"] == pytest.approx(0.9)
    assert r5["angstrom"] == pytest.approx(1.5 / 0.9)
    assert F.resolution(_curve(c5)) == 0.5


def test_fsc_doc_rows():
    c = _curve([1.0, 0.8, -0.2, 0.5], voxels=[0, 12, 30, 50])
    rows = F.fsc_doc(c, pixel_size=2.5)
    assert sorted(rows) == [2, 3, 4]                    # shells with voxels, keyed L
    f2 = 1 / 6                                          # (L-1)/(INC-1)/2 for L = 2, INC = 4
    assert rows[2][0] == pytest.approx(f2) and rows[2][1] == pytest.approx(2.5 / f2)
    assert rows[2][3] == pytest.approx(np.sqrt(0.8)) and rows[2][4] == pytest.approx(np.sqrt(1.6 / 1.8))
    assert rows[3][2] == pytest.approx(-0.2) and rows[3][3:] == [0.0, 0.0]   # clipped at zero
    assert len(F.fsc_doc(c, sqrts=False)[4]) == 3


def test_fsc_operation_on_files(tmp_path):
    """FSC on two SPIDER volume files: the document rows and the resolution
    are those of the in-memory form."""
    from spider_amd import spiderio
    d = np.load(GOLD)
    p1, p2, doc = (str(tmp_path / n) for n in ("v1.dat", "v2.dat", "fsc.dat"))
    spiderio.write_image(p1, d["v1"].astype(np.float32))
    spiderio.write_image(p2, d["v2"].astype(np.float32))
    r = F.fsc_files(p1, p2, doc, shell_width=0.5, mask_radius=9.0, pixel_size=2.0, cutoff=0.5)
    c = F.fsc(torch.from_numpy(d["v1"].astype(np.float32)), torch.from_numpy(d["v2"].astype(np.float32)),
              0.5, 9.0)
    np.testing.assert_array_equal(r["curve"]["fsc"], c["fsc"])
    assert r["angstrom"] == F.fsc_resolution(c, 0.5, 2.0)["angstrom"]
    rows = spiderio.read_doc(doc)
    want = F.fsc_doc(c, 2.0)
    assert sorted(rows) == sorted(want)
    for k in want:
        np.testing.assert_allclose(rows[k][2:], want[k][2:], rtol=1e-5, atol=1e-6)
