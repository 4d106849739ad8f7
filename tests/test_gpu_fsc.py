"""spider_amd.fsc on device tensors (the FFT and shell sums run on the GPU)
against the reference PR3DB shell sums in tests/golden/fsc.npz (pr3db.f:34),
and the masked form against the same computation on the CPU."""
import os

import numpy as np
import pytest

from spider_amd import fsc as F

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
GOLD = os.path.join(os.path.dirname(__file__), "golden", "fsc.npz")


@pytest.mark.parametrize("tag,wi", [("w05", 0.5), ("w10", 1.0)])
def test_device_shell_sums_match_reference(tag, wi):
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    d = np.load(GOLD)
    out = F.fsc(torch.from_numpy(d["v1"]).cuda(), torch.from_numpy(d["v2"]).cuda(), shell_width=wi)
    np.testing.assert_array_equal(out["voxels"], d[f"{tag}_lr"])
    for k in ("csum", "csum1", "csum2"):
        ref = d[f"{tag}_{k}"].astype(np.float64)
        np.testing.assert_allclose(out[k], ref, rtol=2e-5, atol=1e-3 * np.abs(ref).max())


def test_device_masked_fsc_matches_cpu():
    d = np.load(GOLD)
    g = F.fsc(torch.from_numpy(d["v1"]).cuda(), torch.from_numpy(d["v2"]).cuda(), mask_radius=8.0)
    c = F.fsc(torch.from_numpy(d["v1"]), torch.from_numpy(d["v2"]), mask_radius=8.0)
    np.testing.assert_array_equal(g["voxels"], c["voxels"])
    np.testing.assert_allclose(g["fsc"], c["fsc"], rtol=1e-6, atol=1e-6)
    assert F.fsc_resolution(g, 0.5, 1.2)["angstrom"] == pytest.approx(F.fsc_resolution(c, 0.5, 1.2)["angstrom"],
                                                                      rel=1e-5)
