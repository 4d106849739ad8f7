"""GPU checks of NRMW2's weight test decided by the reference's own sums
(spbp_reference_weights, spbp_threshold_stats).

NRMW2 divides where W > 0.1 (bp3f.f:1359); near 0.1 two summation orders can
decide differently.  The library re-sums such voxels in the reference's order
from its insertion history.  Bars:
  * the re-summed weight of every voxel equals the FP32-faithful oracle's
    serial sum bit for bit (a change in the order of the terms, a dropped or
    extra term, a wrong tap or a wrong half changes an fp32 sum), for one or
    several insert calls, both strategies, BP 32F halves, symmetries, the
    batched path and graph replays;
  * with FP64 geometry, the FP64 oracle's sums to 1e-13 relative;
  * finalisations report no re-sum that disagrees with the GPU weights, and
    after an imported partial volume the history is not used.
"""
import numpy as np
import pytest

import oracle
from spider_amd import _lib, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

C3 = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 120.0], [0.0, 0.0, 240.0]])


@pytest.fixture(scope="module")
def bpmod():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    from spider_amd import bp
    return bp


def dataset(nx, m, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((m, nx, nx)).astype(np.float32), synth.uniform_angles(m, rng)


def all_voxels(nx):
    return np.arange((nx + 1) * (2 * nx) ** 2, dtype=np.int64)


def assert_fp32_sums_equal(w, ref):
    w = np.asarray(w, np.float64)
    np.testing.assert_array_equal(w, w.astype(np.float32))      # fp32 sums, widened
    np.testing.assert_array_equal(w.astype(np.float32), np.asarray(ref, np.float32).ravel())


@pytest.mark.parametrize("strategy", ["direct", "brick"])
def test_reference_weights_bit_exact(bpmod, strategy):
    nx, m = 24, 90
    imgs, ang = dataset(nx, m, 11)
    ti = torch.from_numpy(imgs).cuda()
    with bpmod.Backprojector(nx, strategy=strategy, batch=32) as b:
        b.insert(ti[:50], ang[:50])          # two calls, several chunks each
        b.insert(ti[50:], ang[50:])
        w = b.reference_weights(all_voxels(nx))
        _, Wg = b.export_partial()
    ref = oracle.bp(imgs, ang, prec="f32", want_xw=True)["W"]
    assert_fp32_sums_equal(w, ref)
    # and they agree with the GPU's own sums up to rounding
    Wg = Wg.cpu().numpy().astype(np.float64).ravel()
    assert np.max(np.abs(w - Wg) / np.maximum(Wg, 1e-3)) < 1e-5


def test_reference_weights_halves_and_symmetries(bpmod):
    nx, m = 20, 40
    imgs, ang = dataset(nx, m, 12)
    half = bpmod.random_halves(m, seed=3)
    with bpmod.Backprojector(nx, two_halves=True, symmetries=C3) as b:
        b.insert(torch.from_numpy(imgs).cuda(), ang, half)
        v = all_voxels(nx)
        full, w0, w1 = (b.reference_weights(v, which) for which in (-1, 0, 1))
    assert_fp32_sums_equal(full, oracle.bp_sym(imgs, ang, C3, half=half, want_xw=True)["W"])
    for h, w in ((0, w0), (1, w1)):
        sel = half == h
        assert_fp32_sums_equal(w, oracle.bp_sym(imgs[sel], ang[sel], C3, want_xw=True)["W"])


def test_reference_weights_fp64_geometry(bpmod):
    nx, m = 20, 50
    imgs, ang = dataset(nx, m, 13)
    with bpmod.Backprojector(nx, geom_f64=True) as b:
        b.insert(torch.from_numpy(imgs).cuda(), ang)
        w = b.reference_weights(all_voxels(nx))
    ref = oracle.bp(imgs, ang, prec="f64", want_xw=True)["W"].ravel()
    np.testing.assert_array_equal(w != 0, ref != 0)
    nz = ref != 0
    assert np.max(np.abs(w[nz] - ref[nz]) / ref[nz]) < 1e-13


def test_history_follows_batches_graphs_and_imports(bpmod):
    nx, m = 16, 60
    imgs, ang = dataset(nx, m, 14)
    ti = torch.from_numpy(imgs).cuda()
    ref = oracle.bp(imgs, ang, prec="f32", want_xw=True)["W"]
    a = bpmod.Backprojector(nx, batch=16)
    b = bpmod.Backprojector(nx, batch=16)
    g = None
    try:
        bpmod.insert_batch([(a, ti, ang, None), (b, ti[:30], ang[:30], None)])
        torch.cuda.synchronize()
        assert_fp32_sums_equal(a.reference_weights(all_voxels(nx)), ref)
        assert_fp32_sums_equal(b.reference_weights(all_voxels(nx)),
                               oracle.bp(imgs[:30], ang[:30], prec="f32", want_xw=True)["W"])
        # a graph that zeroes first: after every replay the history is one batch
        g = bpmod.BatchGraph([(a, ti, ang, None)], reset_first=True)
        for _ in range(2):
            g.launch()
            torch.cuda.synchronize()
            assert_fp32_sums_equal(a.reference_weights(all_voxels(nx)), ref)
        vol = a.finalize()["full"].cpu().numpy()
        st = a.threshold_stats()
        assert st["inconsistent"] == 0
        o32 = oracle.bp(imgs, ang, prec="f32")["full"]
        assert np.linalg.norm(vol - o32) / np.linalg.norm(o32) < 1e-6
        # an imported partial volume: the history no longer describes the sums
        X, W = a.export_partial()
        b.import_partial(X, W)
        with pytest.raises(_lib.SpbpError):
            b.reference_weights(all_voxels(nx)[:10])
        b.finalize()
        assert b.threshold_stats()["listed"] == 0
        b.reset()                              # and a reset starts a new one
        b.insert(ti, ang)
        assert_fp32_sums_equal(b.reference_weights(all_voxels(nx)), ref)
    finally:
        if g is not None:
            g.close()
        a.close()
        b.close()


@pytest.mark.parametrize("twosum", [False, True])
def test_finalize_decides_by_reference_sums(bpmod, twosum):
    """Every listed voxel re-summed consistently; the volume at parity with the
    FP32-faithful oracle; the plain test available as an option."""
    nx, m = 48, 120
    imgs, ang = dataset(nx, m, 15)
    ti = torch.from_numpy(imgs).cuda()
    o32 = oracle.bp(imgs, ang, prec="f32")["full"]
    with bpmod.Backprojector(nx, twosum=twosum) as b:
        b.insert(ti, ang)
        vol = b.finalize()["full"].cpu().numpy()
        st = b.threshold_stats()
    assert st["listed"] > 0 and st["inconsistent"] == 0
    assert np.linalg.norm(vol - o32) / np.linalg.norm(o32) < 1e-6
    with bpmod.Backprojector(nx, twosum=twosum, plain_threshold=True) as b:
        b.insert(ti, ang)
        b.finalize()
        assert b.threshold_stats()["listed"] == 0
