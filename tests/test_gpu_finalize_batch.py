"""GPU checks of the batched finalisation (spbp_finalize_batch).

Bars: every volume of a batch over contexts of different box sizes, halves,
symmetries, FP64 geometry and threshold rules is bit-identical to the same
context finalised alone (same kernels, same sums: the batch only changes how
launches and host synchronisations are grouped), the threshold counters
agree, and the batch's volumes meet the FP32-faithful oracle at rel-L2 <= 1e-6.
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


def dataset(nx, m, seed):
    rng = np.random.default_rng(seed)
    imgs = rng.standard_normal((m, nx, nx)).astype(np.float32)
    return imgs, synth.uniform_angles(m, rng)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


# (NX, projections, two halves, symmetries, FP64 geometry, plain threshold, full, halves)
CASES = [
    (24, 200, True, None, False, False, True, True),
    (33, 150, False, None, False, False, True, False),
    (32, 100, True, [[0, 0, 0], [0, 0, 180]], False, False, False, True),
    (20, 90, True, None, True, False, True, True),
    (28, 120, True, None, False, True, True, True),
    (16, 0, True, None, False, False, True, True),      # nothing inserted
]


def _contexts(bpmod):
    out = []
    for i, (nx, m, two, syms, g64, plain, full, halves) in enumerate(CASES):
        imgs, ang = dataset(nx, max(m, 1), 40 + i)
        half = bpmod.random_halves(m, seed=i) if two and m else None
        b = bpmod.Backprojector(nx, two_halves=two, symmetries=syms, geom_f64=g64, plain_threshold=plain)
        if m:
            b.insert(torch.from_numpy(imgs[:m]).cuda(), ang[:m], half)
        out.append(dict(b=b, imgs=imgs[:m], ang=ang[:m], half=half, full=full, halves=halves))
    return out


def test_batch_equals_one_by_one(bpmod):
    ctxs = _contexts(bpmod)
    try:
        alone, stats_alone = [], []
        for d in ctxs:
            alone.append({k: v.cpu().numpy() for k, v in d["b"].finalize(d["full"], d["halves"]).items()})
            stats_alone.append(d["b"].threshold_stats())
        for rep in range(2):   # a second call reuses the job table and re-summation buffers
            outs = bpmod.finalize_batch([(d["b"], d["full"], d["halves"]) for d in ctxs])
            torch.cuda.synchronize()
            for i, (d, o, a) in enumerate(zip(ctxs, outs, alone)):
                assert sorted(o) == sorted(a)
                for k in a:
                    np.testing.assert_array_equal(o[k].cpu().numpy(), a[k], err_msg=f"context {i} volume {k} call {rep}")
                assert d["b"].threshold_stats() == stats_alone[i]
        assert stats_alone[0]["listed"] > 0 and stats_alone[0]["inconsistent"] == 0
        assert stats_alone[4]["listed"] == 0          # plain threshold: no re-summation
        # against the oracle: BP 32F volumes of the first context, BP 3F of the second
        o = oracle.bp(ctxs[0]["imgs"], ctxs[0]["ang"], ctxs[0]["half"], prec="f32")
        for k in ("full", "h1", "h2"):
            assert rel(alone[0][k], o[k]) <= 1e-6, k
        o = oracle.bp(ctxs[1]["imgs"], ctxs[1]["ang"], prec="f32")
        assert rel(alone[1]["full"], o["full"]) <= 1e-6
        assert not np.any(alone[5]["full"])           # an empty context finalises to zeros
    finally:
        for d in ctxs:
            d["b"].close()


def test_batch_reuses_buffers_when_volumes_change(bpmod):
    """A batch after more insertions and with another request pattern: the
    mask bits of the previous call are cleared and the lists regrown."""
    nx = 24
    imgs, ang = dataset(nx, 400, 7)
    half = bpmod.random_halves(400, seed=3)
    a = bpmod.Backprojector(nx, two_halves=True)
    b = bpmod.Backprojector(nx, two_halves=True)
    try:
        ti = torch.from_numpy(imgs).cuda()
        a.insert(ti[:100], ang[:100], half[:100])
        b.insert(ti[:60], ang[:60], half[:60])
        bpmod.finalize_batch([a, b], full=True, halves=True)
        a.insert(ti[100:], ang[100:], half[100:])
        outs = bpmod.finalize_batch([(a, False, True), (b, True, False)])
        torch.cuda.synchronize()
        ref = oracle.bp(imgs, ang, half, prec="f32")
        for k in ("h1", "h2"):
            assert rel(outs[0][k].cpu().numpy(), ref[k]) <= 1e-6
        refb = oracle.bp(imgs[:60], ang[:60], prec="f32")
        assert rel(outs[1]["full"].cpu().numpy(), refb["full"]) <= 1e-6
    finally:
        a.close()
        b.close()


def test_batch_refusals(bpmod):
    from spider_amd import _lib
    a = bpmod.Backprojector(16, two_halves=False)
    b = bpmod.Backprojector(16, two_halves=True)
    try:
        with pytest.raises(_lib.SpbpError, match="share a context"):
            bpmod.finalize_batch([b, b])
        with pytest.raises(_lib.SpbpError):
            bpmod.finalize_batch([(b, True, True), (a, True, True)])
        arr = (_lib.FinalJob * 1)()
        assert _lib.lib().spbp_finalize_batch(arr, 0, None) == _lib.ERR_ARG
        assert _lib.lib().spbp_finalize_batch(arr, 1, None) == _lib.ERR_ARG   # no context
    finally:
        a.close()
        b.close()
