"""GPU checks of the allocation error path (ERRT 46, bp32f.f:180-185).

spbp_debug_fail_alloc makes the k-th device allocation from now on fail as
if the device were out of memory.  For every k up to past the last lazy
allocation of a BP 32F job (insert + finalise) the call must return
SPBP_ERR_ALLOC, the context must stay usable, and a clean retry (reset,
insert, finalise) must match the FP32-faithful oracle.  Both forward
transform paths are covered: the library's own (N = 32) and hipFFT (N = 40).
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
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("nx", [16, 20])
def test_failed_allocation_leaves_context_usable(bpmod, nx):
    from spider_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(700 + nx)
    m = 40
    imgs = rng.standard_normal((m, nx, nx)).astype(np.float32)
    ang = synth.uniform_angles(m, rng)
    half = bpmod.random_halves(m, seed=nx)
    ref = oracle.bp(imgs, ang, half=half, prec="f32")
    ti = torch.from_numpy(imgs).cuda()
    # creation: the accumulators' allocation fails, nothing is left behind
    L.spbp_debug_fail_alloc(0)
    try:
        with pytest.raises(bpmod.SpbpError) as e:
            bpmod.Backprojector(nx, two_halves=True)
        assert e.value.code == _lib.ERR_ALLOC
    finally:
        L.spbp_debug_fail_alloc(-1)
    failures = 0
    for k in range(40):
        print(f"case nx={nx} failing allocation {k}", flush=True)
        with bpmod.Backprojector(nx, two_halves=True, batch=16) as b:
            L.spbp_debug_fail_alloc(k)
            try:
                b.insert(ti, ang, half)
                b.finalize(full=True, halves=True)
                torch.cuda.synchronize()
                ok = True
            except bpmod.SpbpError as err:
                assert err.code == _lib.ERR_ALLOC, str(err)
                ok = False
            finally:
                L.spbp_debug_fail_alloc(-1)
            failures += not ok
            b.reset()
            b.insert(ti, ang, half)
            out = b.finalize(full=True, halves=True)
            torch.cuda.synchronize()
            for key in ("full", "h1", "h2"):
                assert rel(out[key].cpu().numpy(), ref[key]) <= 1e-5, (k, key)
        if ok:   # k is past the job's last allocation
            break
    assert failures >= 10   # the job's lazy buffers were each made to fail
