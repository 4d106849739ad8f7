"""GPU checks of the voxel-driven gather plan (SPBP_STRAT_GATHER).

Bars: the gather adds each voxel's terms in the serial reference's order, so
its weights are the FP32-faithful oracle's serial sums bit for bit (one or
two halves, across chunks and calls) and X is too when both sides scatter
the same spectra; from images, X differs from the oracle's only through the
spectra (the oracle's own FFT against the library's fp32 one) and is held
to rel-L2 <= 1e-6; with symmetries the
weights equal the library's own reference-order re-summation bit for bit
and the volumes meet the oracle's; the adjoint scatter of given spectra,
with and without weights, is the serial reference's bit for bit; the
FP64-geometry form agrees with the direct scatter to fp32 rounding; two
runs are bit-identical; in a batch and a replayed graph beside a brick
context it gives the same bits as alone; the tuner times it.
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


def oracle_raw(imgs, ang, sel=None):
    """Serial FP32 insertion in input order (the reference's order)."""
    nx = imgs.shape[-1]
    tab, fltb = oracle.kb_table(2 * nx, "f32")
    X, W = oracle.empty_grid(nx)
    for p in range(imgs.shape[0]):
        if sel is not None and not sel[p]:
            continue
        dm = oracle.cang(ang[p][2], ang[p][1], ang[p][0], prec="f32")
        oracle.insert(X, W, oracle.spectrum(imgs[p], "f32"), dm, tab, fltb)
    return X, W


def export(b, half=0):
    X, W = b.export_raw(half)
    torch.cuda.synchronize()
    return X.cpu().numpy(), W.cpu().numpy()


def rel(a, b):
    a = np.asarray(a).astype(np.complex128).ravel()
    b = np.asarray(b).astype(np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("nx,m,batch", [(20, 90, 32), (25, 60, 0), (32, 40, 16)])
def test_gather_is_the_serial_reference_bit_for_bit(bpmod, nx, m, batch):
    imgs, ang = dataset(nx, m, 300 + nx)
    ti = torch.from_numpy(imgs).cuda()
    Xr, Wr = oracle_raw(imgs, ang)
    with bpmod.Backprojector(nx, strategy="gather", batch=batch) as b:
        assert b.plan()["strategy"] == "gather"
        b.insert(ti[: m // 3], ang[: m // 3])     # two calls, several chunks each
        b.insert(ti[m // 3:], ang[m // 3:])
        X, W = export(b)
        b.reset()
        b.insert(ti, ang)
        X2, W2 = export(b)
    np.testing.assert_array_equal(W, Wr)
    assert rel(X, Xr) <= 1e-6                     # the spectra's FFTs differ, not the sums
    np.testing.assert_array_equal(W2, W)          # run to run: the same bits
    np.testing.assert_array_equal(X2, X)


def test_gather_halves_and_compensated_weights(bpmod):
    """Per half the fp32 sums are the serial reference's; the compensation
    terms ride along, and the finalised volumes match the brick plan's."""
    nx, m = 24, 80
    imgs, ang = dataset(nx, m, 41)
    half = bpmod.random_halves(m, seed=5)
    ti = torch.from_numpy(imgs).cuda()
    vols = {}
    for strat in ("gather", "brick"):
        with bpmod.Backprojector(nx, strategy=strat, two_halves=True, twosum=True, batch=24) as b:
            b.insert(ti, ang, half)
            if strat == "gather":
                for h in (0, 1):
                    X, W = export(b, h)
                    Xr, Wr = oracle_raw(imgs, ang, sel=(half == h))
                    np.testing.assert_array_equal(W, Wr)
                    assert rel(X, Xr) <= 1e-6
            vols[strat] = {k: v.cpu().numpy() for k, v in b.finalize(full=True, halves=True).items()}
    for k in ("full", "h1", "h2"):
        assert rel(vols["gather"][k], vols["brick"][k]) <= 1e-6, k


def test_gather_with_symmetries_matches_reference_resummation(bpmod):
    nx, m = 20, 30
    imgs, ang = dataset(nx, m, 52)
    half = bpmod.random_halves(m, seed=2)
    syms = [[0, 0, 0], [0, 0, 120], [0, 0, 240]]
    ti = torch.from_numpy(imgs).cuda()
    n = 2 * nx
    kz, ky, kx = np.meshgrid(np.arange(n), np.arange(n), np.arange(1, nx + 1), indexing="ij")
    vox = ((kz * n + ky) * (nx + 1) + kx).ravel()        # kx > 0: no SYMPLANE0 partner
    with bpmod.Backprojector(nx, strategy="gather", two_halves=True, symmetries=syms, batch=16) as b:
        b.insert(ti, ang, half)
        for h in (0, 1):
            _, W = export(b, h)
            ref = b.reference_weights(vox, which=h)
            np.testing.assert_array_equal(W.ravel()[vox], ref.astype(np.float32))
        g = {k: v.cpu().numpy() for k, v in b.finalize(full=True, halves=True).items()}
    o = oracle.bp_sym(imgs, ang, syms, half)
    for k in ("full", "h1", "h2"):
        assert rel(g[k], o[k]) <= 1e-6, k


@pytest.mark.parametrize("geom_f64", [False, True])
def test_gather_agrees_with_direct_scatter(bpmod, geom_f64):
    nx, m = 28, 50
    imgs, ang = dataset(nx, m, 61)
    ti = torch.from_numpy(imgs).cuda()
    out = {}
    for strat in ("gather", "direct"):
        with bpmod.Backprojector(nx, strategy=strat, geom_f64=geom_f64) as b:
            b.insert(ti, ang)
            out[strat] = export(b)
    (Xg, Wg), (Xd, Wd) = out["gather"], out["direct"]
    np.testing.assert_array_equal(Wg != 0, Wd != 0)
    assert rel(Wg, Wd) <= 1e-6 and rel(Xg, Xd) <= 1e-6


def test_gather_in_a_batch_and_a_graph(bpmod):
    nx1, nx2, m = 20, 24, 70
    i1, a1 = dataset(nx1, m, 71)
    i2, a2 = dataset(nx2, m, 72)
    t1, t2 = torch.from_numpy(i1).cuda(), torch.from_numpy(i2).cuda()
    with bpmod.Backprojector(nx1, strategy="gather") as alone:
        alone.insert(t1, a1)
        Xa, Wa = export(alone)
    g = bpmod.Backprojector(nx1, strategy="gather", batch=32)
    k = bpmod.Backprojector(nx2)
    graph = None
    try:
        bpmod.insert_batch([(g, t1, a1), (k, t2, a2)])
        torch.cuda.synchronize()
        X, W = export(g)
        np.testing.assert_array_equal(W, Wa)
        np.testing.assert_array_equal(X, Xa)
        graph = bpmod.BatchGraph([(g, t1, a1), (k, t2, a2)], reset_first=True)
        for _ in range(2):
            graph.launch()
            torch.cuda.synchronize()
            X, W = export(g)
            np.testing.assert_array_equal(W, Wa)
            np.testing.assert_array_equal(X, Xa)
    finally:
        if graph is not None:
            graph.close()
        g.close()
        k.close()


def test_tuner_times_the_gather_plan(bpmod):
    nx, m = 24, 200
    imgs, ang = dataset(nx, m, 81)
    with bpmod.Backprojector(nx) as b:
        plan = b.tune(torch.from_numpy(imgs).cuda(), ang)
        assert plan["gather_ms_per_proj"] > 0 and plan["direct_ms_per_proj"] > 0
        best = min(plan["ms_per_proj"].min(), plan["direct_ms_per_proj"], plan["gather_ms_per_proj"])
        chosen = {"brick": plan["ms_per_proj"][plan["shape"], [16, 32, 64, 128].index(plan["brick_proj"])],
                  "direct": plan["direct_ms_per_proj"], "gather": plan["gather_ms_per_proj"]}[plan["strategy"]]
        assert chosen == best


def test_gather_scatter_of_spectra_with_and_without_weights(bpmod):
    """spbp_scatter (the projector's adjoint, weights optional) on a gather
    context: the serial reference's sums of the given BI spectra, bit for bit."""
    nx, m = 22, 40
    rng = np.random.default_rng(91)
    n = 2 * nx
    B = (rng.standard_normal((m, n, nx + 1)) + 1j * rng.standard_normal((m, n, nx + 1))).astype(np.complex64)
    ang = synth.uniform_angles(m, rng)
    tab, fltb = oracle.kb_table(n, "f32")
    Xr, Wr = oracle.empty_grid(nx)
    for p in range(m):
        oracle.insert(Xr, Wr, B[p], oracle.cang(ang[p][2], ang[p][1], ang[p][0], prec="f32"), tab, fltb)
    Bt = torch.from_numpy(B).cuda()
    with bpmod.Backprojector(nx, strategy="gather", batch=16) as b:
        X = b.scatter(Bt, ang)                       # no weights
        W = torch.zeros(b.grid_shape, dtype=torch.float32, device="cuda")
        X2 = b.scatter(Bt, ang, W=W)
        torch.cuda.synchronize()
    np.testing.assert_array_equal(X.cpu().numpy(), Xr)
    np.testing.assert_array_equal(X2.cpu().numpy(), Xr)
    np.testing.assert_array_equal(W.cpu().numpy(), Wr)
