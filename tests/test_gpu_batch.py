"""GPU checks of the batched insertion (spbp_insert_batch), the plan tuner
(spbp_tune / spbp_set_plan) and captured batch graphs (spbp_graph_*).

Bars: raw accumulators of every plan and of a batch against the FP32-faithful
oracle / the one-context path: rel-L2 <= 1e-6 with identical weight support
(only the fp32 atomic order differs); graph replays against the eager batch,
every replay, with eager work and synchronisations in between.
"""
import numpy as np
import pytest

import oracle
from spider_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def rel(a, b):
    a = np.asarray(a).astype(np.complex128).ravel()
    b = np.asarray(b).astype(np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def bpmod():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    from spider_amd import bp
    return bp


def oracle_raw(imgs, ang):
    nx = imgs.shape[-1]
    tab, fltb = oracle.kb_table(2 * nx, "f32")
    X, W = oracle.empty_grid(nx)
    for p in range(imgs.shape[0]):
        dm = oracle.cang(ang[p][2], ang[p][1], ang[p][0], prec="f32")
        oracle.insert(X, W, oracle.spectrum(imgs[p], "f32"), dm, tab, fltb)
    return X, W


def dataset(nx, m, seed):
    rng = np.random.default_rng(seed)
    imgs = rng.standard_normal((m, nx, nx)).astype(np.float32)
    return imgs, synth.uniform_angles(m, rng)


def _export(ctx, half=0):
    X, W = ctx.export_raw(half)
    torch.cuda.synchronize()
    return X.cpu().numpy(), W.cpu().numpy()


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("cproj", [16, 128])
def test_every_plan_matches_oracle(bpmod, shape, cproj):
    nx, m = 20, 90
    imgs, ang = dataset(nx, m, 3 + shape)
    Xr, Wr = oracle_raw(imgs, ang)
    with bpmod.Backprojector(nx, batch=32) as b:
        b.set_plan("brick", shape, cproj)
        assert b.plan()["brick"] == [(16, 16, 8), (16, 8, 8), (8, 8, 8), (16, 12, 8), (16, 16, 10), (16, 16, 12),
                                    (16, 16, 13), (16, 16, 13)][shape]
        b.insert(torch.from_numpy(imgs).cuda(), ang)
        X, W = _export(b)
    np.testing.assert_array_equal(W != 0, Wr != 0)
    assert rel(W, Wr) <= 1e-6
    assert rel(X, Xr) <= 1e-6


def test_batch_matches_one_context_path(bpmod):
    sizes = [(16, 40, False), (24, 300, True), (32, 77, False), (21, 150, True)]
    ctxs, seq, items = [], [], []
    try:
        for k, (nx, m, two) in enumerate(sizes):
            imgs, ang = dataset(nx, m, 40 + k)
            half = bpmod.random_halves(m, seed=k) if two else None
            ti = torch.from_numpy(imgs).cuda()
            s = bpmod.Backprojector(nx, two_halves=two, batch=64)
            s.insert(ti, ang, half)
            seq.append(s)
            c = bpmod.Backprojector(nx, two_halves=two, batch=64)
            ctxs.append(c)
            items.append((c, ti, ang, half))
        bpmod.insert_batch(items)
        torch.cuda.synchronize()
        for (nx, m, two), s, c in zip(sizes, seq, ctxs):
            for h in range(2 if two else 1):
                Xs, Ws = _export(s, h)
                Xb, Wb = _export(c, h)
                np.testing.assert_array_equal(Wb != 0, Ws != 0)
                assert rel(Wb, Ws) <= 1e-6 and rel(Xb, Xs) <= 1e-6, (nx, h)
    finally:
        for c in ctxs + seq:
            c.close()


def test_tune_keeps_accumulators_and_picks_fastest(bpmod):
    nx, m = 32, 400
    imgs, ang = dataset(nx, m, 5)
    with bpmod.Backprojector(nx) as b:
        plan = b.tune(torch.from_numpy(imgs).cuda(), ang)
        X, W = _export(b)
        assert not np.any(W) and not np.any(X)
        t = plan["ms_per_proj"]
        assert np.all(t > 0) and plan["direct_ms_per_proj"] > 0
        if plan["strategy"] == "brick":
            k = [16, 32, 64, 128].index(plan["brick_proj"])
            assert t[plan["shape"], k] == t.min()
        # the tuned plan reconstructs the same accumulators as the oracle
        b.insert(torch.from_numpy(imgs[:60]).cuda(), ang[:60])
        X, W = _export(b)
    Xr, Wr = oracle_raw(imgs[:60], ang[:60])
    assert rel(W, Wr) <= 1e-6 and rel(X, Xr) <= 1e-6


def test_graph_replays_match_eager_batch(bpmod):
    sizes = [(16, 70, False), (28, 200, True), (20, 33, False)]
    eager, ctxs, items = [], [], []
    g = None
    try:
        for k, (nx, m, two) in enumerate(sizes):
            imgs, ang = dataset(nx, m, 70 + k)
            half = bpmod.random_halves(m, seed=k) if two else None
            ti = torch.from_numpy(imgs).cuda()
            if k == 1:
                ti = ti.to(torch.bfloat16)   # one job with BF16 storage
            e = bpmod.Backprojector(nx, two_halves=two, batch=64)
            eager.append(e)
            c = bpmod.Backprojector(nx, two_halves=two, batch=64)
            ctxs.append(c)
            items.append((c, ti, ang, half))
        bpmod.insert_batch([(e,) + it[1:] for e, it in zip(eager, items)])
        ref = [[_export(e, h) for h in range(2 if two else 1)] for e, (_, _, two) in zip(eager, sizes)]
        g = bpmod.BatchGraph(items, reset_first=True)
        for rep in range(3):
            g.launch()
            # the program's own eager work between replays
            junk = torch.randn(1 << 20, device="cuda")
            junk.mul_(2.0).sum().item()
            torch.cuda.synchronize()
            for c, r, (nx, m, two) in zip(ctxs, ref, sizes):
                for h in range(2 if two else 1):
                    X, W = _export(c, h)
                    np.testing.assert_array_equal(W != 0, r[h][1] != 0)
                    assert rel(W, r[h][1]) <= 1e-6 and rel(X, r[h][0]) <= 1e-6, (rep, nx, h)
    finally:
        if g is not None:
            g.close()
        for c in ctxs + eager:
            c.close()


def test_calibrations_return_rates(bpmod):
    """spbp_calibrate: positive, finite rates in a physically possible range
    (256 CUs at 2.4 GHz issue at most 6.1e11 wave-instructions per second,
    1.6e14 lane operations); undersized and oversized buffers refused."""
    import ctypes as C
    from spider_amd import _lib
    L = _lib.lib()
    r = C.c_double()
    assert L.spbp_calibrate(_lib.CAL_LDS_U64, 0, C.byref(r), None) == 0
    assert 1e9 < r.value < 6.2e11
    assert L.spbp_calibrate(_lib.CAL_GLOBAL_F32, 1 << 28, C.byref(r), None) == 0
    assert 1e9 < r.value < 1.6e14
    assert L.spbp_calibrate(_lib.CAL_GLOBAL_F32, 4096, C.byref(r), None) == _lib.ERR_ARG
    assert L.spbp_calibrate(_lib.CAL_GLOBAL_F32, (1 << 34) + 4, C.byref(r), None) == _lib.ERR_ARG


@pytest.mark.parametrize("nx,two,sym,geom_f64", [(64, True, False, False), (25, False, True, False),
                                                 (48, True, True, False), (40, True, False, True)])
def test_owner_bricks_match_shared_evaluation(bpmod, nx, two, sym, geom_f64):
    """Owner-computes bricks (plan 7) against shared-evaluation 16x16x13 bricks
    (plan 6, 16x16x10 under FP64 geometry) on larger and odd grids, two
    halves, symmetries: the same taps, so identical weight support and sums
    equal up to the fp32 order of the flush atomics."""
    m = 150
    imgs, ang = dataset(nx, m, 40 + nx)
    half = bpmod.random_halves(m, seed=nx) if two else None
    syms = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 180.0], [180.0, 0.0, 0.0]]) if sym else None
    out = []
    for shape in (6, 7):
        with bpmod.Backprojector(nx, two_halves=two, batch=64, symmetries=syms, geom_f64=geom_f64) as b:
            b.set_plan("brick", shape, 0)
            b.insert(torch.from_numpy(imgs).cuda(), ang, half)
            out.append([_export(b, h) for h in ((0, 1) if two else (0,))])
    for (Xs, Ws), (Xo, Wo) in zip(out[0], out[1]):
        np.testing.assert_array_equal(Wo != 0, Ws != 0)
        assert rel(Wo, Ws) <= 1e-6
        assert rel(Xo, Xs) <= 1e-6


def test_non_orthonormal_rotations_are_refused(bpmod):
    """The bricks enumerate each section's lattice along DM's first two rows:
    spbp_insert_dm refuses matrices whose rows 1-2 are not orthonormal."""
    from spider_amd import _lib
    nx, m = 16, 4
    imgs, ang = dataset(nx, m, 77)
    dm = _lib.rotations(ang).reshape(m, 9).copy()
    with bpmod.Backprojector(nx) as b:
        b.insert_rotations(torch.from_numpy(imgs).cuda(), dm)          # CANG matrices pass
        bad = dm.copy()
        bad[2, :3] *= 1.01
        with pytest.raises(_lib.SpbpError) as e:
            b.insert_rotations(torch.from_numpy(imgs).cuda(), bad)
        assert e.value.code == _lib.ERR_ANGLE and "image 3" in str(e.value)


def test_live_graph_refuses_calls_that_free_its_buffers(bpmod):
    """While a captured graph exists, calls that would free or replace the
    chunk buffers, brick lists or forward FFT plans it replays on fail with
    SPBP_ERR_ARG (203) and change nothing; calls that fit the buffers run,
    and the graph still replays exactly.  After spbp_graph_destroy the same
    calls succeed.  No replay follows a refused call that could have
    invalidated the graph."""
    from spider_amd import _lib
    c1 = bpmod.Backprojector(20, batch=64)     # N = 40: hipFFT transforms, tail plan of 36
    c2 = bpmod.Backprojector(16)               # N = 32: own transforms, chunk of 50
    imgs1, ang1 = dataset(20, 100, 91)
    imgs2, ang2 = dataset(16, 50, 92)
    t1, t2 = torch.from_numpy(imgs1).cuda(), torch.from_numpy(imgs2).cuda()
    big, bang = dataset(16, 200, 93)
    g = bpmod.BatchGraph([(c1, t1, ang1), (c2, t2, ang2)], reset_first=True)
    try:
        g.launch()
        ref = [_export(c1), _export(c2)]
        refused = [
            lambda: c1.set_plan("brick", 0, 0),                       # another brick shape
            lambda: c1.tune(t1[:32], ang1[:32]),                     # re-plans every shape
            lambda: c2.set_symmetries([[0, 0, 0], [0, 0, 120], [0, 0, 240]]),
            lambda: c2.insert(torch.from_numpy(big).cuda(), bang),   # larger chunk buffers
            lambda: c1.insert(t1[:50], ang1[:50]),                    # a new hipFFT tail plan
            lambda: c1.close(),
        ]
        for k, call in enumerate(refused):
            with pytest.raises(bpmod.SpbpError) as e:
                call()
            assert e.value.code == _lib.ERR_ARG, (k, str(e.value))
            assert "captured batch graph" in str(e.value), k
        c1.insert(t1[:64], ang1[:64])      # fits the chunk and plan: allowed
        c2.insert(t2[:20], ang2[:20])
        torch.cuda.synchronize()
        g.launch()
        for c, r in zip((c1, c2), ref):
            X, W = _export(c)
            np.testing.assert_array_equal(W != 0, r[1] != 0)
            assert rel(W, r[1]) <= 1e-6 and rel(X, r[0]) <= 1e-6
    finally:
        g.close()
    # released: the refused calls now run
    c1.set_plan("brick", 0, 0)
    c1.reset()
    c1.insert(t1[:50], ang1[:50])
    c2.reset()
    c2.insert(torch.from_numpy(big).cuda(), bang)
    X, W = _export(c2)
    Xr, Wr = oracle_raw(big, bang)
    assert rel(W, Wr) <= 1e-6 and rel(X, Xr) <= 1e-6
    c1.close()
    c2.close()


def test_symmetries_fixed_while_accumulators_hold_sums(bpmod):
    """One set of symmetries per operation (DMS insertion and SYMVOL,
    bp32f.f:866-872, 278-303): spbp_set_symmetries is refused between an
    insert and spbp_reset, and allowed again after the reset."""
    from spider_amd import _lib
    imgs, ang = dataset(16, 10, 95)
    with bpmod.Backprojector(16) as b:
        b.insert(torch.from_numpy(imgs).cuda(), ang)
        with pytest.raises(bpmod.SpbpError) as e:
            b.set_symmetries([[0, 0, 0], [0, 0, 180]])
        assert e.value.code == _lib.ERR_ARG
        b.reset()
        b.set_symmetries([[0, 0, 0], [0, 0, 180]])
