"""Pins the oracle's restatement against the reference's own routines.

Fixtures in tests/golden/ were written by tests/golden/make_golden.py from the
reference's Fortran sources (oneline.f, cang.f, asta.f, fillbessi0.f,
besi1.f) compiled by oracle/Makefile.  The live-reference tests run where that
library was built and are skipped elsewhere (the fixtures still pin).
"""
import numpy as np
import pytest

import oracle


def test_cang_bitexact_vs_reference(golden):
    g = golden("cang.npz")
    for a, dm in zip(g["phi_theta_psi"], g["dm"]):
        np.testing.assert_array_equal(oracle.cang(a[0], a[1], a[2], prec="f32"), dm)


def test_kb_table_vs_reference(golden):
    g = golden("kbtab.npz")
    for n2 in (128, 150, 256):
        ref = g[f"tab{n2}"]
        fltb, ln2, alpha, aaaa, nnn = g[f"meta{n2}"]
        tab, f = oracle.kb_table(n2, "f32")
        assert np.float32(f) == np.float32(fltb) and ln2 == 2 and nnn == 3
        assert alpha == pytest.approx(6.5) and aaaa == pytest.approx(1.8 / n2, rel=1e-6)
        # the REAL evaluation (Cody's I1 as besi1.f) is the reference's bit for bit
        np.testing.assert_array_equal(tab, ref)
        assert (ref != 0).sum() == 3000      # zero beyond |d| = 1.8 voxels, index 2999
        # the FP64 evaluation differs only by single-precision rounding
        t64, _ = oracle.kb_table(n2, "f64")
        assert np.abs(t64 - ref).max() <= 1e-6


def test_library_kb_table_is_the_reference(golden):
    """libspbp's default table (host code, no GPU) equals the reference's TABI."""
    from spider_amd import _lib
    g = golden("kbtab.npz")
    for n2 in (128, 150, 256):
        tab, f = _lib.kb_table(n2)
        np.testing.assert_array_equal(tab, g[f"tab{n2}"])
        assert np.float32(f) == np.float32(g[f"meta{n2}"][0])
        t64, _ = _lib.kb_table(n2, geom_f64=True)
        np.testing.assert_array_equal(t64, oracle.kb_table(n2, "f64")[0].astype(np.float32))


def test_bessel_vs_reference(golden):
    g = golden("bessel.npz")
    for x, i1, i15, j15 in zip(g["x"], g["besi1"], g["ribsl15"], g["rjbsl15"]):
        assert oracle.bessel_i1(x) == pytest.approx(float(i1), rel=2e-6, abs=1e-7)
        assert oracle.bessel_i15(x) == pytest.approx(float(i15), rel=2e-6, abs=1e-7)
        assert oracle.lib().or_bessel_j15(float(x)) == pytest.approx(float(j15), rel=5e-6, abs=2e-7)


def test_asta_mean_vs_reference(golden):
    g = golden("asta.npz")
    for k, L in enumerate((16, 17, 75)):
        aba, dklp = g["aba_dklp"][k]
        assert oracle.asta_mean(g[f"img{L}"]) == aba / dklp


def test_oneline_bitexact_vs_reference(golden):
    """All rows of one projection, serial order: identical bits to ONELINE."""
    g = golden("oneline.npz")
    ci = 0
    while f"c{ci}_nx" in g:
        nx = int(g[f"c{ci}_nx"][0])
        X, W = oracle.empty_grid(nx)
        bi = g[f"c{ci}_bi"]
        oracle.insert(X, W, bi, g[f"c{ci}_dm"], g[f"c{ci}_tab"], g[f"c{ci}_fltb"][0])
        np.testing.assert_array_equal(W, g[f"c{ci}_W"])
        np.testing.assert_array_equal(X, g[f"c{ci}_X"])
        assert W.sum() > 0
        ci += 1
    assert ci == 6


@pytest.mark.skipif(oracle.ref() is None, reason="reference routines not built here (oracle/_ref)")
def test_oneline_bitexact_live_reference():
    R = oracle.ref()
    rng = np.random.default_rng(7)
    for trial in range(40):
        nx = int(rng.integers(9, 21))
        n = 2 * nx
        tab, fltb = R.fillbessil(n)[:2]
        bi = (rng.standard_normal((n, nx + 1)) + 1j * rng.standard_normal((n, nx + 1))).astype(np.complex64)
        ang = rng.uniform(-360, 360, 3).astype(np.float32)
        if trial < 4:   # on-axis and edge-on orientations hit the Friedel/reflection branches hardest
            ang = np.float32([[0, 0, 0], [0, 90, 0], [90, 90, 90], [0, 180, 45]][trial])
        dm = R.cang(*ang)
        X1, W1 = oracle.empty_grid(nx)
        X2, W2 = oracle.empty_grid(nx)
        for J in range(-nx + 1, nx + 1):
            R.oneline(J, X1, W1, bi, dm)
        oracle.insert(X2, W2, bi, dm, tab, fltb)
        np.testing.assert_array_equal(W1, W2)
        np.testing.assert_array_equal(X1, X2)


@pytest.mark.skipif(oracle.ref() is None, reason="reference routines not built here (oracle/_ref)")
def test_cang_and_asta_live_reference():
    R = oracle.ref()
    rng = np.random.default_rng(8)
    for a in rng.uniform(-720, 720, (500, 3)).astype(np.float32):
        np.testing.assert_array_equal(R.cang(*a), oracle.cang(*a, prec="f32"))
    for L in (8, 9, 64, 75, 128):
        img = rng.standard_normal((L, L)).astype(np.float32)
        aba, dklp = R.asta_d(img, float(L // 2))
        assert oracle.asta_mean(img) == aba / dklp


@pytest.mark.skipif(oracle.ref() is None, reason="reference routines not built here (oracle/_ref)")
def test_restatement_speed_against_compiled_oneline():
    """Calibration of the CPU baseline: the restatement's serial insertion
    against the reference's own ONELINE (flang -O2, one call per row J as
    bp32f.f:873-888 makes them) over the same spectra of 64^2 projections
    (configuration 1's shape), one thread each; the sums must be bit-identical.
    SPBP_CALIBRATE_NPROJ sets the count (default 150; the figure quoted for
    the baseline used 2000).  Prints both rates and their ratio."""
    import os
    import time
    from spider_amd import synth
    R = oracle.ref()
    nx, m = 64, int(os.environ.get("SPBP_CALIBRATE_NPROJ", "150"))
    rng = np.random.default_rng(2026)
    imgs = rng.standard_normal((m, nx, nx)).astype(np.float32)
    ang = synth.uniform_angles(m, rng).astype(np.float32)
    specs = [oracle.spectrum(im) for im in imgs]
    dms = [R.cang(*a) for a in ang]
    tab, fltb = R.fillbessil(2 * nx)[:2]
    X1, W1 = oracle.empty_grid(nx)
    X2, W2 = oracle.empty_grid(nx)
    t0 = time.perf_counter()
    for bi, dm in zip(specs, dms):
        for J in range(-nx + 1, nx + 1):
            R.oneline(J, X1, W1, bi, dm)
    t_ref = time.perf_counter() - t0
    t0 = time.perf_counter()
    for bi, dm in zip(specs, dms):
        oracle.insert(X2, W2, bi, dm, tab, fltb)
    t_port = time.perf_counter() - t0
    np.testing.assert_array_equal(W1, W2)
    np.testing.assert_array_equal(X1, X2)
    print(f"\ncalibration, {m} projections of {nx}^2: compiled ONELINE {t_ref:.2f} s ({m / t_ref:.0f} proj/s), "
          f"restatement {t_port:.2f} s ({m / t_port:.0f} proj/s), ratio {t_port / t_ref:.2f}")
