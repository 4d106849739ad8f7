"""Symmetry path of the oracle: TRIQUAD and SYMVOL against the reference's own
routines (compiled from triquad.f / symvol.f), BUILDS + MATMUL against the
definition, and what symmetrisation must do to a reconstruction."""
import numpy as np
import pytest

import oracle
from spider_amd import synth

C3 = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 120.0], [0.0, 0.0, 240.0]])
D2 = np.array([[0.0, 0.0, 0.0], [0.0, 180.0, 0.0], [180.0, 0.0, 0.0], [0.0, 180.0, 180.0]])


@pytest.fixture(scope="module")
def ref():
    r = oracle.ref()
    if r is None:
        pytest.skip("reference checker library not built (reference tree absent)")
    return r


def test_triquad_bit_exact(ref):
    rng = np.random.default_rng(0)
    for _ in range(3000):
        r, s, t = rng.random(3)
        f = (rng.standard_normal(27) * 10 ** rng.uniform(-3, 3)).astype(np.float32)
        assert ref.triquad(r, s, t, f) == oracle.triquad(r, s, t, f)


@pytest.mark.parametrize("nx", [11, 12, 16, 17])
@pytest.mark.parametrize("sym", [C3, D2], ids=["C3", "D2"])
def test_symvol_bit_exact(ref, nx, sym):
    rng = np.random.default_rng(nx)
    q = rng.standard_normal((nx, nx, nx)).astype(np.float32)
    sm = oracle.sym_matrices(sym)
    np.testing.assert_array_equal(oracle.symvol(q, sm), ref.symvol(q, sm))


def test_triquad_reproduces_quadratics():
    """A triquadratic through the 27 points is interpolated exactly."""
    rng = np.random.default_rng(1)
    c = rng.standard_normal((3, 3, 3))
    g = np.array([-1.0, 0.0, 1.0])

    def poly(x, y, z):
        return sum(c[i, j, k] * x ** i * y ** j * z ** k for i in range(3) for j in range(3) for k in range(3))

    # F(L): x fastest (-1, 0, 1), then y, then z (symvol.f:75-86)
    F = np.array([poly(g[l % 3], g[(l // 3) % 3], g[l // 9]) for l in range(27)], np.float32)
    for _ in range(50):
        r, s, t = rng.random(3)
        assert abs(oracle.triquad(r, s, t, F) - poly(r, s, t)) < 1e-5 * (1 + abs(poly(r, s, t)))


def test_builds_and_matmul():
    rng = np.random.default_rng(2)
    sm = oracle.sym_matrices(C3)
    # CANG of (0, 0, 120): rotation by 120 degrees about z (cang.f:72-80)
    m = sm[1].reshape(3, 3, order="F")
    np.testing.assert_allclose(m @ m @ m, np.eye(3), atol=1e-6)
    np.testing.assert_allclose(np.diag(m), [-0.5, -0.5, 1.0], atol=1e-6)
    for _ in range(20):
        a = synth.uniform_angles(1, rng)[0]
        dm = oracle.cang(a[2], a[1], a[0], prec="f32")
        out = oracle.sym_matmul(sm[1], dm).reshape(3, 3, order="F")
        want = sm[1].reshape(3, 3, order="F").astype(np.float64) @ dm.reshape(3, 3, order="F").astype(np.float64)
        np.testing.assert_allclose(out, want, atol=2e-7)
        # the order matters: DM SM is a different matrix
        assert np.abs(out - dm.reshape(3, 3, order="F") @ m).max() > 1e-3


def test_symmetrised_reconstruction_is_symmetric():
    """BP 3F with a C3 document: the volume is invariant under the 120 degree
    rotation (up to the interpolation of the check itself), and a one-row
    document (the identity) is the plain reconstruction."""
    rng = np.random.default_rng(3)
    nx, m = 16, 60
    imgs = rng.standard_normal((m, nx, nx)).astype(np.float32)
    ang = synth.uniform_angles(m, rng)
    v = oracle.bp_sym(imgs, ang, C3)["full"]
    sm = oracle.sym_matrices(C3)
    # SYMVOL with the same group leaves a symmetric volume nearly unchanged
    v2 = oracle.symvol(v, sm)
    inner = np.zeros_like(v, bool)
    c = np.arange(nx) - nx // 2
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    inner = x * x + y * y + z * z <= (nx // 2 - 3) ** 2
    rel = np.linalg.norm((v2 - v)[inner]) / np.linalg.norm(v[inner])
    assert rel < 0.15, rel
    # without symmetries the same check fails badly
    p = oracle.bp(imgs, ang)["full"]
    rel_p = np.linalg.norm((oracle.symvol(p, sm) - p)[inner]) / np.linalg.norm(p[inner])
    assert rel_p > 2 * rel
    # one-row document = no symmetries
    np.testing.assert_array_equal(oracle.bp_sym(imgs, ang, C3[:1])["full"], p)
