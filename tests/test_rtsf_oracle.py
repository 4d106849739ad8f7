"""RB 3F / RB 32F with FBS ON: the oracle's RTSF (rtsf.f:106-205) and its
interpolation FBS2 (fbs2.f:95-160), pinned against the reference's own FBS2
compiled from fbs2.f (bit for bit) and, for what the reference's FFTW-bound
FBS2_PREP and RTSF leave open, against what the mathematics fixes: exact
derivatives of trigonometric polynomials, exact cyclic shifts at zero angle,
quarter turns, and rotations of a band-limited image."""
import numpy as np
import pytest

import oracle


@pytest.fixture(scope="module")
def ref():
    r = oracle.ref()
    if r is None:
        pytest.skip("reference checker library not built (reference tree absent)")
    return r


def trig_image(nx, ny, terms):
    """sum a cos(2 pi (kx x / nx + ky y / ny) + ph) on 0-based pixels, with its
    exact derivatives d/dx, d/dy, d2/dxdy."""
    y, x = np.mgrid[0:ny, 0:nx].astype(np.float64)
    f, fx, fy, fxy = (np.zeros((ny, nx)) for _ in range(4))
    for a, kx, ky, ph in terms:
        wx, wy = 2 * np.pi * kx / nx, 2 * np.pi * ky / ny
        t = wx * x + wy * y + ph
        f += a * np.cos(t)
        fx += -a * wx * np.sin(t)
        fy += -a * wy * np.sin(t)
        fxy += -a * wx * wy * np.cos(t)
    return f, fx, fy, fxy


@pytest.mark.skipif(oracle.ref() is None, reason="reference routines not built here (oracle/_ref)")
def test_fbs2_bit_exact_vs_reference(ref):
    rng = np.random.default_rng(11)
    for ny, nx in [(9, 11), (16, 16), (13, 8)]:
        img, x1, y1, xy2 = (rng.standard_normal((ny, nx)).astype(np.float32) for _ in range(4))
        for _ in range(1500):
            x, y = np.float32(rng.uniform(-30, 40)), np.float32(rng.uniform(-30, 40))
            assert oracle.fbs2(x, y, img, x1, y1, xy2) == ref.fbs2(x, y, img, x1, y1, xy2)
        for x, y in [(1, 1), (nx, ny), (3.5, 2.25), (nx + 0.75, 0.5)]:
            assert oracle.fbs2(x, y, img, x1, y1, xy2) == ref.fbs2(x, y, img, x1, y1, xy2)


@pytest.mark.parametrize("nx,ny", [(32, 32), (31, 31), (24, 20), (17, 30)])
def test_fbs2_prep_derivatives_of_trig_polynomials(nx, ny):
    terms = [(1.0, 1, 0, 0.3), (0.5, 2, 3, -1.1), (0.25, 0, 4, 0.7), (0.8, -3, 2, 2.0)]
    f, fx, fy, fxy = trig_image(nx, ny, terms)
    x1, y1, xy2 = oracle.fbs2_prep(f.astype(np.float32))
    scale = np.abs(fx).max() + np.abs(fy).max()
    np.testing.assert_allclose(x1, fx, atol=2e-6 * scale)
    np.testing.assert_allclose(y1, fy, atol=2e-6 * scale)
    np.testing.assert_allclose(xy2, fxy, atol=2e-6 * np.abs(fxy).max())


@pytest.mark.parametrize("nx", [16, 17, 64])
def test_rtsf_integer_shift_is_cyclic_roll(nx):
    rng = np.random.default_rng(nx)
    img = rng.standard_normal((nx, nx)).astype(np.float32)
    for sx, sy in [(3, -2), (-5, 7), (0, 0), (nx + 2, -2 * nx - 1)]:
        np.testing.assert_array_equal(oracle.rtsf(img, 0.0, sx, sy), np.roll(img, (sy, sx), axis=(0, 1)))


def test_rtsf_quarter_turns_match_rtsq():
    """At multiples of 90 degrees both rotations sample (up to the fp32
    cosine of a right angle) pixel centres, so RTSF and RTSQ agree."""
    rng = np.random.default_rng(5)
    img = rng.standard_normal((32, 32)).astype(np.float32)
    for th in (90.0, 180.0, 270.0, -90.0):
        a, b = oracle.rtsf(img, th, 1.0, -2.0), oracle.rtsq(img, th, 1.0, -2.0)
        assert np.abs(a - b).max() <= 1e-4 * np.abs(img).max(), th


@pytest.mark.parametrize("theta,shx,shy", [(33.3, 2.7, -1.3), (-117.0, 0.0, 0.0), (200.5, -4.25, 3.5)])
def test_rtsf_rotates_band_limited_image(theta, shx, shy):
    """out(p) = f(R(theta) (p - c - s) + c) for the periodic band-limited f,
    c = (NX/2 + 1, NY/2 + 1) 1-based, s the shift; Hermite interpolation with
    exact derivatives leaves only its O(h^4) error."""
    nx = ny = 64
    terms = [(1.0, 1, 2, 0.3), (0.5, -2, 1, 1.0), (0.3, 3, 0, -0.4)]
    f, _, _, _ = trig_image(nx, ny, terms)
    out = oracle.rtsf(f.astype(np.float32), theta, shx, shy)
    th = np.float32(theta) * np.float32(np.pi) / np.float32(180.0)
    c, s = np.cos(np.float64(th)), np.sin(np.float64(th))
    iy, ix = np.mgrid[1:ny + 1, 1:nx + 1].astype(np.float64)
    fx = (ix - shx - (nx // 2 + 1) + nx) % nx
    fy = (iy - shy - (ny // 2 + 1) + ny) % ny
    fx = np.where(fx >= nx / 2, fx - nx, fx)     # the wrap of the shifted output grid
    fy = np.where(fy >= ny / 2, fy - ny, fy)
    xo = c * fx - s * fy + (nx // 2 + 1) - 1      # 0-based source coordinates
    yo = s * fx + c * fy + (ny // 2 + 1) - 1
    want = np.zeros_like(xo)
    for a, kx, ky, ph in terms:
        want += a * np.cos(2 * np.pi * kx * xo / nx + 2 * np.pi * ky * yo / ny + ph)
    assert np.abs(out - want).max() <= 2e-3, np.abs(out - want).max()
