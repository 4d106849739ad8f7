"""RB 3F pre-rotation in the oracle: RTSQ and QUADRI against the reference's
own routines (compiled from rtsq.f / quadri.f), and what the definition fixes."""
import numpy as np
import pytest

import oracle


@pytest.fixture(scope="module")
def ref():
    r = oracle.ref()
    if r is None:
        pytest.skip("reference checker library not built (reference tree absent)")
    return r


@pytest.mark.parametrize("nx", [12, 13, 16, 17, 64])
def test_rtsq_bit_exact(ref, nx):
    rng = np.random.default_rng(nx)
    for k in range(12):
        img = rng.standard_normal((nx, nx)).astype(np.float32)
        th, sx, sy = rng.uniform(-400, 400), rng.uniform(-2 * nx, 2 * nx), rng.uniform(-2 * nx, 2 * nx)
        if k % 4 == 0:
            th, sx, sy = np.round([th, sx, sy])
        np.testing.assert_array_equal(oracle.rtsq(img, th, sx, sy), ref.rtsq(img, th, sx, sy))


def test_quadri_bit_exact(ref):
    rng = np.random.default_rng(1)
    img = rng.standard_normal((9, 11)).astype(np.float32)
    for _ in range(4000):
        x, y = np.float32(rng.uniform(-30, 40)), np.float32(rng.uniform(-30, 40))
        assert oracle.quadri(x, y, img) == ref.quadri(x, y, img)


def test_rtsq_integer_shift_is_cyclic_roll():
    rng = np.random.default_rng(2)
    img = rng.standard_normal((16, 16)).astype(np.float32)
    for sx, sy in [(3, -2), (-5, 7), (0, 0), (17, -33)]:
        np.testing.assert_array_equal(oracle.rtsq(img, 0.0, sx, sy), np.roll(img, (sy, sx), axis=(0, 1)))


def test_quadri_interpolates_quadratics():
    """QUADRI is exact on f = a + b x + c x^2 + d y + e y^2 + g x y away
    from the wrap-around (fp32 rounding aside)."""
    rng = np.random.default_rng(3)
    a, b, c, d, e, g = rng.standard_normal(6)
    yy, xx = np.mgrid[1:21, 1:21].astype(np.float64)
    f = (a + b * xx + c * xx ** 2 + d * yy + e * yy ** 2 + g * xx * yy).astype(np.float32)
    for _ in range(200):
        x, y = rng.uniform(3, 18, 2)
        want = a + b * x + c * x * x + d * y + e * y * y + g * x * y
        assert abs(oracle.quadri(x, y, f) - want) < 2e-5 * (1 + abs(want)) * 50
