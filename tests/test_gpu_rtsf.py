"""GPU checks of RB 3F / RB 32F with FBS ON (RTSF pre-rotation, rtsf.f:106,
fbs2.f:95-245; bp32f.f:811-816).

Bars: RTSF on the GPU against the oracle's RTSF (itself pinned against the
reference's FBS2 and the mathematics): rel-L2 <= 1e-5 and max |diff| <=
1e-4 max|image| -- the derivative images come from the GPU's fp32 transforms
against the oracle's FP64 ones; exact at zero angle with integer shifts
(FBS2 then samples pixel centres); RB 32F with FBS against the oracle's
RTSF + BP 32F <= 1e-5; the driver's FBS mode writes the transformed
images of the Python operation (to the last bits of the FFT library each
process loads) and of the oracle."""
import os

import numpy as np
import pytest

import oracle
from spider_amd import spiderio, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def bpmod():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    from spider_amd import bp
    return bp


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("nx", [16, 17, 64, 75])
def test_rtsf_kernel_matches_oracle(bpmod, nx):
    rng = np.random.default_rng(200 + nx)
    m = 9
    imgs = rng.standard_normal((m, nx, nx)).astype(np.float32)
    prm = np.stack([rng.uniform(-400, 400, m), rng.uniform(-2 * nx, 2 * nx, m),
                    rng.uniform(-2 * nx, 2 * nx, m)], axis=1).astype(np.float32)
    for dt in (torch.float32, torch.bfloat16):
        t = torch.from_numpy(imgs).cuda().to(dt)
        out = bpmod.rtsq(t, prm, fbs=True).cpu().numpy()
        host = t.float().cpu().numpy()
        for i in range(m):
            want = oracle.rtsf(host[i], *prm[i])
            assert rel(out[i], want) <= 1e-5, (dt, i, rel(out[i], want))
            assert np.abs(out[i] - want).max() <= 1e-4 * np.abs(host[i]).max()
    # zero angle, integer shifts: pixel centres, exactly the cyclic roll
    prm0 = np.array([[0.0, 3.0, -2.0], [0.0, -5.0, 7.0], [0.0, nx + 1.0, 0.0]], np.float32)
    out = bpmod.rtsq(torch.from_numpy(imgs[:3]).cuda(), prm0, fbs=True).cpu().numpy()
    for i in range(3):
        np.testing.assert_array_equal(out[i], np.roll(imgs[i], (int(prm0[i, 2]), int(prm0[i, 1])), axis=(0, 1)))


def test_rb32f_with_fbs_matches_oracle(bpmod):
    rng = np.random.default_rng(93)
    nx, m = 24, 60
    imgs = rng.standard_normal((m, nx, nx)).astype(np.float32)
    ang = synth.uniform_angles(m, rng)
    rot = np.stack([rng.uniform(0, 360, m), rng.uniform(-3, 3, m), rng.uniform(-3, 3, m)], axis=1)
    half = bpmod.random_halves(m, seed=6)
    rimg = np.stack([oracle.rtsf(imgs[i], *rot[i].astype(np.float32)) for i in range(m)])
    ref = oracle.bp(rimg, ang, half=half, prec="f32")
    with bpmod.Backprojector(nx, two_halves=True, fbs=True, batch=16) as b:
        b.insert(torch.from_numpy(imgs).cuda(), ang, half, inplane=rot)
        out = {k: v.cpu().numpy() for k, v in b.finalize(full=True, halves=True).items()}
    for k in ("full", "h1", "h2"):
        assert rel(out[k], ref[k]) <= 1e-5, (k, rel(out[k], ref[k]))


def test_driver_fbs_mode_writes_rtsf_images(bpmod, tmp_path):
    d = str(tmp_path)
    rng = np.random.default_rng(94)
    nx, m = 20, 12
    ang = synth.uniform_angles(m, rng)
    # angles and shifts that the document's five significant digits hold exactly, so
    # that the driver and Python parse the same fp32 values
    rot = np.stack([np.round(rng.uniform(0, 360, m) * 4) / 4, np.round(rng.uniform(-2, 2, m) * 8) / 8,
                    np.round(rng.uniform(-2, 2, m) * 8) / 8], axis=1)
    imgs = rng.standard_normal((m, nx, nx)).astype(np.float32)
    for k in range(m):
        spiderio.write_image(os.path.join(d, f"img{k + 1:03d}.dat"), imgs[k])
    spiderio.write_doc(os.path.join(d, "ang.dat"),
                       {k + 1: list(ang[k]) + [k + 1, k + 1] + list(rot[k]) for k in range(m)})
    p = run_fbs(["img***", f"1-{m}", "ang", "*", "vol", "rot***"], d)
    assert p[0] == 0, p[1]
    r = spiderio.inplane_for(spiderio.read_doc(os.path.join(d, "ang.dat")), range(1, m + 1)).astype(np.float32)
    np.testing.assert_array_equal(r, rot.astype(np.float32))
    want = bpmod.rtsq(torch.from_numpy(imgs).cuda(), r, fbs=True).cpu().numpy()
    got = spiderio.read_images(os.path.join(d, "rot***"), f"1-{m}", ".dat")
    # the driver links the system's FFT library, this process the one loaded with
    # torch: the derivative images may differ in the last bits, nothing more
    assert rel(got, want) <= 1e-6 and np.abs(got - want).max() <= 1e-5 * np.abs(imgs).max()
    o = np.stack([oracle.rtsf(imgs[k], *r[k]) for k in range(m)])
    assert rel(got, o) <= 1e-5
    vd = spiderio.read_image(os.path.join(d, "vol.dat"))
    vp = bpmod.rb3f_files(os.path.join(d, "img***"), f"1-{m}", os.path.join(d, "ang.dat"), "*",
                          os.path.join(d, "pv.dat"), ext=".dat", fbs=True)
    assert rel(vd, vp) <= 1e-6


def run_fbs(answers, cwd):
    import subprocess
    from spider_amd import build as spbuild
    p = subprocess.run([spbuild.DRIVER, "RB3F", "dat", "FBS"], input="\n".join(answers) + "\n", cwd=cwd,
                       capture_output=True, text=True, timeout=300)
    return p.returncode, p.stdout + p.stderr
