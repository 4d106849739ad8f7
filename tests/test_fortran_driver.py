"""The Fortran `BP 3F` / `BP 32F` driver (ISO_C_BINDING on libspbp): answer
sequence, SPIDER files in and out, error codes.  The CPU tests stop before
any device call; the GPU tests run the testbp3f known-answer case."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from spider_amd import build as spbuild
from spider_amd import spiderio, synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def driver():
    if not os.path.exists(spbuild.DRIVER):
        pytest.fail("Fortran driver not built (run __graft_entry__.build())")
    return spbuild.DRIVER


def run(op, answers, cwd):
    p = subprocess.run([driver(), op, "dat"], input="\n".join(answers) + "\n", cwd=cwd,
                       capture_output=True, text=True, timeout=300)
    return p.returncode, p.stdout + p.stderr


def kat_files(d, nx=75, n=12):
    img = synth.mo_t_image(nx)
    for k in range(1, n + 1):
        spiderio.write_image(os.path.join(d, f"jnk{k:03d}.dat"), img)
    shutil.copy(os.path.join(GOLDEN, "sav_bp_ang_doc.dat"), os.path.join(d, "savangbprp.dat"))


def test_driver_reports_missing_image(tmp_path):
    rc, out = run("3F", ["jnk***", "1-12", "savangbprp", "*", "out"], str(tmp_path))
    assert rc != 0 and "ERROR (18)" in out and "jnk001.dat" in out


def test_driver_rejects_non_square_and_missing_angles(tmp_path):
    spiderio.write_image(str(tmp_path / "rect001.dat"), np.zeros((10, 12), np.float32))
    shutil.copy(os.path.join(GOLDEN, "sav_bp_ang_doc.dat"), str(tmp_path / "savangbprp.dat"))
    rc, out = run("3F", ["rect***", "1", "savangbprp", "*", "out"], str(tmp_path))
    assert rc != 0 and "ERROR (101)" in out and "SQUARE" in out
    kat_files(str(tmp_path), nx=16, n=2)
    for k in (21, 22):
        spiderio.write_image(str(tmp_path / f"jnk{k:03d}.dat"), np.zeros((16, 16), np.float32))
    rc, out = run("3F", ["jnk***", "1,21", "savangbprp", "*", "out"], str(tmp_path))
    assert rc != 0 and "ERROR (102)" in out and "MISSING ANGLE" in out
    rc, out = run("3F", ["jnk***", "1-2", "savangbprp", "c1sym", "out"], str(tmp_path))
    assert rc != 0 and "ERROR (18)" in out and "c1sym" in out          # symmetries doc not found
    # RB: the angles doc has no registers 6-8, with or without a transformed-images template
    for rot in ("*", "rot***"):
        rc, out = run("RB3F", ["jnk***", "1-2", "savangbprp", "*", "out", rot], str(tmp_path))
        assert rc != 0 and "ERROR (102)" in out and "6-8" in out


def indexed_stack(path, imgs, slots, maxindx):
    """An indexed stack laid out as docs/image_doc.html describes it
    (ISTACK = -MAXINDX, index records of NX integers, then slot headers and
    images); slots maps image number -> storage slot."""
    ny, nx = next(iter(imgs.values())).shape
    labrec = -(-1024 // (4 * nx))
    labbyt = labrec * 4 * nx
    hdr = np.zeros(labbyt // 4, np.float32)
    hdr[[0, 1, 2, 4, 11, 12, 21, 22, 23, 25]] = [1, ny, labrec + ny, 1, nx, labrec, labbyt, 4 * nx, -maxindx,
                                                len(slots)]
    index = np.zeros(-(-maxindx // nx) * nx, np.int32)
    for k, sl in slots.items():
        index[k - 1] = sl
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes() + index.tobytes())
        for k, sl in sorted(slots.items(), key=lambda kv: kv[1]):
            h = np.zeros(labbyt // 4, np.float32)
            h[[0, 1, 4, 11, 12, 21, 22, 26]] = [1, ny, 1, nx, labrec, labbyt, 4 * nx, k]
            fh.write(h.tobytes() + imgs[k].astype(np.float32).tobytes())


def test_driver_reads_indexed_stacks(tmp_path):
    """Images of an indexed stack are read (the run then stops at the device
    here, or finishes where a GPU is present); an image number the index
    does not hold is error 102."""
    d = str(tmp_path)
    img = synth.mo_t_image(16)
    indexed_stack(os.path.join(d, "ind.dat"), {2: img, 7: img}, {7: 1, 2: 2}, maxindx=20)
    shutil.copy(os.path.join(GOLDEN, "sav_bp_ang_doc.dat"), os.path.join(d, "savangbprp.dat"))
    rc, out = run("3F", ["ind@**", "2,7", "savangbprp", "*", "out"], d)
    assert "NUMBER OF IMAGES" in out and "ERROR (102)" not in out and "ERROR (18)" not in out
    assert rc == 0 or "ERROR (202)" in out, out
    rc, out = run("3F", ["ind@**", "2,3", "savangbprp", "*", "out"], d)
    assert rc != 0 and "ERROR (102)" in out


@pytest.mark.gpu
def test_driver_testbp3f_known_answer(tmp_path):
    """proc/testbp3f.tst through the Fortran driver: BP 3F and BP 32F of 12
    `MO T` 75x75 images, range -0.12581E-01 ... 0.13152E-01."""
    d = str(tmp_path)
    kat_files(d)
    rc, out = run("3F", ["jnk***            ; Template", "1-12", "savangbprp", "*", "jnkbp3f_out"], d)
    assert rc == 0, out
    v = spiderio.read_image(os.path.join(d, "jnkbp3f_out.dat"))
    assert v.shape == (75, 75, 75)
    assert float("%.4E" % v.min()) == pytest.approx(-0.012581, abs=5e-7)
    assert float("%.4E" % v.max()) == pytest.approx(0.013152, abs=5e-7)
    rc, out = run("32F", ["jnk***", "1-12", "savangbprp", "*", "jnkbp32fb_out", "jnkbp32f1_out", "jnkbp32f2_out"], d)
    assert rc == 0, out
    vb = spiderio.read_image(os.path.join(d, "jnkbp32fb_out.dat"))
    assert float("%.4E" % vb.min()) == pytest.approx(-0.012581, abs=5e-7)
    assert float("%.4E" % vb.max()) == pytest.approx(0.013152, abs=5e-7)
    for name in ("jnkbp32f1_out", "jnkbp32f2_out"):
        assert spiderio.read_image(os.path.join(d, name + ".dat")).shape == (75, 75, 75)


@pytest.mark.gpu
def test_python_file_operations(tmp_path):
    """bp3f_files == BP 3F; BP 3F P on two subsets + BP 3F M == BP 3F."""
    from spider_amd import bp
    d = str(tmp_path)
    rng = np.random.default_rng(31)
    nx, m = 24, 40
    ang = synth.uniform_angles(m, rng)
    imgs = rng.standard_normal((m, nx, nx)).astype(np.float32)
    spiderio.write_stack(os.path.join(d, "stk.dat"), imgs)
    spiderio.write_doc(os.path.join(d, "ang.dat"), {k + 1: list(ang[k]) for k in range(m)})
    doc = spiderio.read_doc(os.path.join(d, "ang.dat"))
    ang_doc = spiderio.angles_for(doc, range(1, m + 1))   # registers as stored (1PGE12.5)
    v = bp.bp3f_files(os.path.join(d, "stk@***"), f"1-{m}", os.path.join(d, "ang.dat"), "*",
                      os.path.join(d, "vol.dat"), ext=".dat")
    import oracle
    ref = oracle.bp(imgs, ang_doc, prec="f32")["full"]
    assert np.linalg.norm(v - ref) / np.linalg.norm(ref) < 1e-5
    np.testing.assert_array_equal(spiderio.read_image(os.path.join(d, "vol.dat")), v)
    p1 = os.path.join(d, "part1.dat")
    p2 = os.path.join(d, "part2.dat")
    bp.bp3f_partial_files(os.path.join(d, "stk@***"), "1-17", os.path.join(d, "ang.dat"), "*", p1, ext=".dat")
    bp.bp3f_partial_files(os.path.join(d, "stk@***"), f"18-{m}", os.path.join(d, "ang.dat"), "*", p2, ext=".dat")
    vm = bp.bp3f_merge_files([p1, p2], os.path.join(d, "merged.dat"))
    assert np.linalg.norm(vm - v) / np.linalg.norm(v) < 1e-6


@pytest.mark.gpu
def test_driver_rb32f_and_symmetries(tmp_path):
    """RB 32F with registers 6-8 and a C2 symmetries doc through the Fortran
    driver, against the Python operations on the same files and the oracle."""
    from spider_amd import bp
    import oracle
    d = str(tmp_path)
    rng = np.random.default_rng(41)
    nx, m = 24, 30
    ang = synth.uniform_angles(m, rng)
    rot = np.stack([rng.uniform(0, 360, m), rng.uniform(-2, 2, m), rng.uniform(-2, 2, m)], axis=1)
    imgs = rng.standard_normal((m, nx, nx)).astype(np.float32)
    for k in range(m):
        spiderio.write_image(os.path.join(d, f"img{k + 1:03d}.dat"), imgs[k])
    spiderio.write_doc(os.path.join(d, "ang.dat"),
                       {k + 1: list(ang[k]) + [k + 1, k + 1] + list(rot[k]) for k in range(m)})
    spiderio.write_doc(os.path.join(d, "sym.dat"), {1: [0.0, 0.0, 0.0], 2: [0.0, 0.0, 180.0]})
    rc, out = run("RB32F", ["img***", f"1-{m}", "ang", "sym", "b", "e", "o", "rot@***"], d)
    assert rc == 0, out
    vd = spiderio.read_image(os.path.join(d, "b.dat"))
    vp, _, _ = bp.rb32f_files(os.path.join(d, "img***"), f"1-{m}", os.path.join(d, "ang.dat"),
                              os.path.join(d, "sym.dat"), os.path.join(d, "pb.dat"), os.path.join(d, "pe.dat"),
                              os.path.join(d, "po.dat"), ext=".dat", rot_out=os.path.join(d, "prot***"))
    assert np.linalg.norm(vd - vp) / np.linalg.norm(vp) < 1e-6
    doc = spiderio.read_doc(os.path.join(d, "ang.dat"))
    a = spiderio.angles_for(doc, range(1, m + 1))
    r = spiderio.inplane_for(doc, range(1, m + 1)).astype(np.float32)
    rimg = np.stack([oracle.rtsq(imgs[k], *r[k]) for k in range(m)])
    ref = oracle.bp_sym(rimg, a, np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 180.0]]))["full"]
    assert np.linalg.norm(vp - ref) / np.linalg.norm(ref) < 1e-5
    # the transformed images (bp32f.f:826-827): the driver's stack and the
    # Python operation's image series both hold RTSQ's output, bit for bit
    hd = spiderio.read_header(os.path.join(d, "rot.dat"))
    assert hd["istack"] > 0 and hd["maxim"] == m
    np.testing.assert_array_equal(spiderio.read_images(os.path.join(d, "rot@***"), f"1-{m}", ".dat"), rimg)
    np.testing.assert_array_equal(spiderio.read_images(os.path.join(d, "prot***"), f"1-{m}", ".dat"), rimg)
