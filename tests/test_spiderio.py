"""SPIDER file formats (docs/image_doc.html, docs/docfile.html)."""
import os

import numpy as np
import pytest

from spider_amd import spiderio

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_reads_reference_image_big_endian():
    """proc/white_font.img of the reference: big-endian 2D image, 1184 x 36."""
    path = os.path.join(GOLDEN, "white_font.img")
    h = spiderio.read_header(path)
    assert (h["byteorder"], h["nx"], h["ny"], h["nz"], h["iform"], h["labbyt"]) == (">", 1184, 36, 1, 1, 4736)
    a = spiderio.read_image(path)
    assert a.shape == (36, 1184) and a.dtype == np.float32
    u = np.unique(a)
    assert u.size == 2 and u[0] == 0.0 and u[1] == pytest.approx(3.14318, abs=1e-5)


@pytest.mark.parametrize("shape", [(75, 75), (7, 9, 11), (1, 300)])
def test_image_roundtrip_and_header_rules(tmp_path, shape):
    a = np.random.default_rng(1).standard_normal(shape).astype(np.float32)
    p = str(tmp_path / "x.dat")
    spiderio.write_image(p, a)
    np.testing.assert_array_equal(spiderio.read_image(p), a)
    h = spiderio.read_header(p)
    nx = shape[-1]
    labrec = -(-1024 // (4 * nx))        # ceil(1024 / LENBYT)
    assert h["labbyt"] == labrec * 4 * nx and h["labbyt"] >= 1024
    assert os.path.getsize(p) == h["labbyt"] + a.size * 4


def test_stack_and_templates(tmp_path):
    imgs = np.random.default_rng(2).standard_normal((5, 12, 12)).astype(np.float32)
    stk = str(tmp_path / "stk.dat")
    spiderio.write_stack(stk, imgs)
    for k in range(5):
        np.testing.assert_array_equal(spiderio.read_image(stk, k + 1), imgs[k])
    np.testing.assert_array_equal(spiderio.read_images(str(tmp_path / "stk@***"), "1-5", ".dat"), imgs)
    for k in range(5):
        spiderio.write_image(str(tmp_path / f"jnk{k + 1:03d}.dat"), imgs[k])
    np.testing.assert_array_equal(spiderio.read_images(str(tmp_path / "jnk***"), "1,2,3-5", ".dat"), imgs)
    with pytest.raises(spiderio.SpiderFormatError):
        spiderio.read_image(stk, 6)


def test_doc_files(tmp_path):
    doc = spiderio.read_doc(os.path.join(GOLDEN, "sav_bp_ang_doc.dat"))
    assert len(doc) == 20 and doc[3] == [0.0, 2.0, 71.98] and doc[20] == [0.0, 6.0, 42.341]
    old = tmp_path / "old.dat"
    old.write_text(" ;tst/tst   16-JUN-1999 AT 13:03:18   jnkdoc001.tst\n"
                   "    9 1  11.000    \n"
                   "9999996  11.000      12.000      13.000      14.000      15.000      16.000    \n"
                   "   987654 11    11.0 12.0 13.000  14.000 15.000 16.000 17.000 -1.76E-13 -19.0 10.0E2 21.2\n")
    d = spiderio.read_doc(str(old))
    assert d[9] == [11.0] and d[999999] == [11.0, 12.0, 13.0, 14.0, 15.0, 16.0]
    assert d[987654][7] == pytest.approx(-1.76e-13) and len(d[987654]) == 11
    p = str(tmp_path / "new.dat")
    spiderio.write_doc(p, {1: [1.5, -2.25, 360.0], 12: [0.1, 0.2, 0.3]})
    back = spiderio.read_doc(p)
    assert back[1] == [1.5, -2.25, 360.0] and back[12] == pytest.approx([0.1, 0.2, 0.3])
    ang = spiderio.angles_for(doc, "1-3")
    assert ang.shape == (3, 3)
    with pytest.raises(KeyError):
        spiderio.angles_for(doc, [21])


def test_partial_volume_files(tmp_path):
    rng = np.random.default_rng(3)
    X = (rng.standard_normal((16, 16, 9)) + 1j * rng.standard_normal((16, 16, 9))).astype(np.complex64)
    W = rng.uniform(0, 1, (16, 16, 9)).astype(np.float32)
    p = str(tmp_path / "part.dat")
    wp = spiderio.write_partial(p, X, W)
    assert wp.endswith("part_W.dat")
    h = spiderio.read_header(p)
    assert (h["nx"], h["ny"], h["nz"], h["iform"]) == (18, 16, 16, 3)   # 2(NX+1) x N x N reals
    X2, W2 = spiderio.read_partial(p)
    np.testing.assert_array_equal(X2, X)
    np.testing.assert_array_equal(W2, W)


def test_weights_file_names():
    """_W goes before the '@' of a stack name, else at the end of the name
    (bp3f.f:204-210, 576-583)."""
    assert spiderio._weights_name("part.dat") == "part_W.dat"
    assert spiderio._weights_name("dir.x/part") == "dir.x/part_W"
    assert spiderio._weights_name("stk@3") == "stk_W@3"
    assert spiderio._weights_name("stk.dat@12") == "stk_W.dat@12"
