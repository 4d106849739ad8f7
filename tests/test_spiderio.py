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


@pytest.mark.parametrize("byteorder", ["<", ">"])
def test_indexed_stack(tmp_path, byteorder):
    """An indexed stack built byte by byte from the layout in
    docs/image_doc.html: overall header with ISTACK = -MAXINDX, then
    INDXREC = ceil(MAXINDX / NX) records of NX 32-bit integers (image number
    -> storage slot, 0 = absent; LUNREDINDX, lunsethdr.f), then per slot a
    LABREC-record header and the image.  Images are stored out of order."""
    nx, ny, maxindx = 6, 5, 15                # INDXREC = 3 records of 6 entries
    labrec = -(-1024 // (4 * nx))             # 43 records of 24 bytes
    labbyt, lenbyt = labrec * 4 * nx, 4 * nx
    indxrec = -(-maxindx // nx)
    rng = np.random.default_rng(4)
    imgs = {3: rng.standard_normal((ny, nx)), 11: rng.standard_normal((ny, nx)), 14: rng.standard_normal((ny, nx))}
    slots = {14: 1, 3: 2, 11: 3}              # image number -> slot
    f4, i4 = np.dtype(byteorder + "f4"), np.dtype(byteorder + "i4")
    hdr = np.zeros(labbyt // 4, f4)
    hdr[0], hdr[1], hdr[2], hdr[4] = 1, ny, labrec + ny, 1            # NZ, NY, IREC, IFORM
    hdr[11], hdr[12], hdr[21], hdr[22] = nx, labrec, labbyt, lenbyt   # NX, LABREC, LABBYT, LENBYT
    hdr[23], hdr[25], hdr[27] = -maxindx, 3, 14                       # ISTACK, MAXIM, LASTINDX
    index = np.zeros(indxrec * nx, i4)
    for k, s in slots.items():
        index[k - 1] = s
    body = b""
    for k, s in sorted(slots.items(), key=lambda kv: kv[1]):
        h = np.zeros(labbyt // 4, f4)
        h[0], h[1], h[4], h[11], h[12], h[21], h[22], h[26] = 1, ny, 1, nx, labrec, labbyt, lenbyt, k
        body += h.tobytes() + imgs[k].astype(f4).tobytes()
    path = str(tmp_path / "ind.dat")
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes() + index.tobytes() + body)
    for k in slots:
        np.testing.assert_array_equal(spiderio.read_image(path, k), imgs[k].astype(np.float32))
    got = spiderio.read_images(str(tmp_path / "ind@**"), "14,3,11", ext=".dat")
    np.testing.assert_array_equal(got, np.stack([imgs[k] for k in (14, 3, 11)]).astype(np.float32))
    with pytest.raises(spiderio.SpiderFormatError, match="no image 5"):
        spiderio.read_image(path, 5)
    with pytest.raises(spiderio.SpiderFormatError, match="exceeds index limit"):
        spiderio.read_image(path, 16)
    with pytest.raises(spiderio.SpiderFormatError, match="give an image number"):
        spiderio.read_image(path)


def test_write_images_series_and_sparse_stack(tmp_path):
    """Transformed-images output: a `***` series and a stack whose image
    numbers skip slots (unused slots carry IMGNUM 0)."""
    rng = np.random.default_rng(9)
    imgs = rng.standard_normal((3, 8, 8)).astype(np.float32)
    spiderio.write_images(str(tmp_path / "r***"), "2,5,9", imgs, ext=".dat")
    np.testing.assert_array_equal(spiderio.read_images(str(tmp_path / "r***"), [2, 5, 9], ".dat"), imgs)
    assert os.path.exists(str(tmp_path / "r005.dat"))
    spiderio.write_images(str(tmp_path / "s@**"), [4, 1, 6], imgs, ext=".dat")
    h = spiderio.read_header(str(tmp_path / "s.dat"))
    assert h["istack"] > 0 and h["maxim"] == 6
    np.testing.assert_array_equal(spiderio.read_images(str(tmp_path / "s@**"), [4, 1, 6], ".dat"), imgs)
    np.testing.assert_array_equal(spiderio.read_image(str(tmp_path / "s.dat"), 3), np.zeros((8, 8), np.float32))
