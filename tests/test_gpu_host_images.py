"""GPU checks of insertion from host memory (spbp_insert_host,
spbp_insert_rb_host): the images are uploaded in sub-chunks on a stream of
their own, into two device buffers by turns, overlapping the insertion;
pageable images go through the context's two page-locked staging buffers,
page-locked ones straight to the device.

Bars: with the gather plan, whose sums are the serial reference's whatever
the chunking, the host path gives the device path's accumulators bit for bit
(fp32 and BF16 images, pageable and page-locked, one and two halves, RB
in-plane transforms, several sub-chunks so that both buffers are reused, and
a second call on the same context reusing the staging; device-resident
images handed to the host entry point are copied as they are); with the default
plan the finalised volumes agree to rel-L2 1e-6 and meet the FP32-faithful
oracle; a failed upload-buffer allocation returns SPBP_ERR_ALLOC and leaves
the context usable.
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


def rel(a, b):
    a = np.asarray(a).astype(np.complex128).ravel()
    b = np.asarray(b).astype(np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def dataset(nx, m, seed):
    rng = np.random.default_rng(seed)
    imgs = rng.standard_normal((m, nx, nx)).astype(np.float32)
    return imgs, synth.uniform_angles(m, rng), rng


def raw(b, halves):
    out = []
    for h in range(2 if halves else 1):
        X, W = b.export_raw(h)
        torch.cuda.synchronize()
        out.append((X.cpu().numpy(), W.cpu().numpy()))
    return out


@pytest.mark.parametrize("bf16", [False, True])
@pytest.mark.parametrize("halves", [False, True])
@pytest.mark.parametrize("pinned", [False, True])
def test_host_images_are_the_device_path_bit_for_bit(bpmod, bf16, halves, pinned):
    nx, m = 24, 70
    imgs, ang, _ = dataset(nx, m, 11 + 2 * bf16 + halves)
    half = bpmod.random_halves(m, seed=3) if halves else None
    host = torch.from_numpy(imgs)
    if bf16:
        host = host.to(torch.bfloat16)
    if pinned:
        host = host.pin_memory()
    out = {}
    for where in ("device", "host"):
        # batch 16: five sub-chunks, both upload buffers used twice; the host
        # path in two calls, the second reusing the context's staging
        with bpmod.Backprojector(nx, strategy="gather", two_halves=halves, batch=16) as b:
            if where == "device":
                b.insert(host.cuda(), ang, half)
            else:
                k = 37
                b.insert_host(host[:k], ang[:k], None if half is None else half[:k])
                b.insert_host(host[k:], ang[k:], None if half is None else half[k:])
            out[where] = raw(b, halves)
    for (Xd, Wd), (Xh, Wh) in zip(out["device"], out["host"]):
        np.testing.assert_array_equal(Wh, Wd)
        np.testing.assert_array_equal(Xh.view(np.uint64), Xd.view(np.uint64))


def test_insert_host_takes_device_memory_too(bpmod):
    """Device-resident images handed to spbp_insert_host are copied as they
    are (not staged through host memory): the same accumulators."""
    from spider_amd import _lib
    import ctypes as C
    nx, m = 20, 40
    imgs, ang, _ = dataset(nx, m, 61)
    dev = torch.from_numpy(imgs).cuda()
    out = {}
    for where in ("insert", "insert_host"):
        with bpmod.Backprojector(nx, strategy="gather", batch=16) as b:
            if where == "insert":
                b.insert(dev, ang)
            else:
                a = np.ascontiguousarray(ang, dtype=np.float64)
                rc = _lib.lib().spbp_insert_host(b._ctx, C.c_void_p(dev.data_ptr()), _lib.F32, m,
                                                 a.ctypes.data_as(C.c_void_p), None, bpmod._stream(b.device))
                assert rc == 0, rc
            out[where] = raw(b, False)
    np.testing.assert_array_equal(out["insert_host"][0][1], out["insert"][0][1])
    np.testing.assert_array_equal(out["insert_host"][0][0].view(np.uint64), out["insert"][0][0].view(np.uint64))


def test_host_rb_images_are_the_device_path_bit_for_bit(bpmod):
    nx, m = 20, 45
    imgs, ang, rng = dataset(nx, m, 29)
    inplane = np.stack([rng.uniform(-180, 180, m), rng.uniform(-2, 2, m), rng.uniform(-2, 2, m)], 1).astype(np.float32)
    out = {}
    for where in ("device", "host"):
        with bpmod.Backprojector(nx, strategy="gather", batch=16) as b:
            if where == "device":
                b.insert(torch.from_numpy(imgs).cuda(), ang, inplane=inplane)
            else:
                b.insert_host(imgs, ang, inplane=inplane)
            out[where] = raw(b, False)
    np.testing.assert_array_equal(out["host"][0][1], out["device"][0][1])
    np.testing.assert_array_equal(out["host"][0][0].view(np.uint64), out["device"][0][0].view(np.uint64))


def test_host_images_default_plan_meet_oracle(bpmod):
    nx, m = 32, 120
    imgs, ang, _ = dataset(nx, m, 41)
    half = bpmod.random_halves(m, seed=8)
    ref = oracle.bp(imgs, ang, half=half, prec="f32")
    vols = {}
    for where in ("device", "host"):
        with bpmod.Backprojector(nx, two_halves=True, batch=40) as b:
            if where == "device":
                b.insert(torch.from_numpy(imgs).cuda(), ang, half)
            else:
                b.insert_host(imgs, ang, half)
            vols[where] = {k: v.cpu().numpy() for k, v in b.finalize(full=True, halves=True).items()}
    for k in ("full", "h1", "h2"):
        assert rel(vols["host"][k], vols["device"][k]) <= 1e-6, k
        assert rel(vols["host"][k], ref[k]) <= 1e-5, k


def test_host_upload_allocation_failure_leaves_context_usable(bpmod):
    from spider_amd import _lib
    L = _lib.lib()
    nx, m = 16, 40
    imgs, ang, _ = dataset(nx, m, 53)
    ref = oracle.bp(imgs, ang, prec="f32")["full"]
    with bpmod.Backprojector(nx, batch=16) as b:
        for k in range(3):   # the first upload buffer, the second, then whatever comes next
            print(f"case failing allocation {k}", flush=True)
            L.spbp_debug_fail_alloc(k)
            try:
                b.insert_host(imgs, ang)
                done = True
            except bpmod.SpbpError as err:
                assert err.code == _lib.ERR_ALLOC, str(err)
                done = False
            finally:
                L.spbp_debug_fail_alloc(-1)
            b.reset()
            if done:
                break
        b.insert_host(imgs, ang)
        vol = b.finalize()["full"].cpu().numpy()
    assert rel(vol, ref) <= 1e-5
