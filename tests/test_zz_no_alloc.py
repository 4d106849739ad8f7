"""spbp_scatter_noalloc (no context, scratch in a static device array)
against spbp_scatter on a context: two calls in a row, the second smaller
(and weight-less, with FP64 geometry), each to the insertion's parity bar --
identical weight support, rel-L2 <= 1e-6 for X and W (the same kernels; only
the fp32 order of the flush atomics differs); sizes beyond the scratch and
flags that need a context are refused before any GPU work."""
import ctypes as C

import numpy as np
import pytest

from spider_amd import _lib, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def rel(a, b):
    a = np.asarray(a).astype(np.complex128).ravel()
    b = np.asarray(b).astype(np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_noalloc_scatter_matches_context_scatter():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    from spider_amd import bp
    L = _lib.lib()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for call, (nx, m, geom_f64, weights) in enumerate([(32, 300, False, True), (24, 40, True, False)]):
        print(f"case noalloc call {call}: nx={nx} m={m} geom_f64={geom_f64} weights={weights}", flush=True)
        rng = np.random.default_rng(500 + call)
        N, H = 2 * nx, nx + 1
        spec = (rng.standard_normal((m, N, H)) + 1j * rng.standard_normal((m, N, H))).astype(np.complex64)
        ang = np.ascontiguousarray(synth.uniform_angles(m, rng))
        Bt = torch.from_numpy(spec).cuda()
        with bp.Backprojector(nx, geom_f64=geom_f64) as ctx:
            Wr = torch.zeros((N, N, H), dtype=torch.float32, device="cuda") if weights else None
            Xr = ctx.scatter(Bt, angles=ang, W=Wr)
            torch.cuda.synchronize()
        X = torch.zeros((N, N, H), dtype=torch.complex64, device="cuda")
        W = torch.zeros((N, N, H), dtype=torch.float32, device="cuda") if weights else None
        flags = _lib.GEOM_F64 if geom_f64 else 0
        rc = L.spbp_scatter_noalloc(nx, flags, C.c_void_p(Bt.data_ptr()), m, ang.ctypes.data_as(C.c_void_p), None,
                                    C.c_void_p(X.data_ptr()), None if W is None else C.c_void_p(W.data_ptr()), stream)
        _lib.check(rc)
        torch.cuda.synchronize()
        assert rel(X.cpu().numpy(), Xr.cpu().numpy()) <= 1e-6, call
        if weights:
            w, wr = W.cpu().numpy(), Wr.cpu().numpy()
            np.testing.assert_array_equal(w > 0, wr > 0)
            assert rel(w, wr) <= 1e-6
    # refused before any GPU work: a box beyond the scratch, flags that need a context
    x = C.c_void_p(1)
    assert L.spbp_scatter_noalloc(1024, 0, x, 1, ang.ctypes.data_as(C.c_void_p), None, x, None, stream) == _lib.ERR_SIZE
    assert L.spbp_scatter_noalloc(32, _lib.TWO_HALVES, x, 1, ang.ctypes.data_as(C.c_void_p), None, x, None,
                                  stream) == _lib.ERR_ARG
