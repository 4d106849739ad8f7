"""GPU checks of the central-slice projector (spbp_gather / spbp_scatter and
spider_amd.project) against the FP64 oracle.

Tolerances:
  * gather vs oracle.gather (same fp32 weights, FP64 sums on both sides):
    rel-L2 <= 2e-7 (one fp32 rounding of each output);
  * scatter vs oracle.scatter: rel-L2 <= 1e-6 (fp32 atomics in any order);
  * <G x, b> = <x, S b> from the GPU outputs: rel <= 1e-6;
  * projector chain vs oracle.project.forward64 / adjoint64: rel-L2 <= 1e-5
    (fp32 FFTs);
  * autograd gradient of 0.5 ||A v - y||^2 vs central finite differences of
    the FP64 loss (exact for a quadratic up to rounding): rel <= 1e-4.
"""
import numpy as np
import pytest

import oracle
from oracle import project as op
from spider_amd import _lib, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def rel(a, b):
    a = np.asarray(a).astype(np.complex128).ravel()
    b = np.asarray(b).astype(np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    from spider_amd import bp, project
    return bp, project


def _angles(rng, m):
    ang = synth.uniform_angles(m, rng)
    return np.concatenate([ang, [[0.0, 0.0, 0.0], [0.0, 90.0, 0.0], [90.0, 90.0, 90.0], [17.0, 180.0, 5.0]]])


def _rand_c(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("nx", [16, 25, 128])
def test_gather_matches_oracle(mods, nx):
    """nx = 128: the configuration-4 box (256^3 grid)."""
    bp, _ = mods
    rng = np.random.default_rng(100 + nx)
    dm = _lib.rotations(_angles(rng, 36 if nx < 128 else 12))
    N = 2 * nx
    X = _rand_c(rng, (N, N, nx + 1))
    with bp.Backprojector(nx) as ctx:
        out = ctx.gather(torch.from_numpy(X).cuda(), dm=dm).cpu().numpy()
    ref = oracle.gather(X.astype(np.complex128), dm)
    assert rel(out, ref) <= 2e-7
    # samples ONELINE skips are exactly zero
    np.testing.assert_array_equal(out[ref == 0], 0)


@pytest.mark.parametrize("strategy", ["direct", "brick"])
def test_scatter_matches_oracle(mods, strategy):
    bp, _ = mods
    nx = 20
    rng = np.random.default_rng(7)
    dm = _lib.rotations(_angles(rng, 60))
    N = 2 * nx
    B = _rand_c(rng, (dm.shape[0], N, nx + 1))
    with bp.Backprojector(nx, strategy=strategy) as ctx:
        W = torch.zeros(ctx.grid_shape, dtype=torch.float32, device="cuda")
        X = ctx.scatter(torch.from_numpy(B).cuda(), dm=dm, W=W)
        X2 = ctx.scatter(torch.from_numpy(B).cuda(), dm=dm)          # no weights
        Xg, Wg, X2g = X.cpu().numpy(), W.cpu().numpy(), X2.cpu().numpy()
    Xr, Wr = oracle.scatter(B.astype(np.complex128), dm, nx)
    assert rel(Xg, Xr) <= 1e-6
    assert rel(Wg, Wr) <= 1e-6
    assert rel(X2g, Xr) <= 1e-6


@pytest.mark.parametrize("geom_f64", [False, True])
@pytest.mark.parametrize("strategy", ["direct", "brick"])
def test_adjoint_identity(mods, strategy, geom_f64):
    """<G x, b> = <x, S b>; with FP64 geometry (FP64 rotations from the doc
    angles, FP64 sample positions) the gather transposes that context's
    insertion too."""
    bp, _ = mods
    nx = 32
    rng = np.random.default_rng(9)
    ang = _angles(rng, 100)
    rot = {"angles": ang} if geom_f64 else {"dm": _lib.rotations(ang)}
    N = 2 * nx
    X = _rand_c(rng, (N, N, nx + 1))
    B = _rand_c(rng, (ang.shape[0], N, nx + 1))
    with bp.Backprojector(nx, strategy=strategy, geom_f64=geom_f64) as ctx:
        GX = ctx.gather(torch.from_numpy(X).cuda(), **rot).cpu().numpy().astype(np.complex128)
        SB = ctx.scatter(torch.from_numpy(B).cuda(), **rot).cpu().numpy().astype(np.complex128)
    B64, X64 = B.astype(np.complex128), X.astype(np.complex128)
    lhs = np.sum(GX.real * B64.real + GX.imag * B64.imag)
    rhs = np.sum(X64.real * SB.real + X64.imag * SB.imag)
    scale = np.sqrt(np.sum(np.abs(GX) ** 2) * np.sum(np.abs(B64) ** 2))
    assert abs(lhs - rhs) <= 1e-6 * scale


def _projector(project, nx, m, seed):
    rng = np.random.default_rng(seed)
    ang = _angles(rng, m)
    return project.Projector(nx, angles=ang, batch=16), rng


def test_projector_matches_fp64_chain(mods):
    _, project = mods
    nx = 24
    P, rng = _projector(project, nx, 30, 21)
    try:
        D = P.D.double().cpu().numpy()
        v = rng.standard_normal((nx,) * 3).astype(np.float32)
        r = rng.standard_normal((P.nproj, nx, nx)).astype(np.float32)
        Av = P.apply(torch.from_numpy(v).cuda()).cpu().numpy()
        assert rel(Av, op.forward64(v, P.dm, D)) <= 1e-5
        Atr = P.adjoint(torch.from_numpy(r).cuda()).cpu().numpy()
        assert rel(Atr, op.adjoint64(r, P.dm, D)) <= 1e-5
        # the autograd path (SliceGather) gives the same values and, as its
        # vector-Jacobian product, the same A^T r
        vt = torch.from_numpy(v).cuda().requires_grad_(True)
        out = P(vt)
        assert rel(out.detach().cpu().numpy(), Av) <= 1e-6
        (g,) = torch.autograd.grad(out, vt, torch.from_numpy(r).cuda())
        assert rel(g.cpu().numpy(), Atr) <= 1e-5
    finally:
        P.close()


@pytest.mark.parametrize("nx", [16, 32, 64])
def test_library_adjoint_matches_torch_chain_and_fp64(mods, nx):
    """A^T r with the library's own padding, transforms and spectrum folding
    (spbp_scatter_images, power-of-two N) against the same chain in torch
    operations before spbp_scatter (rel-L2 <= 1e-6: two fp32 FFTs) and the
    FP64 adjoint (<= 1e-5), over several library calls (batch 16)."""
    _, project = mods
    P, rng = _projector(project, nx, 40, 31 + nx)
    try:
        D = P.D.double().cpu().numpy()
        r = rng.standard_normal((P.nproj, nx, nx)).astype(np.float32)
        rt = torch.from_numpy(r).cuda()
        P.library_chain = True
        lib = P.adjoint(rt).cpu().numpy()
        P.library_chain = False
        chain = P.adjoint(rt).cpu().numpy()
        assert rel(lib, chain) <= 1e-6
        assert rel(lib, op.adjoint64(r, P.dm, D)) <= 1e-5
    finally:
        P.close()


@pytest.mark.parametrize("nx", [16, 24, 64])
def test_library_forward_matches_torch_chain_and_fp64(mods, nx):
    """A v with the library's image synthesis (spbp_gather_images: completion,
    C2R, crop) against the same steps as torch operations after spbp_gather
    (rel-L2 <= 1e-6) and the FP64 forward chain (<= 1e-5), over several
    library calls (batch 16); also for a padded size that is not a power of
    two."""
    _, project = mods
    P, rng = _projector(project, nx, 40, 51 + nx)
    try:
        D = P.D.double().cpu().numpy()
        v = rng.standard_normal((nx,) * 3).astype(np.float32)
        vt = torch.from_numpy(v).cuda()
        P.library_chain = True
        lib = P.apply(vt).cpu().numpy()
        P.library_chain = False
        chain = P.apply(vt).cpu().numpy()
        assert rel(lib, chain) <= 1e-6
        assert rel(lib, op.forward64(v, P.dm, D)) <= 1e-5
    finally:
        P.close()


def test_library_adjoint_refuses_other_sizes(mods):
    """spbp_scatter_images needs the own transforms (power-of-two N): a 48^2
    padded size is refused with SPBP_ERR_SIZE, and the projector's adjoint
    then runs its torch chain."""
    bp, project = mods
    with bp.Backprojector(24) as ctx:
        with pytest.raises(_lib.SpbpError) as e:
            ctx.scatter_images(torch.zeros((2, 24, 24), device="cuda"), dm=_lib.rotations(np.zeros((2, 3))))
        assert e.value.code == _lib.ERR_SIZE


def test_gradient_against_fp64_finite_differences(mods):
    """32^3: the hand-written backward (scatter) through autograd against
    central differences of the FP64 loss along random directions."""
    _, project = mods
    nx = 32
    P, rng = _projector(project, nx, 20, 33)
    try:
        D = P.D.double().cpu().numpy()
        v = rng.standard_normal((nx,) * 3)
        y = rng.standard_normal((P.nproj, nx, nx))
        vt = torch.tensor(v, dtype=torch.float32, device="cuda", requires_grad=True)
        loss = 0.5 * torch.sum((P(vt) - torch.tensor(y, dtype=torch.float32, device="cuda")) ** 2)
        loss.backward()
        g = vt.grad.double().cpu().numpy()

        def L64(x):
            return 0.5 * np.sum((op.forward64(x, P.dm, D) - y) ** 2)

        for k in range(3):
            d = rng.standard_normal(v.shape)
            h = 1e-2
            fd = (L64(v + h * d) - L64(v - h * d)) / (2 * h)
            an = float(np.sum(g * d))
            assert abs(fd - an) <= 1e-4 * abs(fd), (k, fd, an)
    finally:
        P.close()


def test_cgls_fits_the_data(mods):
    """CGLS on noise-free projections of a phantom: the residual falls
    monotonically and the reconstruction correlates with the phantom."""
    _, project = mods
    nx = 32
    rng = np.random.default_rng(44)
    ph = synth.blob_phantom(nx, 10, 8.0, rng)
    ang = synth.uniform_angles(400, rng)
    P = project.Projector(nx, angles=ang, batch=128)
    try:
        imgs = synth.project(ph, P.dm.astype(np.float64), device="cuda")
        x, hist = project.cgls(P, imgs, iters=8)
        res = [h["residual"] for h in hist]
        assert all(b <= a * (1 + 1e-6) for a, b in zip(res, res[1:]))
        assert res[-1] < 0.2 * float(torch.linalg.norm(imgs.double()))
        xv = x.cpu().numpy().ravel()
        pv = ph.volume().ravel()
        cc = np.corrcoef(xv, pv)[0, 1]
        assert cc > 0.9, cc
    finally:
        P.close()


def test_density_preconditioned_cg(mods):
    """The density preconditioner is symmetric and positive semi-definite
    (<u, M v> = <M u, v>, <u, M u> > 0 on random u); PCG on noise-free (consistent)
    data lowers the data misfit ||A x - y|| at every iteration (it minimises
    the A^T A-norm error over growing Krylov spaces) and approaches the
    phantom; heavy damping shrinks the solution toward zero.""" 
    _, project = mods
    nx = 32
    rng = np.random.default_rng(45)
    ph = synth.blob_phantom(nx, 10, 8.0, rng)
    ang = synth.uniform_angles(400, rng)
    P = project.Projector(nx, angles=ang, batch=128)
    try:
        M = project.DensityPreconditioner(P)
        u = torch.randn((nx,) * 3, device="cuda", dtype=torch.float32)
        v = torch.randn((nx,) * 3, device="cuda", dtype=torch.float32)
        a, b = float(torch.sum(u.double() * M(v).double())), float(torch.sum(M(u).double() * v.double()))
        assert abs(a - b) <= 1e-5 * (abs(a) + abs(b)), (a, b)
        assert float(torch.sum(u.double() * M(u).double())) > 0
        imgs = synth.project(ph, P.dm.astype(np.float64), device="cuda")
        xs = []
        xp, _ = project.pcg(P, imgs, iters=6, precond=M, callback=lambda it, x: xs.append(x.clone()))
        mis = [float(torch.linalg.norm((P.apply(x) - imgs).double())) for x in xs]
        assert all(b <= a * (1 + 1e-6) for a, b in zip(mis, mis[1:])), mis
        assert mis[-1] < 0.5 * float(torch.linalg.norm(imgs.double()))
        pv = ph.volume().ravel()
        cc = [np.corrcoef(x.cpu().numpy().ravel(), pv)[0, 1] for x in xs]
        assert cc[0] > 0.7 and cc[-1] > 0.9, cc
        lam = float(torch.sum(u * P.adjoint(P.apply(u))) / torch.sum(u * u))
        xd, _ = project.pcg(P, imgs, iters=6, precond=M, damp=(1e3 * lam) ** 0.5)
        assert float(torch.linalg.norm(xd)) < 0.05 * float(torch.linalg.norm(xp))
    finally:
        P.close()
