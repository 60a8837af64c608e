"""Pins the oracle's gradient-domain path (IR::gradReconstruct, reconstruct.cu:
158-328; forwardProjGrad / backwardProjGrad / sinogramGradientCoef, ir.cu:425-594;
poissonSolveLDLT, poisson_cpu.cpp:15-112) against closed forms, the already
pinned IR forward projector, adjointness, a fixed point, and a known potential.
"""
import numpy as np
import pytest

from oracle import oracle as orc


def _tilted_case(N=16, U=28, V=26, P=12, mirrored_offset=False):
    ang = 0.3
    R = np.array([[1, 0, 0], [0, np.cos(ang), -np.sin(ang)], [0, np.sin(ang), np.cos(ang)]])
    t = np.array([0.3, -0.2, 0.5])
    off = np.array([0.7, 0.4, -0.3])
    if mirrored_offset:
        off = np.array([-off[0], -off[1], off[2]])
    return orc.Case(1000.0, 500.0, 0.4, N, U, P, V=V, cond_R=R.reshape(-1), cond_t=t, cond_offset=off)


def test_grad_coef_closed_form():
    P = 36
    for n in range(P):
        th = 2 * np.pi * n / P
        assert abs(orc.grad_coef(n, P, 0) - np.cos(th)) < 1e-6
        assert abs(orc.grad_coef(n, P, 1) - np.sin(th)) < 1e-6
        assert orc.grad_coef(n, P, 2) == 1.0


@pytest.mark.parametrize("n", [0, 1, 5, -3, 7])
def test_grad_forward_is_mirrored_half_turn_ir_forward(n):
    """The gradient grid puts voxel x at (x - (N+1)/2)*vs, the IR grid of side N
    does too, and the source sits at R(0,-sod,0): the gradient view at angle
    theta is the IR view at theta + pi read right to left, times vs, once the
    offset is turned by pi about z.  Any slip in the half-voxel shift, the
    source side, the detector axis or the vs factor breaks this."""
    N = 16
    cg = _tilted_case(N)
    ci = _tilted_case(N, mirrored_offset=True)
    rng = np.random.default_rng(1)
    G = np.zeros((N + 1,) * 3, np.float32)
    G[:N, :N, :N] = rng.random((N, N, N))
    pg = cg.grad_forward_view(G, n)
    pi, _ = ci.forward_view(G[:N, :N, :N], n + cg.P // 2, normalise=False)
    want = cg.g.vox_size * pi[:, ::-1]
    assert np.linalg.norm(pg - want) / np.linalg.norm(want) < 1e-5


def test_grad_backward_is_adjoint_of_forward():
    """<F g, r> = vs*ratio*<g, B r>: the gather uses the splat's weights."""
    N = 14
    c = _tilted_case(N, U=24, V=22, P=8)
    g = c.g
    ratio = np.float64(np.float32(g.vox_size) ** 2 / (g.det_size * (g.sod / g.sdd) * g.det_size * (g.sod / g.sdd)))
    rng = np.random.default_rng(2)
    G = rng.random((N + 1,) * 3).astype(np.float32)
    for n in (0, 3):
        r = rng.standard_normal((c.V, c.U)).astype(np.float32)
        tmp = np.zeros((N + 1) ** 3)
        fac = np.zeros((N + 1) ** 3)
        c.grad_backward_view(r, n, tmp, fac)
        lhs = float((c.grad_forward_view(G, n) * r).sum())
        rhs = g.vox_size * ratio * float((G.reshape(-1).astype(np.float64) * tmp).sum())
        assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
        assert set(np.unique(fac)) <= {0.0, 1.0}


def test_grad_extent_skips_the_outer_faces():
    """With N a multiple of blockSize the launches stop at x, z < N: those faces
    neither project nor gather; y runs to N."""
    N = 16
    c = orc.Case(1000.0, 500.0, 0.4, N, 28, 6)
    assert c.grad_extent(16) == N and c.grad_extent(32) == N + 1 and c.grad_extent(5) == N + 1
    rng = np.random.default_rng(3)
    G = rng.random((N + 1,) * 3).astype(np.float32)
    Gz = G.copy()
    Gz[N, :, :] = 0  # z = N
    Gz[:, :, N] = 0  # x = N
    a = c.grad_forward_view(G, 2, ext=N)
    b = c.grad_forward_view(Gz, 2, ext=N + 1)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12)
    assert not np.allclose(a, c.grad_forward_view(G, 2, ext=N + 1))
    tmp = np.zeros((N + 1) ** 3)
    fac = np.zeros((N + 1) ** 3)
    c.grad_backward_view(np.ones((c.V, c.U), np.float32), 2, tmp, fac, ext=N)
    fac = fac.reshape((N + 1,) * 3)
    assert fac[N].max() == 0 and fac[:, :, N].max() == 0 and fac[:N, N, :N].max() == 1


@pytest.mark.parametrize("art", [True, False])
def test_grad_reconstruct_fixed_point_of_the_third_slot(art):
    """Slot 2 (coefficient 1) starts at 0.01 on the launched voxels; with the
    sinogram set to that volume's own projections its residual is zero (ART)
    or its ratio one (MLEM), so it stays put while slots 0 and 1 move."""
    N, P, bs = 12, 8, 4
    c = orc.Case(1000.0, 500.0, 0.4, N, 20, P)
    ext = c.grad_extent(bs)
    assert ext == N
    x0 = np.full((N + 1,) * 3, 0.01, np.float32)
    sino = np.stack([c.grad_forward_view(x0, n, ext=ext) for n in range(P)]).astype(np.float32)
    grads, _ = c.grad_reconstruct(sino, 2, 2, art=art, lam=0.5, block_size=bs)
    assert np.abs(grads[2] - 0.01).max() < 1e-9
    assert np.abs(grads[0] - 0.01).max() > 1e-4
    # the faces the launches skip keep their initial value in every slot
    assert np.all(grads[:, N, :, :] == np.float64(np.float32(0.01)))
    assert np.all(grads[:, :, :, N] == np.float64(np.float32(0.01)))


def test_grad_reconstruct_one_subset_is_one_landweber_step():
    """batch 1, one epoch: x1 = x0 + lambda/P * B(c_d * y - F x0) per slot."""
    N, P, lam = 10, 6, 0.3
    c = orc.Case(1000.0, 500.0, 0.4, N, 18, P)
    rng = np.random.default_rng(4)
    sino = rng.random((P, c.V, c.U)).astype(np.float32)
    grads, _ = c.grad_reconstruct(sino, 1, 1, art=True, lam=lam, block_size=32)
    x0 = np.full((N + 1,) * 3, 0.01, np.float32)
    for d in range(3):
        tmp = np.zeros((N + 1) ** 3)
        fac = np.zeros((N + 1) ** 3)
        for n in range(P):
            y = (np.float32(orc.grad_coef(n, P, d)) * sino[n]).astype(np.float32)
            c.grad_backward_view(y - c.grad_forward_view(x0, n), n, tmp, fac)
        want = x0.astype(np.float64).reshape(-1) + np.float32(lam / P) * tmp
        # the view wrappers take the residual as FP32, the loop keeps FP64
        assert np.allclose(grads[d].reshape(-1), want, rtol=1e-5, atol=1e-8)
        assert not np.allclose(grads[(d + 1) % 3].reshape(-1), want, rtol=1e-3, atol=1e-6)


def _potential(n, seed):
    rng = np.random.default_rng(seed)
    f = np.zeros((n, n, n))
    f[:, 1:n - 1, 1:n - 1] = rng.standard_normal((n, n - 2, n - 2))
    g0 = np.zeros((n + 1,) * 3, np.float32)
    g1 = np.zeros((n + 1,) * 3, np.float32)
    # backward differences: g0(x) = f(x) - f(x-1), g1(y) = f(y) - f(y-1)
    g0[:n, :n, 1:n] = f[:, :, 1:] - f[:, :, :-1]
    g1[:n, 1:n, :n] = f[:, 1:, :] - f[:, :-1, :]
    return f, g0, g1


def test_poisson_recovers_a_known_potential():
    """b = div of the backward-difference gradients of f equals A f, so the
    solve returns f: a swapped gradient, a sign or an index slip breaks it."""
    n = 12
    f, g0, g1 = _potential(n, 5)
    out = orc.poisson_solve(g0, g1, n)
    assert np.abs(out - f).max() < 1e-5 * np.abs(f).max()


def test_poisson_residual_and_border():
    n = 10
    rng = np.random.default_rng(6)
    g0 = rng.standard_normal((n + 1,) * 3).astype(np.float32)
    g1 = rng.standard_normal((n + 1,) * 3).astype(np.float32)
    out = orc.poisson_solve(g0, g1, n)
    A = orc.poisson_matrix(n)
    b = orc.poisson_rhs(g0, g1, n)
    for z in range(n):
        r = A @ out[z].reshape(-1) - b[z].reshape(-1).astype(np.float64)
        assert np.abs(r).max() < 1e-9
    assert np.all(out[:, 0, :] == 0) and np.all(out[:, -1, :] == 0)
    assert np.all(out[:, :, 0] == 0) and np.all(out[:, :, -1] == 0)
    # the matrix: 4 on the interior diagonal, 1 on the border
    d = A.diagonal().reshape(n, n)
    assert np.all(d[1:-1, 1:-1] == 4) and d[0, 0] == 1
