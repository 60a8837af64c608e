"""GPU parity of the gradient-domain phase-contrast path (IR::gradReconstruct,
reconstruct.cu:158-328, with poissonSolveLDLT, poisson_cpu.cpp:15-112) against
the CPU oracle, through the C ABI.

Tolerances: the XR_CONV_GRAD projector pair <= 1e-5 relative L2 against the
FP64 oracle (as the IR pair); the Poisson solve <= 1e-6 relative L2 against
SciPy's sparse LU on the same FP32 right sides (the GPU solves in FP64 and
rounds once); the gradient slots after 2 OS-ART epochs <= 1e-4, and the
integrated volume <= 1e-4.
"""
import numpy as np
import pytest

from conftest import rel_l2
from test_gpu_parity import record, rot_x, to_dev

pytestmark = pytest.mark.gpu


class GradSetup:
    """Oracle case of the result side N and the GPU geometry of side N+1."""

    def __init__(self, orc, xr, N, U, P, V=None, vs=0.5, sdd=600.0, sod=300.0, det=0.8, cond_R=None, cond_t=None,
                 cond_offset=None, n_cond=1):
        kw = dict(cond_R=None if cond_R is None else np.asarray(cond_R, np.float64).reshape(-1), cond_t=cond_t,
                  cond_offset=cond_offset)
        if cond_R is not None:
            n_cond = len(kw["cond_R"]) // 9
        self.case = orc.Case(sdd, sod, det, N, U, P, vox_size=vs, V=V, n_cond=n_cond, **kw)
        self.geom = xr.make_geometry(sdd, sod, det, N, U, P, vox_size=vs, n_det_v=V)
        self.geom1 = xr.make_geometry(sdd, sod, det, N + 1, U, P, vox_size=vs, n_det_v=V)
        self.scan = xr.Scan(n_cond=n_cond, **kw)
        self.N, self.U, self.V, self.P, self.C = N, U, U if V is None else V, P, n_cond


OPS = {
    "plain": dict(N=24, U=40, P=16),
    "tilted_offsets": dict(N=21, U=36, V=32, P=12, cond_R=rot_x(15.0), cond_t=[0.4, -0.3, 0.2],
                           cond_offset=[0.3, 0.6, -0.2]),
}


@pytest.mark.parametrize("name", list(OPS))
def test_grad_projector_pair_matches_oracle(orc, xr, ctx, torch_dev, name):
    import torch

    st = GradSetup(orc, xr, **OPS[name])
    N1 = st.N + 1
    rng = np.random.default_rng(11)
    G = rng.standard_normal((N1, N1, N1)).astype(np.float32)  # signed, like a gradient
    views = sorted({0, 1, st.P // 3, st.P - 1})
    for direction in (xr.CW, xr.CCW):
        plan = ctx.plan(st.geom1, st.scan, convention=xr.CONV_GRAD, direction=direction)
        vref = to_dev(G, torch_dev)
        vcol = torch.empty_like(vref)
        ctx.to_columns(vref, vcol, N1, 1)
        proj = torch.zeros(st.P * st.V * st.U, device=torch_dev)
        plan.project(vcol, proj, plan.codes(views), out_mode=xr.OUT_NORMALISED)
        ctx.synchronize()
        got = proj.cpu().numpy().reshape(st.P, st.V, st.U)
        # the splat already carries vs and nothing normalises it: RAW = NORMALISED
        raw = torch.zeros_like(proj)
        plan.project(vcol, raw, plan.codes(views), out_mode=xr.OUT_RAW)
        ctx.synchronize()
        assert np.array_equal(raw.cpu().numpy().reshape(got.shape), got)
        sgn = 1 if direction == xr.CW else -1
        worst = 0.0
        for v in views:
            worst = max(worst, rel_l2(got[v], st.case.grad_forward_view(G, sgn * v)))
        record("grad_forward_f64/%s/%d" % (name, direction), worst)
        assert worst <= 1e-5, (name, direction, worst)
        # the gather of random images
        img = rng.random((st.P, st.V, st.U)).astype(np.float32)
        tmp = torch.zeros(N1 ** 3, device=torch_dev)
        fac = torch.zeros(N1 ** 3, device=torch_dev)
        plan.backproject(to_dev(img, torch_dev), plan.codes(views), tmp, fac)
        tref = torch.empty_like(tmp)
        fref = torch.empty_like(fac)
        ctx.from_columns(tmp, tref, N1)
        ctx.from_columns(fac, fref, N1)
        want_t = np.zeros(N1 ** 3)
        want_f = np.zeros(N1 ** 3)
        for v in views:
            st.case.grad_backward_view(img[v], sgn * v, want_t, want_f)
        err = rel_l2(tref.cpu().numpy(), want_t)
        record("grad_backward_f64/%s/%d" % (name, direction), err)
        assert err <= 1e-5, (name, direction, err)
        assert np.count_nonzero(fref.cpu().numpy() != want_f) <= 1e-4 * N1 ** 3


def test_grad_repeated_view_adds_the_new_splat_only(orc, xr, ctx, torch_dev):
    """A view repeated within an epoch adds vs*ratio*sum(c mu) to the image it
    already holds (forwardProjGrad accumulates, nothing rescales afterwards)."""
    import torch

    st = GradSetup(orc, xr, N=16, U=26, P=8)
    N1 = st.N + 1
    rng = np.random.default_rng(12)
    G = rng.random((N1, N1, N1)).astype(np.float32)
    plan = ctx.plan(st.geom1, st.scan, convention=xr.CONV_GRAD)
    vref = to_dev(G, torch_dev)
    vcol = torch.empty_like(vref)
    ctx.to_columns(vref, vcol, N1, 1)
    base = rng.standard_normal((st.P, st.V, st.U)).astype(np.float32)
    proj = to_dev(base, torch_dev)
    acc = np.ones(st.P, np.uint8)
    plan.project(vcol, proj, plan.codes(range(st.P)), out_mode=xr.OUT_NORMALISED, accumulate=acc)
    ctx.synchronize()
    got = proj.cpu().numpy().reshape(st.P, st.V, st.U)
    for v in (0, 3, 7):
        want = base[v] + st.case.grad_forward_view(G, v)
        assert rel_l2(got[v], want) <= 1e-5


def gpu_poisson(xr, ctx, dev, n, g0, g1):
    import torch

    out = torch.full((n ** 3,), 7.0, device=dev)  # the border must come back zero
    ctx.poisson_solve(n, to_dev(g0, dev), to_dev(g1, dev), out)
    ctx.synchronize()
    return out.cpu().numpy().reshape(n, n, n)


@pytest.mark.parametrize("n", [3, 24, 70])
def test_poisson_matches_sparse_lu(orc, xr, ctx, torch_dev, n):
    rng = np.random.default_rng(n)
    g0 = rng.standard_normal((n + 1,) * 3).astype(np.float32)
    g1 = rng.standard_normal((n + 1,) * 3).astype(np.float32)
    got = gpu_poisson(xr, ctx, torch_dev, n, g0, g1)
    want = orc.poisson_solve(g0, g1, n)
    err = rel_l2(got, want)
    record("poisson_f64/%d" % n, err)
    assert err <= 1e-6, err
    assert np.all(got[:, 0, :] == 0) and np.all(got[:, -1, :] == 0)
    assert np.all(got[:, :, 0] == 0) and np.all(got[:, :, -1] == 0)


def test_poisson_recovers_a_known_potential(orc, xr, ctx, torch_dev):
    """Backward-difference gradients of f (zero on the slice borders) integrate
    back to f: 70 > 64 exercises the partial tiles of the transform."""
    n = 70
    rng = np.random.default_rng(5)
    f = np.zeros((n, n, n))
    f[:, 1:n - 1, 1:n - 1] = rng.standard_normal((n, n - 2, n - 2))
    g0 = np.zeros((n + 1,) * 3, np.float32)
    g1 = np.zeros((n + 1,) * 3, np.float32)
    g0[:n, :n, 1:n] = f[:, :, 1:] - f[:, :, :-1]
    g1[:n, 1:n, :n] = f[:, 1:, :] - f[:, :-1, :]
    got = gpu_poisson(xr, ctx, torch_dev, n, g0, g1)
    err = rel_l2(got, f)
    record("poisson_potential/%d" % n, err)
    assert err <= 1e-5, err


SOLVES = {
    "cw_full_extent": dict(geo=dict(N=16, U=26, P=12), direction=0, block_size=32),
    "cw_two_conditions_skipped_faces": dict(
        geo=dict(N=16, U=26, V=24, P=12, cond_R=np.concatenate([np.eye(3).reshape(-1), rot_x(10.0).reshape(-1)]),
                 cond_t=[0.0, 0.0, 0.0, 0.3, -0.2, 0.1], cond_offset=[0.0, 0.0, 0.0, 0.2, 0.1, -0.3]),
        direction=0, block_size=8),
    # 3 subsets of 4 views over 10: views 0 and 1 come twice an epoch
    "cw_repeated_views": dict(geo=dict(N=15, U=24, P=10), direction=0, block_size=4),
    "ccw": dict(geo=dict(N=15, U=24, P=12), direction=1, block_size=4),
    "ccw_repeated_views": dict(geo=dict(N=15, U=24, P=10), direction=1, block_size=4),
}


def _sinogram(st, seed):
    """Smooth positive data: the IR projections of a random blob volume."""
    rng = np.random.default_rng(seed)
    vol = np.zeros((st.N,) * 3, np.float32)
    z, y, x = np.mgrid[0:st.N, 0:st.N, 0:st.N] - (st.N - 1) / 2
    for _ in range(4):
        c = rng.uniform(-st.N / 5, st.N / 5, 3)
        r = rng.uniform(st.N / 8, st.N / 4)
        vol += (((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) < r * r).astype(np.float32) * rng.uniform(0.5, 1)
    return np.stack([st.case.forward_view(vol, n)[0] for n in range(st.P)]).astype(np.float32)


@pytest.mark.parametrize("name", list(SOLVES))
def test_grad_reconstruct_matches_oracle(orc, xr, ctx, name):
    cfg = SOLVES[name]
    st = GradSetup(orc, xr, **cfg["geo"])
    sino = _sinogram(st, 21)
    epochs, batch, lam = 2, 3, 1e-2
    N1 = st.N + 1
    vol = np.zeros(st.N ** 3, np.float32)
    grads = [np.zeros(N1 ** 3, np.float32) for _ in range(3 * st.C)]
    losses = ctx.grad_reconstruct(st.geom, st.scan, sino, vol, epochs, batch, direction=cfg["direction"],
                                  method=xr.METHOD_ART, lam=lam, block_size=cfg["block_size"], grads=grads)
    want_g, want_loss = st.case.grad_reconstruct(sino, epochs, batch, direction=cfg["direction"], art=True, lam=lam,
                                                 block_size=cfg["block_size"])
    worst = max(rel_l2(grads[i], want_g[i].reshape(-1)) for i in range(3 * st.C))
    record("grad_slots/%s" % name, worst)
    assert worst <= 1e-4, worst
    ext = st.case.grad_extent(cfg["block_size"])
    if ext == st.N:
        for gsl in grads:
            g = gsl.reshape(N1, N1, N1)
            assert np.all(g[st.N] == np.float32(0.01)) and np.all(g[:, :, st.N] == np.float32(0.01))
    # the per-epoch loss summed over slots is the reference's loss1
    assert np.allclose(losses.sum(axis=0), want_loss, rtol=1e-4)
    want_v = orc.poisson_solve(want_g[0], want_g[1], st.N)
    err = rel_l2(vol, want_v.reshape(-1))
    record("grad_volume/%s" % name, err)
    assert err <= 1e-4, err


def test_grad_reconstruct_mlem_third_slot(orc, xr, ctx):
    """OS-EM on the slot with coefficient 1 (positive data) against the oracle.
    The x and y slots hold signed data, where the reference's ratio y/p has no
    lower clamp and OS-EM is not a meaningful solver; their values are only
    checked to be finite."""
    st = GradSetup(orc, xr, N=14, U=24, P=10)
    sino = _sinogram(st, 22)
    N1 = st.N + 1
    vol = np.zeros(st.N ** 3, np.float32)
    grads = [np.zeros(N1 ** 3, np.float32) for _ in range(3)]
    ctx.grad_reconstruct(st.geom, st.scan, sino, vol, 2, 2, method=xr.METHOD_MLEM, block_size=32, grads=grads)
    want_g, _ = st.case.grad_reconstruct(sino, 2, 2, art=False, block_size=32)
    err = rel_l2(grads[2], want_g[2].reshape(-1))
    record("grad_mlem_slot2", err)
    assert err <= 1e-4, err
    assert all(np.isfinite(g).all() for g in grads)
