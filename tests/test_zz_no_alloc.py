"""xr_project_noalloc against xr_project, through the C ABI.

The variant keeps its per-call buffers (view list, flags, FP64 loss, max
|voxel|) in the library's fixed per-GPU scratch array; its results must equal
xr_project's bit for bit (same kernels), call after call, and a smaller call
after a larger one must not see the larger call's leftovers.
"""
import numpy as np
import pytest

from conftest import ellipsoid_phantom
from test_gpu_parity import Setup, columns, to_dev

pytestmark = pytest.mark.gpu


def test_noalloc_matches_project(orc, xr, ctx, torch_dev):
    import torch

    st = Setup(orc, xr, N=48, U=72, P=40, vs=0.5)
    vol, _ = ellipsoid_phantom(orc, xr, st.N, st.vs, seed=901)
    vcol = columns(ctx, vol, st.N, 1, torch_dev)
    plan = ctx.plan(st.geom, st.scan)
    n = st.P * st.V * st.U
    y = torch.rand(n, device=torch_dev, generator=torch.Generator(device=torch_dev).manual_seed(5)) * 2.0
    # the plan's tables do not exist yet: the variant refuses instead of allocating them
    p0 = torch.zeros(n, device=torch_dev)
    with pytest.raises(xr.XRError, match="normalisation table"):
        plan.project_noalloc(vcol, p0, plan.codes([0]), out_mode=xr.OUT_RATIO, sino=y)
    plan.project(vcol, p0, plan.codes([0]), out_mode=xr.OUT_NORMALISED)  # builds the tables
    # two calls in a row, the second smaller; each against xr_project
    for views, mode in ((list(range(st.P)), xr.OUT_RATIO), ([3, 17, 29], xr.OUT_RESIDUAL),
                        (list(range(0, st.P, 2)), xr.OUT_NORMALISED), ([11], xr.OUT_RATIO)):
        codes = plan.codes(views)
        want = torch.zeros(n, device=torch_dev)
        got = torch.zeros(n, device=torch_dev)
        lw = plan.project(vcol, want, codes, out_mode=mode, sino=y)
        lg = plan.project_noalloc(vcol, got, codes, out_mode=mode, sino=y)
        ctx.synchronize()
        assert torch.equal(got, want), (views, mode)
        # FP64 sums of the same terms, atomically added in another order
        assert abs(lg - lw) <= 1e-12 * max(abs(lw), 1e-300), (views, mode, lg, lw)
    # accumulate flags (a view repeated within an epoch adds to its image)
    codes = plan.codes([4, 5, 6, 7])
    acc = np.array([1, 0, 1, 0], np.uint8)
    base = torch.rand(n, device=torch_dev, generator=torch.Generator(device=torch_dev).manual_seed(6))
    want, got = base.clone(), base.clone()
    plan.project(vcol, want, codes, out_mode=xr.OUT_RAW, accumulate=acc)
    plan.project_noalloc(vcol, got, codes, out_mode=xr.OUT_RAW, accumulate=acc)
    ctx.synchronize()
    assert torch.equal(got, want)
    # too many views for the scratch array: a clear error, nothing launched
    with pytest.raises(xr.XRError, match="at most 16384 views"):
        plan.project_noalloc(vcol, got, np.zeros(16385, np.int32), out_mode=xr.OUT_RAW)
    plan.close()


def test_noalloc_xtt_matches_project(orc, xr, ctx, torch_dev):
    import importlib

    import torch

    wl = importlib.import_module(xr.__name__ + ".workloads")
    st = Setup(orc, xr, N=32, U=56, P=24, vs=0.5, cond_R=np.tile(np.eye(3).reshape(-1), 2),
               basis=wl.XTT_BASIS13, grating=wl.grating([0, 90]))
    K = 13
    rng = np.random.default_rng(902)
    vol = (rng.random((K, st.N, st.N, st.N)) * 0.02).astype(np.float32)
    vcol = columns(ctx, vol, st.N, K, torch_dev)
    plan = ctx.plan(st.geom, st.scan)
    n = 2 * st.P * st.V * st.U
    y = to_dev(rng.random(n) * 0.1, torch_dev)
    plan.project(vcol, torch.zeros(n, device=torch_dev), plan.codes([0]), xtt=True)
    for codes in (np.arange(2 * st.P, dtype=np.int32), np.array([1, st.P + 5], np.int32)):
        want = torch.zeros(n, device=torch_dev)
        got = torch.zeros(n, device=torch_dev)
        lw = plan.project(vcol, want, codes, out_mode=xr.OUT_RATIO, sino=y, xtt=True)
        lg = plan.project_noalloc(vcol, got, codes, out_mode=xr.OUT_RATIO, sino=y, xtt=True)
        ctx.synchronize()
        assert torch.equal(got, want)
        assert abs(lg - lw) <= 1e-12 * max(abs(lw), 1e-300)
    plan.close()
