"""The batched workload's problem shapes: volumes 64^3 ... 256^3, 90 ... 360
views, detector 1.5x the volume edge.  Every (N, P) pair of the draws runs in
one batch (persistent launches, HIP graph, autotuned tiles) for 2 epochs; each
problem must equal its standalone xr_ir_reconstruct solve (IR::reconstruct,
reconstruct.cu:23-156) bit for bit."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = (64, 96, 128, 192, 256)
VIEWS = (90, 180, 360)


def test_every_shape_equals_standalone(xr, ctx, torch_dev):
    import importlib

    import torch

    wl = importlib.import_module(xr.__name__ + ".workloads")
    pairs = list(itertools.product(SIZES, VIEWS))
    vs = xr.default_vox_size(1000.0, 500.0, 0.4)
    geoms = [xr.make_geometry(1000.0, 500.0, 0.4, N, 3 * N // 2, P, vox_size=vs) for N, P in pairs]
    assert set((g.n_vox, g.n_proj) for g in geoms) == set(pairs)
    assert set(wl.C4_SIZES) == set(SIZES) and set(wl.C4_VIEWS) == set(VIEWS)
    sinos = []
    for i, (g, (N, P)) in enumerate(zip(geoms, pairs)):
        ref = torch.zeros(N ** 3, device=torch_dev)
        ctx.phantom_rasterise(xr.ellipsoid_params(400 + i, 10, fov_radius=0.5 * N * vs), N, vs, ref)
        col = torch.empty_like(ref)
        ctx.to_columns(ref, col, N)
        plan = ctx.plan(g, xr.Scan())
        sino = torch.zeros(P * g.n_det_u * g.n_det_v, device=torch_dev)
        plan.project(col, sino, plan.codes(range(P)))
        ctx.synchronize()
        plan.close()
        sinos.append(sino.cpu().numpy())
        del ref, col, sino
    epochs = 2
    init = [np.full(g.n_vox ** 3, 0.01, np.float32) for g in geoms]
    print("batched solve of %d problems" % len(pairs), flush=True)
    got = [v.copy() for v in init]
    losses = ctx.batch_reconstruct(geoms, [s.copy() for s in sinos], got, epochs,
                                   flags=xr.BATCH_GRAPH | xr.BATCH_AUTOTUNE)
    for i, (g, (N, P)) in enumerate(zip(geoms, pairs)):
        v = [init[i].copy()]
        want_loss = ctx.ir_reconstruct(g, xr.Scan(), [sinos[i].copy()], v, epochs, 1)
        print("problem N=%d P=%d max |diff| %g" % (N, P, float(np.max(np.abs(got[i] - v[0])))), flush=True)
        assert np.array_equal(got[i], v[0]), (N, P)
        np.testing.assert_allclose(losses[:, i], want_loss, rtol=1e-9)
