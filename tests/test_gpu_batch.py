"""Batched multi-problem solver, HIP-graph replay and the tile autotuner.

The batched solver has no counterpart in the reference: each batched problem
is IR::reconstruct (reconstruct.cu:23-156) on its own, so the bar is that every
problem equals its standalone xr_ir_reconstruct solve bit for bit (same
kernels' arithmetic, same subset order, exact fixed-point forward sums), and
that the standalone solve itself matches the oracle.  Tile shapes only change
the work split, so every tile setting must give bitwise identical results.
"""
import numpy as np
import pytest

from conftest import ellipsoid_phantom, rel_l2

pytestmark = pytest.mark.gpu

SDD, SOD, DET = 600.0, 300.0, 0.8


def make_problem(orc, xr, ctx, torch_dev, N, P, seed, U=None):
    """(geometry, sinogram (host, normalised forward projection), phantom)."""
    import torch

    U = U or (3 * N) // 2
    vs = 0.5 * 64.0 / N  # same field of view for every N
    g = xr.make_geometry(SDD, SOD, DET, N, U, P, vox_size=vs)
    vol, _ = ellipsoid_phantom(orc, xr, N, vs, seed)
    scan = xr.Scan()
    plan = ctx.plan(g, scan)
    ref = torch.from_numpy(vol.reshape(-1)).to(torch_dev)
    col = torch.empty_like(ref)
    ctx.to_columns(ref, col, N)
    proj = torch.zeros(P * U * U, dtype=torch.float32, device=torch_dev)
    plan.project(col, proj, plan.codes(range(P)), out_mode=xr.OUT_NORMALISED)
    ctx.synchronize()
    sino = proj.cpu().numpy()
    plan.close()
    return g, sino, vol


# (N, P, subsets): uneven sizes; 45 views in 4 subsets wrap past P, so revisited
# views accumulate onto their image (reconstruct.cu:83)
CASES = [(64, 36, 1), (32, 24, 1), (48, 45, 4), (80, 30, 3), (40, 24, 2)]


@pytest.fixture(scope="module")
def problems(orc, xr, ctx, torch_dev):
    return [make_problem(orc, xr, ctx, torch_dev, N, P, 100 + i) for i, (N, P, _) in enumerate(CASES)]


@pytest.mark.parametrize("method", ["mlem", "art"])
def test_batch_equals_standalone(xr, ctx, problems, method):
    m = xr.METHOD_MLEM if method == "mlem" else xr.METHOD_ART
    epochs = 3
    geoms = [p[0] for p in problems]
    batch = [c[2] for c in CASES]
    init = [np.full(g.n_vox ** 3, 0.01, np.float32) for g in geoms]
    want, want_loss = [], []
    for (g, sino, _), b, v0 in zip(problems, batch, init):
        v = [v0.copy()]
        want_loss.append(ctx.ir_reconstruct(g, xr.Scan(), [sino.copy()], v, epochs, b, method=m, lam=0.5))
        want.append(v[0])
    got = [v0.copy() for v0 in init]
    losses = ctx.batch_reconstruct(geoms, [p[1].copy() for p in problems], got, epochs, batch=batch, method=m,
                                   lam=0.5)
    for i in range(len(CASES)):
        print("problem", i, CASES[i], "max |diff|", float(np.max(np.abs(got[i] - want[i]))), flush=True)
        assert np.array_equal(got[i], want[i]), "problem %d differs from its standalone solve" % i
        # loss sums: same terms, atomically added in another order
        np.testing.assert_allclose(losses[:, i], want_loss[i], rtol=1e-9)


def test_batch_matches_oracle(orc, xr, ctx, problems):
    """One batched problem against the CPU oracle's IR solve (FP64)."""
    g, sino, _ = problems[2]
    N, P = g.n_vox, g.n_proj
    case = orc.Case(SDD, SOD, DET, N, g.n_det_u, P, vox_size=g.vox_size)
    v0 = np.full(N ** 3, 0.01, np.float32)
    want, _ = case.ir_reconstruct(sino.reshape(P, g.n_det_u, g.n_det_u), v0, epochs=2, batch=CASES[2][2])
    got = [v0.copy()]
    ctx.batch_reconstruct([g], [sino.copy()], got, 2, batch=[CASES[2][2]])
    assert rel_l2(got[0], want.reshape(-1)) <= 1e-4


def test_graph_replay_matches_eager(xr, ctx, problems, torch_dev):
    """The captured epoch loop, replayed several times with eager work and
    synchronisation in between, equals the eager (uncaptured) enqueue."""
    import torch

    geoms = [p[0] for p in problems]
    batch = [c[2] for c in CASES]
    results = {}
    for flags in (0, xr.BATCH_GRAPH):
        b = ctx.batch(geoms, batch=batch, flags=flags)
        outs = []
        for rep in range(3):
            for i, (g, sino, _) in enumerate(problems):
                vptr, sptr = b.buffers(i)
                n3 = g.n_vox ** 3
                v = torch.full((n3,), 0.01, dtype=torch.float32, device=torch_dev)
                s = torch.from_numpy(sino).to(torch_dev)
                torch.cuda.synchronize()
                # eager work between replays: fresh inputs copied in
                ctx.copy(vptr, v)
                ctx.copy(sptr, s)
            torch.cuda.synchronize()
            b.restart()  # each replay is a fresh 2-epoch solve
            losses = b.run(2)
            outs.append(([read_back(ctx, b.buffers(i)[0], geoms[i].n_vox ** 3) for i in range(len(geoms))],
                         losses))
        b.close()
        results[flags] = outs
    eager = results[0][0]
    for rep, (vols, losses) in enumerate(results[xr.BATCH_GRAPH]):
        for i in range(len(geoms)):
            assert np.array_equal(vols[i], eager[0][i]), "replay %d problem %d" % (rep, i)
        np.testing.assert_allclose(losses, eager[1], rtol=1e-9)
    for rep, (vols, _) in enumerate(results[0]):
        for i in range(len(geoms)):
            assert np.array_equal(vols[i], eager[0][i])


def read_back(ctx, addr, n):
    out = np.empty(n, dtype=np.float32)
    ctx.copy(out, addr, 4 * n)
    return out


def test_tiles_do_not_change_results(orc, xr, ctx, torch_dev, problems):
    import torch

    g, sino, vol = problems[3]
    N, P, U = g.n_vox, g.n_proj, g.n_det_u
    plan = ctx.plan(g, xr.Scan())
    ref = torch.from_numpy(vol.reshape(-1)).to(torch_dev)
    col = torch.empty_like(ref)
    ctx.to_columns(ref, col, N)
    y = torch.from_numpy(sino).to(torch_dev)
    codes = plan.codes(range(P))
    base_fwd = base_bwd = None
    for rows, scan, split in ((0, 1, 1), (0, 0, 1), (0, 0, 2), (0, 0, 4), (0, 0, 0), (64, 0, 2), (40, 0, 4),
                              (17, 0, 1), (17, 1, 4)):
        plan.set_tiles(fwd_rows=rows, fwd_scan=scan, fwd_split=split)
        p = torch.zeros(P * U * U, dtype=torch.float32, device=torch_dev)
        plan.project(col, p, codes, out_mode=xr.OUT_RATIO, sino=y)
        ctx.synchronize()
        if base_fwd is None:
            base_fwd = p.clone()
        assert torch.equal(p, base_fwd), "fwd_rows=%d fwd_scan=%d fwd_split=%d changed the forward projection" % (
            rows, scan, split)
    for zpl in (1, 2, 4):
        for gx in (1, 2, 4, 8, 16):
            plan.set_tiles(bwd_zpl=zpl, bwd_gx=gx)
            v = col.clone()
            plan.backproject_update(base_fwd, codes, v, method=xr.METHOD_MLEM)
            ctx.synchronize()
            if base_bwd is None:
                base_bwd = v.clone()
            assert torch.equal(v, base_bwd), "tile (zpl %d, gx %d) changed the MLEM update" % (zpl, gx)
    plan.set_tiles()
    assert plan.tiles()[0] in (1, 2, 4)
    best, ms = plan.autotune(col, y)
    assert best[0] in (1, 2, 4) and best[1] in (2, 4, 8) and ms > 0
    v = col.clone()
    plan.backproject_update(base_fwd, codes, v, method=xr.METHOD_MLEM)
    ctx.synchronize()
    assert torch.equal(v, base_bwd)
    plan.close()


def test_batch_autotune_keeps_results(xr, ctx, problems):
    geoms = [p[0] for p in problems]
    batch = [c[2] for c in CASES]
    init = [np.full(g.n_vox ** 3, 0.01, np.float32) for g in geoms]
    a = [v.copy() for v in init]
    b = [v.copy() for v in init]
    ctx.batch_reconstruct(geoms, [p[1].copy() for p in problems], a, 2, batch=batch, flags=xr.BATCH_GRAPH)
    ctx.batch_reconstruct(geoms, [p[1].copy() for p in problems], b, 2, batch=batch,
                          flags=xr.BATCH_GRAPH | xr.BATCH_AUTOTUNE)
    for i in range(len(geoms)):
        assert np.array_equal(a[i], b[i])


def test_runs_continue_the_subset_order(xr, ctx, problems):
    """Runs of 1 + 2 epochs equal one 3-epoch run (and the standalone solve):
    the subset order continues across runs until restart."""
    geoms = [p[0] for p in problems]
    batch = [c[2] for c in CASES]
    b = ctx.batch(geoms, batch=batch)
    init = [np.full(g.n_vox ** 3, 0.01, np.float32) for g in geoms]

    def load():
        import torch

        for i, (g, sino, _) in enumerate(problems):
            vptr, sptr = b.buffers(i)
            ctx.copy(vptr, torch.from_numpy(init[i]).cuda())
            ctx.copy(sptr, torch.from_numpy(sino).cuda())

    load()
    b.run(1)
    b.run(2)
    split = [read_back(ctx, b.buffers(i)[0], geoms[i].n_vox ** 3) for i in range(len(geoms))]
    load()
    b.restart()
    b.run(3)
    whole = [read_back(ctx, b.buffers(i)[0], geoms[i].n_vox ** 3) for i in range(len(geoms))]
    for i in range(len(geoms)):
        assert np.array_equal(split[i], whole[i]), i
    b.close()


def rot_x(deg):
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]], np.float64)


@pytest.mark.parametrize("method", ["mlem", "art"])
def test_mixed_tilt_batch_equals_standalone(orc, xr, ctx, torch_dev, method):
    """A batch with a tilted problem (condition rotated 20 degrees about x)
    among untilted ones runs the general persistent forward and the general
    backprojection items; every problem still equals its standalone solve."""
    import torch

    m = xr.METHOD_MLEM if method == "mlem" else xr.METHOD_ART
    specs = [(40, 24, 1, None), (32, 18, 2, rot_x(20.0).reshape(-1)), (48, 30, 3, None)]
    geoms, scans, sinos, batch = [], [], [], []
    for i, (N, P, b, R) in enumerate(specs):
        U = (3 * N) // 2
        vs = 0.5 * 64.0 / N
        g = xr.make_geometry(SDD, SOD, DET, N, U, P, vox_size=vs)
        scan = xr.Scan(cond_R=R)
        vol, _ = ellipsoid_phantom(orc, xr, N, vs, 300 + i)
        plan = ctx.plan(g, scan)
        ref = torch.from_numpy(vol.reshape(-1)).to(torch_dev)
        col = torch.empty_like(ref)
        ctx.to_columns(ref, col, N)
        proj = torch.zeros(P * U * U, dtype=torch.float32, device=torch_dev)
        plan.project(col, proj, plan.codes(range(P)), out_mode=xr.OUT_NORMALISED)
        ctx.synchronize()
        plan.close()
        geoms.append(g)
        scans.append(scan)
        sinos.append(proj.cpu().numpy())
        batch.append(b)
    epochs = 2
    init = [np.full(g.n_vox ** 3, 0.01, np.float32) for g in geoms]
    want = []
    for g, scan, sino, b, v0 in zip(geoms, scans, sinos, batch, init):
        v = [v0.copy()]
        ctx.ir_reconstruct(g, scan, [sino.copy()], v, epochs, b, method=m, lam=0.5)
        want.append(v[0])
    got = [v0.copy() for v0 in init]
    ctx.batch_reconstruct(geoms, [s.copy() for s in sinos], got, epochs, batch=batch, method=m, lam=0.5,
                          scans=scans)
    for i in range(len(specs)):
        assert np.array_equal(got[i], want[i]), "problem %d differs from its standalone solve" % i


@pytest.mark.parametrize("tilt,det", [(0.0, 1.5), (20.0, 1.5), (0.0, 0.75)])
def test_forward_work_split_is_exact(orc, xr, ctx, torch_dev, tilt, det):
    """The forward's band split (fwd_split 1, 2, 4 and the automatic choice for
    few views) and its per-row candidate table (fwd_scan 0) change only the work
    split: every view set projects bitwise as the unsplit whole-row scan, for
    untilted views (table + split), tilted ones (split) and a detector narrower
    than the volume's shadow (truncated rows)."""
    import torch

    N, P = 40, 24
    U = int(det * N)
    vs = 0.5 * 64.0 / N
    g = xr.make_geometry(SDD, SOD, DET, N, U, P, vox_size=vs)
    scan = xr.Scan(cond_R=rot_x(tilt).reshape(-1) if tilt else None)
    vol, _ = ellipsoid_phantom(orc, xr, N, vs, 77)
    plan = ctx.plan(g, scan)
    ref = torch.from_numpy(vol.reshape(-1)).to(torch_dev)
    col = torch.empty_like(ref)
    ctx.to_columns(ref, col, N)
    y = torch.zeros(P * U * U, dtype=torch.float32, device=torch_dev)
    plan.project(col, y, plan.codes(range(P)), out_mode=xr.OUT_NORMALISED)
    for views in (list(range(P)), [3, 11, 17], [5]):
        codes = plan.codes(views)
        base = None
        for scan_mode, split in ((1, 1), (0, 1), (0, 2), (0, 4), (0, 0), (1, 4)):
            plan.set_tiles(fwd_scan=scan_mode, fwd_split=split)
            p = torch.zeros(P * U * U, dtype=torch.float32, device=torch_dev)
            loss = plan.project(col, p, codes, out_mode=xr.OUT_RATIO, sino=y)
            ctx.synchronize()
            if base is None:
                base, base_loss = p.clone(), loss
                assert float(p.abs().sum()) > 0
            assert torch.equal(p, base), "views %s: fwd_scan=%d fwd_split=%d changed the projection" % (
                views, scan_mode, split)
            assert abs(loss - base_loss) <= 1e-9 * max(1.0, abs(base_loss))
    plan.close()


@pytest.mark.parametrize("method,subsets", [("mlem", 1), ("art", 1), ("mlem", 4), ("art", 3)])
def test_solver_max_bound_equals_measured(orc, xr, ctx, torch_dev, method, subsets):
    """The solver takes max |voxel| for the forward's fixed-point scale from its
    previous backprojection instead of a pass over the volume.  Three epochs of
    xr_ir_reconstruct with `subsets` ordered subsets equal the same subset
    steps (same order: libstdc++ std::shuffle with one default-seeded
    mt19937_64 per call, reconstruct.cu:61-89) driven through xr_project /
    xr_backproject_update, which measure the maximum from the volume at every
    call, bit for bit."""
    import torch

    m = xr.METHOD_MLEM if method == "mlem" else xr.METHOD_ART
    out_mode = xr.OUT_RATIO if method == "mlem" else xr.OUT_RESIDUAL
    N, P = 40, 24
    U = (3 * N) // 2
    vs = 0.5 * 64.0 / N
    g = xr.make_geometry(SDD, SOD, DET, N, U, P, vox_size=vs)
    vol, _ = ellipsoid_phantom(orc, xr, N, vs, 91)
    plan = ctx.plan(g, xr.Scan())
    codes = plan.codes(range(P))
    ref = torch.from_numpy(vol.reshape(-1)).to(torch_dev)
    col = torch.empty_like(ref)
    ctx.to_columns(ref, col, N)
    y = torch.zeros(P * U * U, dtype=torch.float32, device=torch_dev)
    plan.project(col, y, codes, out_mode=xr.OUT_NORMALISED)
    ctx.synchronize()
    epochs = 3
    assert P % subsets == 0  # no view repeats within an epoch
    ss = P // subsets
    v0 = np.full(N ** 3, 0.01, np.float32)
    solved = [v0.copy()]
    ctx.ir_reconstruct(g, xr.Scan(), [y.cpu().numpy()], solved, epochs, subsets, method=m, lam=0.5)
    cur = torch.empty_like(ref)
    ctx.to_columns(torch.from_numpy(v0).to(torch_dev), cur, N)
    orders = orc.subset_orders(subsets, epochs, False)
    for ep in range(epochs):
        p = torch.zeros(P * U * U, dtype=torch.float32, device=torch_dev)
        for sub in orders[ep]:
            sc = plan.codes([(int(sub) + subsets * j) % P for j in range(ss)])
            plan.project(cur, p, sc, out_mode=out_mode, sino=y)
            plan.backproject_update(p, sc, cur, method=m, lam=0.5, subset_size=ss)
    out = torch.empty_like(cur)
    ctx.from_columns(cur, out, N)
    ctx.synchronize()
    got = out.cpu().numpy()
    assert float(np.abs(got).max()) > 0
    assert np.array_equal(got, solved[0]), "max |diff| %g" % float(np.max(np.abs(got - solved[0])))
    plan.close()
