"""XTT problems in the batched persistent launches.  Each XTT problem of a
batch (its own volume size, view count, grating conditions and basis count)
must equal its standalone xr_xtt_reconstruct solve (XTT::reconstruct,
reconstruct.cu:799-954: per-epoch subset engine, condition loop inside each
subset, voxelSqrt, the sinogram overwritten with the final ratios) bit for
bit, next to IR problems in the same batch."""
import numpy as np
import pytest

from conftest import ellipsoid_phantom

pytestmark = pytest.mark.gpu


def xtt_problem(orc, xr, wl, N, U, P, phis, kidx, seed):
    basis = wl.XTT_BASIS13[list(kidx)]
    scan = xr.Scan(cond_R=np.tile(np.eye(3).reshape(-1), len(phis)), basis=basis, grating=wl.grating(phis))
    vs = 0.5
    geom = xr.make_geometry(600.0, 300.0, 0.8, N, U, P, vox_size=vs)
    case = orc.Case(600.0, 300.0, 0.8, N, U, P, vox_size=vs, cond_R=np.tile(np.eye(3).reshape(-1), len(phis)),
                    basis=basis, grating=wl.grating(phis))
    fp, _ = xr.fibre_params(seed, 10, fov_radius=0.5 * N * vs, r_lo=2 * vs, r_hi=4 * vs)
    truth = (orc.phantom_fibres(N, vs, fp, basis) * 0.02).astype(np.float32)
    sino = np.stack([np.stack([case.forward_view(truth, n, cond=c, xtt=True)[0] for n in range(P)])
                     for c in range(len(phis))]).astype(np.float32)
    return geom, scan, sino


@pytest.mark.parametrize("method", ["mlem", "art"])
def test_batched_xtt_equals_standalone(orc, xr, ctx, method):
    import importlib

    wl = importlib.import_module(xr.__name__ + ".workloads")
    m = xr.METHOD_MLEM if method == "mlem" else xr.METHOD_ART
    lam = 0.05
    specs = [(24, 40, 18, (0.0, 45.0, 90.0, 135.0), range(13), 3), (20, 32, 12, (0.0, 90.0), (0, 1, 2, 3, 4, 9, 11), 2)]
    xtt = [xtt_problem(orc, xr, wl, N, U, P, phis, k, 60 + i) for i, (N, U, P, phis, k, _) in enumerate(specs)]
    plain = []
    for i, (N, P) in enumerate([(32, 20), (16, 12)]):
        U = (3 * N) // 2
        g = xr.make_geometry(600.0, 300.0, 0.8, N, U, P, vox_size=0.5)
        vol, _ = ellipsoid_phantom(orc, xr, N, 0.5, 70 + i)
        case = orc.Case(600.0, 300.0, 0.8, N, U, P, vox_size=0.5)
        plain.append((g, xr.Scan(), np.stack([case.forward_view(vol, n)[0] for n in range(P)]).astype(np.float32)))
    epochs = 2
    subsets = [s[5] for s in specs] + [2, 1]
    geoms = [p[0] for p in xtt] + [p[0] for p in plain]
    scans = [p[1] for p in xtt] + [p[1] for p in plain]
    sinos = [p[2].reshape(-1).copy() for p in xtt + plain]
    init = [np.full(g.n_vox ** 3 * max(s.n_basis, 1), 0.01, np.float32) for g, s in zip(geoms, scans)]
    got = [v.copy() for v in init]
    got_sino = [s.copy() for s in sinos]
    losses = ctx.batch_reconstruct(geoms, got_sino, got, epochs, batch=subsets, method=m, lam=lam, scans=scans)
    for i, (g, sc, sino) in enumerate(xtt):
        K, N3 = sc.n_basis, g.n_vox ** 3
        v = [init[i][k * N3:(k + 1) * N3].copy() for k in range(K)]
        s = [sino[c].reshape(-1).copy() for c in range(sc.n_cond)]
        md = [np.zeros(N3, np.float32) for _ in range(3)]
        want_loss = ctx.xtt_reconstruct(g, sc, s, v, md, epochs, subsets[i], method=m, lam=lam)
        print("xtt problem %d max |diff| %g" % (i, float(np.max(np.abs(got[i] - np.concatenate(v))))), flush=True)
        assert np.array_equal(got[i], np.concatenate(v)), i
        assert np.array_equal(got_sino[i], np.concatenate(s)), i  # final ratio / residual images
        np.testing.assert_allclose(losses[:, i], want_loss, rtol=1e-9)
    for j, (g, sc, sino) in enumerate(plain):
        i = len(xtt) + j
        v = [init[i].copy()]
        want_loss = ctx.ir_reconstruct(g, sc, [sino.reshape(-1).copy()], v, epochs, subsets[i], method=m, lam=lam)
        assert np.array_equal(got[i], v[0]), i
        np.testing.assert_allclose(losses[:, i], want_loss, rtol=1e-9)
