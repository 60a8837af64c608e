"""Directional (XTT) reconstruction at its full workload: 256^3 x 13
coefficients, 4 grating orientations x 360 views, 512^2 detector.

The XTT forward (forwardProjXTT + correlationProjByLength, ir.cu:1753-1804,
1533-1603) of sampled views of every condition is compared with the FP64
oracle at <= 1e-5 relative L2; the raw XTT backprojection (backwardProjXTT,
ir.cu:1807-1850) of sampled views per channel likewise.  A short
XTT::reconstruct (reconstruct.cu:799-954) with PCA (pca.cu:12-65) runs on a
smaller volume of the same scan against the oracle at <= 1e-4.
"""
import numpy as np
import pytest

from conftest import rel_l2
from test_gpu_parity import Setup, proj_error, record

pytestmark = pytest.mark.gpu

PHIS = (0.0, 45.0, 90.0, 135.0)


@pytest.fixture(scope="module")
def wl(xr):
    import importlib

    return importlib.import_module(xr.__name__ + ".workloads")


@pytest.fixture(scope="module")
def c3(xr, ctx, wl):
    import torch

    geom, scan, plan, truth, sino, axes = wl.xtt_workload(xr, ctx, noise=False)
    K, N = scan.n_basis, geom.n_vox
    ref = torch.empty_like(truth)
    ctx.from_columns(truth, ref, N, K)
    ctx.synchronize()
    out = dict(geom=geom, scan=scan, plan=plan, truth=truth, sino=sino, ref=ref.cpu().numpy().reshape(K, N, N, N))
    yield out
    plan.close()
    del out


def c3_case(orc, wl, geom):
    return orc.Case(geom.sdd, geom.sod, geom.det_size, geom.n_vox, geom.n_det_u, geom.n_proj,
                    vox_size=geom.vox_size, cond_R=np.tile(np.eye(3).reshape(-1), len(PHIS)),
                    basis=wl.XTT_BASIS13, grating=wl.grating(PHIS))


def test_c3_forward_sampled_views(orc, xr, ctx, wl, c3):
    geom = c3["geom"]
    case = c3_case(orc, wl, geom)
    P, U = geom.n_proj, geom.n_det_u
    sino = c3["sino"].view(len(PHIS), P, U, U)
    worst = 0.0
    for cond, views in ((0, (0, 97)), (1, (211,)), (2, (45, 300)), (3, (359,))):
        for v in views:
            print("c3 forward cond %d view %d" % (cond, v), flush=True)
            got = sino[cond, v].cpu().numpy()
            want, fac = case.forward_view(c3["ref"], v, cond=cond, xtt=True)
            want32, _ = case.forward_view(c3["ref"], v, cond=cond, xtt=True, prec=orc.REF32)
            worst = max(worst, proj_error(got, want, want32, fac))
    record("c3_xtt_forward_f64", worst)
    assert worst <= 1e-5


def test_c3_backprojection_sampled_views(orc, xr, ctx, wl, c3, torch_dev):
    import torch

    geom, scan, plan = c3["geom"], c3["scan"], c3["plan"]
    case = c3_case(orc, wl, geom)
    K, N, P, U = scan.n_basis, geom.n_vox, geom.n_proj, geom.n_det_u
    N3 = N ** 3
    rng = np.random.default_rng(33)
    picks = [(0, 13), (1, 190), (2, 77), (3, 301)]
    img = torch.zeros(len(PHIS) * P * U * U, device=torch_dev)
    imgs = rng.random((len(picks), U, U)).astype(np.float32)
    iv = img.view(len(PHIS), P, U, U)
    for i, (c, v) in enumerate(picks):
        iv[c, v] = torch.from_numpy(imgs[i]).to(torch_dev)
    codes = np.array([c * P + v for c, v in picks], np.int32)
    tmp = torch.zeros(K * N3, device=torch_dev)
    fac = torch.zeros(K * N3, device=torch_dev)
    print("c3 backprojection of %d views" % len(picks), flush=True)
    plan.backproject(img, codes, tmp, fac, xtt=True)
    tref, fref = torch.empty_like(tmp), torch.empty_like(fac)
    ctx.from_columns(tmp, tref, N, K)
    ctx.from_columns(fac, fref, N, K)
    ctx.synchronize()
    del tmp, fac
    got_t = tref.cpu().numpy().reshape(K, N3)
    got_f = fref.cpu().numpy().reshape(K, N3)
    del tref, fref
    wt = np.zeros(K * N3)
    wf = np.zeros(K * N3)
    for i, (c, v) in enumerate(picks):
        case.backward_view(imgs[i], v, wt, wf, cond=c, xtt=True)
    wt = wt.reshape(K, N3)
    wf = wf.reshape(K, N3)
    et = max(rel_l2(got_t[k], wt[k]) for k in range(K))
    ef = max(rel_l2(got_f[k], wf[k]) for k in range(K))
    record("c3_xtt_backward_tmp_worst_channel", et)
    record("c3_xtt_backward_factor_worst_channel", ef)
    assert et <= 1e-5 and ef <= 1e-5


def test_c3_scan_solver_and_pca(orc, xr, ctx, wl):
    """XTT::reconstruct with PCA on a 48^3 volume of the same 4-grating scan."""
    K = 13
    st = Setup(orc, xr, N=48, U=96, P=60, vs=0.4, sdd=1000.0, sod=500.0, det=0.8,
               cond_R=np.tile(np.eye(3).reshape(-1), len(PHIS)), basis=wl.XTT_BASIS13, grating=wl.grating(PHIS))
    fp, _ = xr.fibre_params(77, 30, n_axes=3, kappa=20.0, fov_radius=0.5 * st.N * st.vs, centre_frac=0.55,
                            r_lo=2 * st.vs, r_hi=6 * st.vs, half_length=0.35 * st.N * st.vs)
    truth = (orc.phantom_fibres(st.N, st.vs, fp, wl.XTT_BASIS13) * 0.02).astype(np.float32)
    sino = np.stack([np.stack([st.case.forward_view(truth, n, cond=c, xtt=True)[0] for n in range(st.P)])
                     for c in range(len(PHIS))]).astype(np.float32)
    vol0 = np.full((K, st.N ** 3), 0.01, dtype=np.float32)
    want, want_proj, wloss = st.case.xtt_reconstruct(sino, vol0, 2, 4)
    v = [vol0[k].copy() for k in range(K)]
    s = [sino[c].reshape(-1).copy() for c in range(len(PHIS))]
    md = [np.zeros(st.N ** 3, np.float32) for _ in range(3)]
    ev = [np.zeros(st.N ** 3, np.float32) for _ in range(3)]
    losses = ctx.xtt_reconstruct(st.geom, st.scan, s, v, md, 2, 4, evalue=ev)
    err = rel_l2(np.stack(v), want.reshape(K, -1))
    record("c3_scan_xtt_solver", err)
    assert err <= 1e-4
    assert rel_l2(np.stack(s), want_proj.reshape(len(PHIS), -1)) <= 1e-4
    assert np.allclose(losses, wloss, rtol=1e-4)
    omd, oev = orc.pca(np.stack(v), wl.XTT_BASIS13)
    valid = np.isfinite(oev).all(axis=0) & (np.abs(omd).sum(axis=0) > 1e-6)
    assert valid.sum() > 1000
    assert rel_l2(np.stack(md)[:, valid], omd[:, valid]) <= 1e-4
