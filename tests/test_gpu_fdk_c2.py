"""FDK at the 512^3 / 720-view / 768^2 workload against the FP64 oracle:
the Ram-Lak and reference filters on sampled views (calcWeight + projConv,
fdk.cu:13-43) and the weighted backprojection of sampled views
(fdk.cu:51-170), <= 1e-5 relative L2."""
import numpy as np
import pytest

from conftest import rel_l2
from test_gpu_parity import record

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2(orc, xr, ctx):
    import torch

    dev = torch.device("cuda:0")
    N, U, P = 512, 768, 720
    sdd, sod, det = 1000.0, 500.0, 0.4
    geom = xr.make_geometry(sdd, sod, det, N, U, P)
    vs = geom.vox_size
    params = xr.ellipsoid_params(1234, 40, fov_radius=0.5 * N * vs)
    vref = torch.zeros(N ** 3, device=dev)
    ctx.phantom_rasterise(params, N, vs, vref)
    vcol = torch.empty_like(vref)
    ctx.to_columns(vref, vcol, N)
    del vref
    plan = ctx.plan(geom, xr.Scan())
    sino = torch.zeros(P * U * U, device=dev)
    views = [0, 181, 359, 600]
    plan.project(vcol, sino, plan.codes(views), out_mode=xr.OUT_NORMALISED)
    ctx.synchronize()
    plan.close()
    del vcol
    case = orc.Case(sdd, sod, det, N, U, P)
    return dict(geom=geom, sino=sino, views=views, case=case)


@pytest.mark.parametrize("kind", [1, 0])
def test_c2_filter_sampled_views(orc, xr, ctx, c2, kind):
    import torch

    geom, sino, views = c2["geom"], c2["sino"], c2["views"]
    U, P = geom.n_det_u, geom.n_proj
    fplan = ctx.plan(geom, xr.Scan(), convention=xr.CONV_FDK)
    filt = torch.zeros_like(sino)
    fplan.fdk_filter(sino, filt, np.array(views, np.int32), filter=kind)
    ctx.synchronize()
    fv = filt.view(P, U, U)
    sv = sino.view(P, U, U)
    worst = 0.0
    for v in views[:2]:
        want = c2["case"].fdk_filter_view(sv[v].cpu().numpy(), kind=kind)
        worst = max(worst, rel_l2(fv[v].cpu().numpy(), want))
    record("c2_fdk_filter/%d" % kind, worst)
    assert worst <= 1e-5
    fplan.close()


def test_c2_backprojection_sampled_views(orc, xr, ctx, c2):
    """The weighted backprojection of two Ram-Lak filtered views of the
    phantom's sinogram.  At 512^3 the FP32 pixel coordinate (u ~ 768) carries
    a rounding of ~6e-5 pixel into every bilinear weight, so the reference's
    own FP32 arithmetic (ref32) is compared with FP64 too."""
    import torch

    geom, sino, views = c2["geom"], c2["sino"], c2["views"]
    N, U, P = geom.n_vox, geom.n_det_u, geom.n_proj
    picks = views[:2]
    fplan = ctx.plan(geom, xr.Scan(), convention=xr.CONV_FDK)
    filt = torch.zeros_like(sino)
    fplan.fdk_filter(sino, filt, np.array(picks, np.int32), filter=xr.FILTER_RAMLAK)
    vol = torch.zeros(N ** 3, device="cuda")
    fplan.fdk_backproject(filt, np.array(picks, np.int32), vol)
    ref = torch.empty_like(vol)
    ctx.from_columns(vol, ref, N)
    ctx.synchronize()
    got = ref.cpu().numpy()
    imgs = [filt.view(P, U, U)[v].cpu().numpy() for v in picks]
    del vol, ref, filt
    want = np.zeros(N ** 3)
    want32 = np.zeros(N ** 3)
    for img, v in zip(imgs, picks):
        c2["case"].fdk_backproject_view(img, v, want)
        c2["case"].fdk_backproject_view(img, v, want32, prec=orc.REF32)
    err, ref_err = rel_l2(got, want), rel_l2(want32, want)
    record("c2_fdk_backprojection", err)
    record("c2_fdk_backprojection_ref32", ref_err)
    assert err <= max(1e-5, 2.0 * ref_err)
    fplan.close()
