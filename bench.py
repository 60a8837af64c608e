"""Benchmark: one plain-MLEM iteration at 512^3 / 720 views / 768^2 detector.

A step is one full MLEM iteration (batch = 1, i.e. one subset holding all 720
views): the fused forward projection + L/(factor+1e-5) normalisation + MLEM
ratio (one launch over all views) followed by the fused backprojection +
multiplicative update (one launch).  Inputs are resident in HBM before timing.

Prints one JSON line.  value = GUPS over the step = 2 * N^3 * P voxel-view
updates (forward + back) per second / 1e9.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-residual 1e-3 (learned warm-start vs plain MLEM) 512³; GUPS fwd/back"
HBM_PEAK_GBS = 8000.0
BYTES_PER_VOXEL_VIEW = 4.0  # one FP32 volume stream per view (plain forward / back)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(N, U, P, sdd, sod, det, vs, n_views=16, extras=True):
    """Time the oracle's FP32 restatement (the reference's operators on the CPU)
    on a bounded sample: forward + back of n_views views of the same workload
    (value, GUPS).  With `extras`: the 64^3 / 90-view / 96^2 configuration
    timed in full (20 MLEM iterations, batch 1) and the per-voxel PCA at
    256^3 x 13 coefficients."""
    import __graft_entry__ as entry

    orc = entry.load_oracle()
    xr = entry.load_package()
    case = orc.Case(sdd, sod, det, N, U, P, vox_size=vs)
    params = xr.ellipsoid_params(1234, 40, fov_radius=0.5 * N * vs)
    vol = orc.phantom_ellipsoids(N, vs, params)
    tmp = np.zeros(N ** 3)
    fac = np.zeros(N ** 3)
    img = np.ones((U, U), np.float32)
    views = [int(round(i * P / n_views)) for i in range(n_views)]
    t0 = time.perf_counter()
    for v in views:
        case.forward_view(vol, v, prec=orc.REF32)
        case.backward_view(img, v, tmp, fac, prec=orc.REF32)
    dt = time.perf_counter() - t0
    gups = 2.0 * N ** 3 * len(views) / dt / 1e9
    out = {"value": gups, "unit": "GUPS", "cores": int(orc.num_threads()), "kind": "port",
           "cpu_model": cpu_model(),
           "sample": "FP32 oracle forward+back of %d of %d views at %d^3, %dx%d detector (%.1f s)"
                     % (len(views), P, N, U, U, dt)}
    if extras:
        del vol, tmp, fac
        # 64^3 / 90 views / 96^2, 20 MLEM iterations (batch 1), timed in full
        c1 = orc.Case(600.0, 300.0, 0.8, 64, 96, 90, vox_size=0.5)
        truth = orc.phantom_ellipsoids(64, 0.5, xr.ellipsoid_params(1001, 10, fov_radius=16.0, value_lo=0.0,
                                                                   value_hi=0.05))
        sino = np.stack([c1.forward_view(truth, n, prec=orc.REF32)[0] for n in range(90)]).astype(np.float32)
        t0 = time.perf_counter()
        c1.ir_reconstruct(sino, np.full(64 ** 3, 0.01, np.float32), 20, 1, prec=orc.REF32)
        out["c1_mlem20_s"] = round(time.perf_counter() - t0, 3)
        # per-voxel PCA over 256^3 voxels x 13 coefficients
        import importlib

        wl = importlib.import_module(xr.__name__ + ".workloads")
        mu = np.random.default_rng(5).random((13, 256 ** 3), dtype=np.float32)
        t0 = time.perf_counter()
        orc.pca(mu, wl.XTT_BASIS13)
        out["pca_256_13_s"] = round(time.perf_counter() - t0, 3)
    return out


MODEL = os.path.join(ROOT, "xray-directional-reconstruction_amd", "models", "warmstart_c24.xunt")


def c2_noisy_sinogram(ctx, xr, geom, plan, N, vs, photons=1e4, seed=2002):
    """The 512^3 / 720-view workload's data: 40 seeded ellipsoids plus a
    cylinder shell, the normalised forward projection, Poisson noise at
    `photons` per ray and the driver's clamp (main.cu:35)."""
    import torch

    dev = torch.device("cuda:0")
    params = xr.ellipsoid_params(1234, 40, fov_radius=0.5 * N * vs)
    vref = torch.zeros(N ** 3, device=dev)
    ctx.phantom_rasterise(params, N, vs, vref)
    ctx.phantom_cylinder_shell(N, vs, 0.45 * N * vs, 4 * vs, 0.02, vref)
    truth = torch.empty_like(vref)
    ctx.to_columns(vref, truth, N)
    del vref
    P, U, V = geom.n_proj, geom.n_det_u, geom.n_det_v
    p = torch.zeros(P * U * V, device=dev)
    plan.project(truth, p, plan.codes(range(P)), out_mode=xr.OUT_NORMALISED)
    del truth
    y = torch.empty_like(p)
    ctx.noise_poisson(y, p, photons, seed)
    ctx.sino_clamp(y)
    ctx.synchronize()
    return y


def ttr_frozen(ctx, xr, geom, sino, tau=1e-3, subsets=80, long_epochs=60, pres=(2,), model=None):
    """Time to the frozen relative residual criterion (warmstart.plain_curve /
    warm_curve): R_inf from `long_epochs` plain epochs, plain epochs to
    (R_k - R_inf)/(R_0 - R_inf) <= tau, then the warm start for each pre."""
    import importlib

    ws = importlib.import_module(xr.__name__ + ".warmstart")
    scene = ws.Scene(ctx, geom, sino, subsets=subsets)
    R, ms = ws.plain_curve(scene, long_epochs)
    R0, Rinf = R[0], R[-1]
    k = ws.first_within(R, R0, Rinf, tau)
    out = {"tau": tau, "criterion": "(R_k - R_inf)/(R_0 - R_inf), R = ||y - Ax||^2/||y||^2 (FP64), "
                                    "R_inf after %d plain epochs" % long_epochs,
           "method": "MLEM, %d ordered subsets" % subsets, "R0": R0, "Rinf": Rinf,
           "plain_epochs": k, "plain_s": None if k is None else round(ms[k] / 1e3, 3),
           "plain_R": [round(r, 9) for r in R[:40]], "epoch_ms": round(ms[-1] / long_epochs, 2)}
    if model and os.path.exists(model):
        net = xr.UNet(ctx, 24)
        net.load(model)
        train_n = ws.trained_edge(model)
        if train_n and (train_n > geom.n_vox or geom.n_vox % train_n):
            train_n = None
        out["network"] = {"file": os.path.relpath(model, ROOT), "trained_at": train_n,
                          "applied_at": train_n or geom.n_vox}
        out["warm"] = []
        for pre in pres:
            w = ws.warm_curve(scene, net, R0, Rinf, tau, long_epochs, pre=pre, train_n=train_n)
            w["R"] = [round(r, 9) for r in w["R"]]
            w["pre"] = pre
            w["s"] = round(w.pop("ms") / 1e3, 3)
            if k is not None and w["reached"]:
                w["speedup_time"] = round(ms[k] / 1e3 / w["s"], 3)
                w["speedup_epochs"] = round(k / max(w["epochs"], 1), 3)
            out["warm"].append(w)
        net.close()
    scene.close()
    return out


def ttr_workload(xr, args):
    """Time to the frozen residual criterion on the noisy 512^3 / 720-view data."""
    N, U, P = args.n_vox, args.n_det, args.n_proj
    sdd, sod, det = 1000.0, 500.0, 0.4
    vs = xr.default_vox_size(sdd, sod, det)
    geom = xr.make_geometry(sdd, sod, det, N, U, P)
    ctx = xr.Context(0)
    plan = ctx.plan(geom, xr.Scan())
    sino = c2_noisy_sinogram(ctx, xr, geom, plan, N, vs, photons=args.photons)
    plan.close()
    pres = tuple(int(v) for v in args.pres.split(","))
    r = ttr_frozen(ctx, xr, geom, sino, tau=args.tau, subsets=args.subsets or 1,
                   long_epochs=args.long_epochs, pres=pres, model=args.model or MODEL)
    r["data"] = "synthetic (40 seeded ellipsoids plus a cylinder shell, Poisson noise at %g photons)" % args.photons
    return r


HELDOUT = ((9101, "mixed"), (9102, "fibres"), (9103, "mixed"))


def heldout_workload(xr, args):
    """The frozen-criterion comparison (ttr_frozen) on seeded phantoms the
    network was not trained on (warmstart.phantom_spec draws; the committed
    network saw seeds 1000-2999), at the bench geometry, with Poisson noise
    at --photons per ray."""
    import importlib

    import torch

    ws = importlib.import_module(xr.__name__ + ".warmstart")
    N, U, P = args.n_vox, args.n_det, args.n_proj
    sdd, sod, det = 1000.0, 500.0, 0.4
    vs = xr.default_vox_size(sdd, sod, det)
    geom = xr.make_geometry(sdd, sod, det, N, U, P)
    ctx = xr.Context(0)
    pres = tuple(int(v) for v in args.pres.split(","))
    out = []
    for seed, family in HELDOUT:
        spec = ws.phantom_spec(seed, N, vs, family)
        vol = ws.make_phantom(ctx, N, vs, seed, **spec)
        clean = ws.project(ctx, geom, vol)
        del vol
        sino = torch.empty_like(clean)
        ctx.noise_poisson(sino, clean, args.photons, seed)
        ctx.sino_clamp(sino)
        del clean
        r = ttr_frozen(ctx, xr, geom, sino, tau=args.tau, subsets=args.subsets or 1, long_epochs=args.long_epochs,
                       pres=pres, model=args.model or MODEL)
        r.pop("plain_R", None)
        r.update({"seed": seed, "family": family})
        out.append(r)
        del sino
        torch.cuda.empty_cache()
    return {"metric": "time-to-residual %g, plain MLEM vs learned warm start, held-out phantoms" % args.tau,
            "unit": "s", "n_gpus": 1,
            "data": "synthetic (warmstart.phantom_spec seeds %s, Poisson noise at %g photons)"
                    % ([s for s, _ in HELDOUT], args.photons),
            "config": {"workload": "%d^3 volume, %d views, %dx%d detector" % (N, P, U, U)}, "results": out}


def batch_workload(xr, args, n=48, epochs=30, seed=4):
    """48 independent MLEM problems with volumes drawn from {64, 96, 128, 192,
    256}^3 and {90, 180, 360} views, detector 1.5x the volume edge, seeded
    ellipsoid phantoms, 30 epochs each: one persistent forward launch and one
    backprojection launch per z-depth class per epoch for all problems, the
    epoch loop replayed as a HIP graph, tile shapes from the autotuner.
    Compared with the same problems solved one after another (one-problem
    batches, also graph-replayed)."""
    import torch

    rng = np.random.default_rng(seed)
    Ns = [int(v) for v in rng.choice([64, 96, 128, 192, 256], n)]
    Ps = [int(v) for v in rng.choice([90, 180, 360], n)]
    ctx = xr.Context(0)
    vs = xr.default_vox_size(1000.0, 500.0, 0.4)
    geoms = [xr.make_geometry(1000.0, 500.0, 0.4, N, 3 * N // 2, P, vox_size=vs) for N, P in zip(Ns, Ps)]
    sinos = []
    for i, (g, N) in enumerate(zip(geoms, Ns)):
        ref = torch.zeros(N ** 3, device="cuda")
        ctx.phantom_rasterise(xr.ellipsoid_params(100 + i, 10, fov_radius=0.5 * N * vs), N, vs, ref)
        col = torch.empty_like(ref)
        ctx.to_columns(ref, col, N)
        plan = ctx.plan(g, xr.Scan())
        sino = torch.zeros(g.n_proj * g.n_det_u * g.n_det_v, device="cuda")
        plan.project(col, sino, plan.codes(range(g.n_proj)))
        ctx.synchronize()
        plan.close()
        sinos.append(sino)

    def load(b, idx):
        for j, i in enumerate(idx):
            va, sa = b.buffers(j)
            ctx.copy(sa, sinos[i])
            ctx.copy(va, torch.full((Ns[i] ** 3,), 0.01, device="cuda"))
        b.restart()

    b = ctx.batch(geoms, flags=xr.BATCH_GRAPH | xr.BATCH_AUTOTUNE)
    load(b, range(n))
    b.run(epochs)  # autotune, capture, warm-up
    load(b, range(n))
    b.run(epochs)
    batched_ms = b.elapsed_ms()
    b.close()
    seq_ms = 0.0
    for i in range(n):
        one = ctx.batch([geoms[i]], flags=xr.BATCH_GRAPH | xr.BATCH_AUTOTUNE)
        load(one, [i])
        one.run(epochs)
        load(one, [i])
        one.run(epochs)
        seq_ms += one.elapsed_ms()
        one.close()
    vv = sum(2.0 * N ** 3 * P * epochs for N, P in zip(Ns, Ps))
    return {"metric": "batched uneven MLEM (48 problems x 30 epochs)", "value": round(vv / (batched_ms * 1e-3) / 1e9, 2),
            "unit": "GUPS", "n_gpus": 1, "batched_s": round(batched_ms / 1e3, 3),
            "sequential_s": round(seq_ms / 1e3, 3), "speedup_vs_sequential": round(seq_ms / batched_ms, 3),
            "problems": [[N, P] for N, P in zip(Ns, Ps)], "dtype": "f32",
            "data": "synthetic (10 seeded ellipsoids per problem, sinograms from the projector)",
            "config": {"workload": "48 MLEM problems, N in {64,96,128,192,256}, P in {90,180,360}, U = 1.5N, "
                                   "30 epochs, batch 1, persistent launches + HIP graph + autotuned tiles"}}


def xtt_measure(xr, steps=3, warmup=1, N=256, U=512, P=360):
    """Directional (XTT) operator pair at 256^3 x 13 coefficients, 4 grating
    orientations x 360 views, 512^2 detector: the fused forward + ratio launch
    over all 1440 (condition, view) images and the fused backprojection +
    per-channel MLEM update launch, plus the per-voxel PCA (calcEigenVector)
    over the whole volume.  G voxel-views/s = N^3 x conditions x views per pass
    / s; roofline on 52 B per voxel-view (13 FP32 coefficients per view)."""
    import importlib

    import torch

    wl = importlib.import_module(xr.__name__ + ".workloads")
    ctx = xr.Context(0)
    geom, scan, plan, truth, sino, _ = wl.xtt_workload(xr, ctx, N=N, U=U, P=P)
    K, C = scan.n_basis, scan.n_cond
    N3 = N ** 3
    vol = torch.full((K * N3,), 0.01, device="cuda")
    proj = torch.empty_like(sino)
    codes = np.arange(C * P, dtype=np.int32)

    def step():
        loss = plan.project(vol, proj, codes, out_mode=xr.OUT_RATIO, sino=sino, xtt=True)
        plan.backproject_update(proj, codes, vol, method=xr.METHOD_MLEM, subset_size=P, xtt=True)
        return loss

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ctx.timing(True)
    ctx.timing_reset()
    losses = [step() for _ in range(steps)]
    torch.cuda.synchronize()
    fwd_ms, fwd_n = ctx.timing_read(xr.TIMING_FORWARD)
    bwd_ms, bwd_n = ctx.timing_read(xr.TIMING_BACKWARD)
    # per-voxel principal orientation over the whole volume
    md = torch.empty(3 * N3, device="cuda")
    ev = torch.empty(3 * N3, device="cuda")
    ctx.timing_reset()
    for _ in range(steps):
        ctx.pca(vol, scan.basis.reshape(-1, 3), N3, N3, md, ev)
    torch.cuda.synchronize()
    pca_ms, pca_n = ctx.timing_read(xr.TIMING_OTHER)
    ctx.timing(False)
    vv = float(N3) * C * P
    fwd_avg, bwd_avg = fwd_ms / max(fwd_n, 1), bwd_ms / max(bwd_n, 1)
    dom_name, dom_ms = ("xtt backprojection+update", bwd_avg) if bwd_avg >= fwd_avg else ("xtt forward+ratio", fwd_avg)
    achieved = 4.0 * K * vv / (dom_ms * 1e-3) / 1e9
    pca_avg = pca_ms / max(pca_n, 1)
    out = {"workload": "XTT MLEM iteration, %d^3 x %d coefficients, %d conditions x %d views, %dx%d detector"
                       % (N, K, C, P, U, U),
           "gvv_forward": round(vv / (fwd_avg * 1e-3) / 1e9, 2), "gvv_backward": round(vv / (bwd_avg * 1e-3) / 1e9, 2),
           "unit": "G voxel-views/s", "forward_ms": round(fwd_avg, 3), "backward_ms": round(bwd_avg, 3),
           "loss_first_last": [losses[0], losses[-1]],
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel": dom_name, "kernel_ms": round(dom_ms, 3),
                        "algorithmic_bytes": "%d B per voxel-view x %d^3 voxels x %d views per launch"
                                             % (4 * K, N, C * P)},
           "pca_ms": round(pca_avg, 3),
           "pca_gbs": round((4 * K + 24) * N3 / (pca_avg * 1e-3) / 1e9, 1)}
    plan.close()
    ctx.close()
    return out


def fdk_art_measure(ctx, xr, geom, sino, art_iters=50, lam=1e-3):
    """FDK with the Ram-Lak filter (calcWeight + projConv + weighted
    backprojection, fdk.cu:13-170) over every view, and `art_iters` ART
    iterations (projSubtract + voxelPlus, ir.cu:1656-1693; batch 1, so the step
    is lambda / views, reconstruct.cu:124) on the same data, from zero.
    lambda = 1e-3 is the driver's value (main.cu:60): the update adds
    lambda/views times an unnormalised backprojection, which grows with the
    volume edge, and at 512^3 with one subset lambda = 0.2 already diverges.
    Roofline models: 8 B per filtered detector sample, 4 B per voxel-view."""
    import torch

    N, U, V, P = geom.n_vox, geom.n_det_u, geom.n_det_v, geom.n_proj
    codes = np.arange(P, dtype=np.int32)
    fplan = ctx.plan(geom, xr.Scan(), convention=xr.CONV_FDK)
    filt = torch.empty_like(sino)
    vol = torch.empty(N ** 3, device=sino.device)
    fplan.fdk_filter(sino, filt, codes, filter=xr.FILTER_RAMLAK)  # warm-up
    fplan.fdk_backproject(filt, codes, vol)
    torch.cuda.synchronize()
    ctx.timing(True)
    ctx.timing_reset()
    fplan.fdk_filter(sino, filt, codes, filter=xr.FILTER_RAMLAK)
    fplan.fdk_backproject(filt, codes, vol)
    torch.cuda.synchronize()
    f_ms, _ = ctx.timing_read(xr.TIMING_FILTER)
    b_ms, _ = ctx.timing_read(xr.TIMING_BACKWARD)
    fplan.close()
    del filt
    # ART from zero
    plan = ctx.plan(geom, xr.Scan())
    vol.zero_()
    resid = torch.empty_like(sino)
    plan.project(vol, resid, codes, out_mode=xr.OUT_RESIDUAL, sino=sino)  # tables, warm-up
    torch.cuda.synchronize()
    ctx.timing_reset()
    t0 = time.perf_counter()
    losses = []
    for _ in range(art_iters):
        losses.append(plan.project(vol, resid, codes, out_mode=xr.OUT_RESIDUAL, sino=sino))
        plan.backproject_update(resid, codes, vol, method=xr.METHOD_ART, lam=lam, subset_size=P)
    torch.cuda.synchronize()
    art_s = time.perf_counter() - t0
    af_ms, af_n = ctx.timing_read(xr.TIMING_FORWARD)
    ab_ms, ab_n = ctx.timing_read(xr.TIMING_BACKWARD)
    ctx.timing(False)
    plan.close()
    vv = float(N) ** 3 * P
    return {"fdk": {"filter": "ram-lak", "filter_ms": round(f_ms, 3), "backprojection_ms": round(b_ms, 3),
                    "total_ms": round(f_ms + b_ms, 3),
                    "filter_gbs": round(8.0 * U * V * P / (f_ms * 1e-3) / 1e9, 1),
                    "backprojection_gups": round(vv / (b_ms * 1e-3) / 1e9, 2),
                    "backprojection_roofline_frac": round(4.0 * vv / (b_ms * 1e-3) / 8e12, 4)},
            "art": {"iterations": art_iters, "lambda": lam, "s": round(art_s, 3),
                    "forward_ms": round(af_ms / max(af_n, 1), 3), "backward_ms": round(ab_ms / max(ab_n, 1), 3),
                    "gups": round(2.0 * vv * art_iters / art_s / 1e9, 2),
                    "loss_first_last": [losses[0], losses[-1]]}}


def grad_measure(ctx, xr, geom, sino, lam=1e-3):
    """The two GPU stages of IR::gradReconstruct (reconstruct.cu:158-328) at
    this geometry: one OS-ART epoch (batch 1) of one gradient slot on the
    (N+1)^3 gradient grid (XR_CONV_GRAD plan: forwardProjGrad + projSubtract,
    backwardProjGrad + voxelPlus), and the Poisson integration of the N^3
    result (poissonSolveLDLT, poisson_cpu.cpp:15-112) by four FP64 sine-
    transform products per z slice: 8 (N-2)^3 flops per slice; then the whole
    entry point (three slots of one epoch and the integration) on the same
    data."""
    import torch

    N, U, V, P = geom.n_vox, geom.n_det_u, geom.n_det_v, geom.n_proj
    N1 = N + 1
    g1 = xr.make_geometry(geom.sdd, geom.sod, geom.det_size, N1, U, P, vox_size=geom.vox_size, n_det_v=V)
    codes = np.arange(P, dtype=np.int32)
    plan = ctx.plan(g1, xr.Scan(), convention=xr.CONV_GRAD)
    vol = torch.full((N1 ** 3,), 0.01, device=sino.device)
    resid = torch.empty_like(sino)
    plan.project(vol, resid, codes, out_mode=xr.OUT_RESIDUAL, sino=sino)  # tables, warm-up
    torch.cuda.synchronize()
    ctx.timing(True)
    ctx.timing_reset()
    plan.project(vol, resid, codes, out_mode=xr.OUT_RESIDUAL, sino=sino)
    plan.backproject_update(resid, codes, vol, method=xr.METHOD_ART, lam=lam, subset_size=P)
    torch.cuda.synchronize()
    f_ms, _ = ctx.timing_read(xr.TIMING_FORWARD)
    b_ms, _ = ctx.timing_read(xr.TIMING_BACKWARD)
    plan.close()
    del resid
    gx = torch.randn(N1 ** 3, device=sino.device)
    gy = torch.randn(N1 ** 3, device=sino.device)
    out = torch.empty(N ** 3, device=sino.device)
    ctx.poisson_solve(N, gx, gy, out)  # warm-up
    ctx.timing_reset()
    t0 = time.perf_counter()
    ctx.poisson_solve(N, gx, gy, out)
    call_s = time.perf_counter() - t0
    p_ms, _ = ctx.timing_read(xr.TIMING_OTHER)
    ctx.timing(False)
    del gx, gy, out, vol
    torch.cuda.empty_cache()
    # the entry point itself (xr_grad_reconstruct): three slots of one epoch
    # (batch 1) and the Poisson integration; with a block size (32) dividing N
    # the layers z >= N are held at zero and neither projector visits them
    sino_h = sino.cpu().numpy()
    rec = {}
    for epochs in (1, 2):  # the difference is one epoch of the three slots
        vol_h = np.zeros(N ** 3, np.float32)
        ctx.timing(True)
        ctx.timing_reset()
        t0 = time.perf_counter()
        ctx.grad_reconstruct(geom, xr.Scan(), sino_h, vol_h, epochs, 1, method=xr.METHOD_ART, lam=lam,
                             block_size=32)
        rec[epochs] = (time.perf_counter() - t0, ctx.timing_read(xr.TIMING_FORWARD)[0],
                       ctx.timing_read(xr.TIMING_BACKWARD)[0])
        ctx.timing(False)
    del sino_h, vol_h
    rec_s = rec[1][0]
    rf_ms = (rec[2][1] - rec[1][1]) / 3
    rb_ms = (rec[2][2] - rec[1][2]) / 3
    tables_ms = rec[1][1] - 3 * rf_ms  # the plan's weight pass (fixed-point bound of the forward)
    vv = float(N1) ** 3 * P
    flops = 8.0 * (N - 2) ** 3 * N
    return {"workload": "gradient slot on the %d^3 grid (%d views, %dx%d detector) + Poisson integration of %d^3"
                        % (N1, P, U, V, N),
            "slot_epoch_ms": round(f_ms + b_ms, 3), "forward_ms": round(f_ms, 3), "backward_ms": round(b_ms, 3),
            "slot_gups": round(2.0 * vv / ((f_ms + b_ms) * 1e-3) / 1e9, 2),
            "reconstruct": {"epochs": 1, "batch": 1, "block_size": 32, "wall_s": round(rec_s, 3),
                            "slot_epoch_ms": round(rf_ms + rb_ms, 3), "forward_ms": round(rf_ms, 3),
                            "backward_ms": round(rb_ms, 3), "plan_tables_ms": round(tables_ms, 3),
                            "slot_gups": round(2.0 * vv / ((rf_ms + rb_ms) * 1e-3) / 1e9, 2)},
            "poisson_kernels_ms": round(p_ms, 3), "poisson_call_ms": round(call_s * 1e3, 3),
            "poisson_tflops_f64": round(flops / (p_ms * 1e-3) / 1e12, 2),
            "poisson_frac_f64_peak": round(flops / (p_ms * 1e-3) / 78.6e12, 4)}


def unet_flops_per_voxel(C):
    """Algorithmic FLOP per full-resolution voxel of one U-Net inference: 2 x 27
    x Cin x Cout per convolution output voxel at its level (1, 1/8, 1/64 of the
    voxels), plus the 1x1x1 head; the MFMA padding of Cout to 16 is not
    counted."""
    lv0 = [(8, C), (C, C), (3 * C, C), (C, C)]
    lv1 = [(C, 2 * C), (2 * C, 2 * C), (6 * C, 2 * C), (2 * C, 2 * C)]
    lv2 = [(2 * C, 4 * C), (4 * C, 4 * C)]
    f = lambda layers: sum(54.0 * ci * co for ci, co in layers)
    return f(lv0) + f(lv1) / 8 + f(lv2) / 64 + 2.0 * C


def unet_measure(xr, N=512, reps=5):
    """One U-Net inference (the warm-start network, base width 24) on an N^3
    volume: decoder upsampling fused into the next convolution's fetch, and
    materialised (as training runs it).  Wall time per inference between
    stream synchronisations; FLOP/s against the 2.5 PFLOP/s BF16 dense peak."""
    import torch

    ctx = xr.Context(0)
    net = xr.UNet(ctx, 24)
    if os.path.exists(MODEL):
        net.load(MODEL)
    g = torch.Generator(device="cuda").manual_seed(11)
    x = torch.rand(N ** 3, device="cuda", generator=g) * 0.05
    y = torch.empty_like(x)
    flop = unet_flops_per_voxel(24) * float(N) ** 3
    out = {"workload": "U-Net inference (base width 24, 3 levels, BF16 MFMA) on a %d^3 volume" % N,
           "gflop": round(flop / 1e9, 1)}
    outs = {}
    for fused in (False, True):
        net.set_fused_upsampling(fused)
        net.forward(x, y, (N, N, N), scale=0.05)  # warm-up (buffers sized on first use)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            net.forward(x, y, (N, N, N), scale=0.05)
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        key = "fused" if fused else "materialised"
        out[key + "_ms"] = round(ms, 3)
        out[key + "_tflops"] = round(flop / (ms * 1e-3) / 1e12, 1)
        out[key + "_frac_bf16_peak"] = round(flop / (ms * 1e-3) / 2.5e15, 4)
        outs[key] = y.clone()
    out["bitwise_equal"] = bool(torch.equal(outs["fused"], outs["materialised"]))
    net.close()
    ctx.close()
    return out


def measured_traffic(kernel, N, P):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by a 1 GiB copy), if one
    matches this workload; else None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_%d_%d.md" % (N, P))
    if not os.path.exists(path):
        return None
    key = "traffic_bytes[%s]:" % kernel
    with open(path) as f:
        for line in f:
            if line.startswith(key):
                return float(line.split(":", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-vox", type=int, default=512)
    ap.add_argument("--n-det", type=int, default=768)
    ap.add_argument("--n-proj", type=int, default=720)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-views", type=int, default=16)
    ap.add_argument("--no-warmstart", action="store_true",
                    help="skip the time-to-residual comparison (learned warm start vs plain MLEM)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the FDK / ART, XTT and 48-problem batch measurements of the default line")
    ap.add_argument("--tau", type=float, default=1e-3)
    ap.add_argument("--subsets", type=int, default=0,
                    help="ordered subsets of the plain method (default 1: plain MLEM, every view per update)")
    ap.add_argument("--photons", type=float, default=1e4)
    ap.add_argument("--long-epochs", type=int, default=200)
    ap.add_argument("--pres", default="5")
    ap.add_argument("--model", default="")
    ap.add_argument("--workload", choices=["mlem", "batch", "heldout", "xtt", "ttr", "fdkart", "grad", "unet"],
                    default="mlem",
                    help="mlem: one MLEM iteration at 512^3/720 views (default); batch: 48 uneven MLEM problems, "
                         "30 epochs in persistent launches (HIP graph, autotuned tiles); heldout: time to "
                         "residual, plain vs warm start, on three phantoms outside the training draws")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 1 GiB device-to-device copy (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (views would shard "
              "across GPUs with an all-reduce of the backprojection).", file=sys.stderr)
        sys.exit(2)

    import torch

    import __graft_entry__ as entry

    xr = entry.load_package()
    if args.workload == "batch":
        print(json.dumps(batch_workload(xr, args)))
        return
    if args.workload == "heldout":
        print(json.dumps(heldout_workload(xr, args)))
        return
    if args.workload == "fdkart":
        N, U, P = args.n_vox, args.n_det, args.n_proj
        geom = xr.make_geometry(1000.0, 500.0, 0.4, N, U, P)
        ctx = xr.Context(0)
        plan = ctx.plan(geom, xr.Scan())
        sino = c2_noisy_sinogram(ctx, xr, geom, plan, N, geom.vox_size, photons=args.photons)
        plan.close()
        print(json.dumps(fdk_art_measure(ctx, xr, geom, sino)))
        return
    if args.workload == "grad":
        N, U, P = args.n_vox, args.n_det, args.n_proj
        geom = xr.make_geometry(1000.0, 500.0, 0.4, N, U, P)
        ctx = xr.Context(0)
        plan = ctx.plan(geom, xr.Scan())
        sino = c2_noisy_sinogram(ctx, xr, geom, plan, N, geom.vox_size, photons=args.photons)
        plan.close()
        print(json.dumps(grad_measure(ctx, xr, geom, sino)))
        return
    if args.workload == "ttr":
        print(json.dumps(ttr_workload(xr, args)))
        return
    if args.workload == "unet":
        print(json.dumps(unet_measure(xr, N=args.n_vox, reps=max(args.steps, 1))))
        return
    if args.workload == "xtt":
        print(json.dumps(xtt_measure(xr, steps=args.steps, warmup=args.warmup)))
        return
    dev = torch.device("cuda:0")
    N, U, P = args.n_vox, args.n_det, args.n_proj
    sdd, sod, det = 1000.0, 500.0, 0.4
    vs = xr.default_vox_size(sdd, sod, det)  # 0.2 mm (geometry.h:13)
    geom = xr.make_geometry(sdd, sod, det, N, U, P)
    scan = xr.Scan()
    ctx = xr.Context(0)
    plan = ctx.plan(geom, scan)
    N3 = N ** 3
    lenD = P * U * U

    # synthetic phantom (40 seeded ellipsoids plus a cylinder shell), its
    # sinogram by the same operator, Poisson noise at --photons per ray
    sino = c2_noisy_sinogram(ctx, xr, geom, plan, N, vs, photons=args.photons)
    codes = plan.codes(range(P))
    vol = torch.full((N3,), 0.01, device=dev)
    proj = torch.zeros(lenD, device=dev)
    torch.cuda.synchronize()

    def step():
        loss = plan.project(vol, proj, codes, out_mode=xr.OUT_RATIO, sino=sino)
        plan.backproject_update(proj, codes, vol, method=xr.METHOD_MLEM, subset_size=P)
        return loss

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.timing(True)
    ctx.timing_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    fwd_ms, fwd_n = ctx.timing_read(xr.TIMING_FORWARD)
    bwd_ms, bwd_n = ctx.timing_read(xr.TIMING_BACKWARD)
    ctx.timing(False)

    if args.calibrate_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        a.fill_(1.0)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    vv = float(N3) * P  # voxel-view updates per pass (one launch each way)
    ms_per_step = dt * 1e3 / args.steps
    gups = 2.0 * vv * args.steps / dt / 1e9
    fwd_avg, bwd_avg = fwd_ms / max(fwd_n, 1), bwd_ms / max(bwd_n, 1)
    dom_name, dom_ms = ("backprojection+update", bwd_avg) if bwd_avg >= fwd_avg else ("forward+ratio", fwd_avg)
    achieved = BYTES_PER_VOXEL_VIEW * vv / (dom_ms * 1e-3) / 1e9
    traffic = measured_traffic(dom_name, N, P)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom_name,
                "kernel_ms": round(dom_ms, 3), "launches_per_pass": 1,
                "algorithmic_bytes": "4 B per voxel-view x %d^3 voxels x %d views per launch" % (N, P)}
    out = {
        "metric": METRIC, "value": round(gups, 2), "unit": "GUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (40 seeded ellipsoids plus a cylinder shell, sinogram from the projector with Poisson "
                "noise at %g photons per ray)" % args.photons,
        "config": {"workload": "plain MLEM iteration (batch 1), %d^3 volume, %d views, %dx%d detector, "
                               "SOD 500 SDD 1000 det 0.4 mm" % (N, P, U, U),
                   "n_vox": N, "n_proj": P, "n_det": U, "parallelism": "single GPU"},
        "gups_forward": round(vv / (fwd_avg * 1e-3) / 1e9, 2), "gups_backward": round(vv / (bwd_avg * 1e-3) / 1e9, 2),
        "forward_ms": round(fwd_avg, 3), "backward_ms": round(bwd_avg, 3),
        "loss_first_last": [losses[0], losses[-1]],
        "roofline": roofline,
    }
    plan.close()
    del proj, vol
    torch.cuda.empty_cache()
    if not args.no_warmstart and os.path.exists(args.model or MODEL):
        pres = tuple(int(v) for v in args.pres.split(","))
        out["time_to_residual"] = ttr_frozen(ctx, xr, geom, sino, tau=args.tau, subsets=args.subsets or 1,
                                             long_epochs=args.long_epochs, pres=pres, model=args.model or MODEL)
    if not args.no_extras:
        out["fdk_art"] = fdk_art_measure(ctx, xr, geom, sino)
        out["grad"] = grad_measure(ctx, xr, geom, sino)
        del sino
        torch.cuda.empty_cache()
        out["xtt"] = xtt_measure(xr)
        b = batch_workload(xr, args)
        out["batch48"] = {k: b[k] for k in ("value", "unit", "batched_s", "sequential_s", "speedup_vs_sequential")}
        out["batch48"]["config"] = b["config"]["workload"]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(N, U, P, sdd, sod, det, vs, args.cpu_views)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
