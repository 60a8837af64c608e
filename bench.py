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


def cpu_baseline(N, U, P, sdd, sod, det, vs, n_views=2):
    """Time the oracle's FP32 restatement (the reference's operators on the CPU)
    on a bounded sample: forward + back of n_views views of the same workload."""
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
    return {"value": gups, "unit": "GUPS", "cores": int(orc.num_threads()), "kind": "port",
            "sample": "FP32 oracle forward+back of %d of %d views at %d^3, %dx%d detector (%.1f s)"
                      % (len(views), P, N, U, U, dt)}


MODEL = os.path.join(ROOT, "xray-directional-reconstruction_amd", "models", "warmstart_c24.xunt")


def time_to_residual(ctx, xr, geom, sino, tau, subsets, max_epochs=300, pre=5):
    """Epochs and GPU seconds until the data residual r = sum|y - Ax|/sum|y|
    reaches tau: the plain method (MLEM, `subsets` ordered subsets, x0 = 0.01)
    against the learned warm start (plain for `pre` epochs, then the trained
    network judged by the residual with fallback, then plain again)."""
    import importlib

    ws = importlib.import_module(xr.__name__ + ".warmstart")
    net = xr.UNet(ctx, 24)
    net.load(MODEL)
    scene = ws.Scene(ctx, geom, sino, subsets=subsets)
    plain = ws.plain_mlem(scene, tau, max_epochs)
    warm = ws.warm_start(scene, net, tau, max_epochs, pre=pre)
    scene.close()
    net.close()
    return {"tau": tau, "residual": "sum|y-Ax|/sum|y| (FP64)", "method": "MLEM, %d ordered subsets" % subsets,
            "plain_epochs": plain["iterations"], "plain_s": round(plain["ms"] / 1e3, 3),
            "plain_reached": plain["reached"],
            "warm_epochs": warm["iterations"], "warm_s": round(warm["ms"] / 1e3, 3), "warm_reached": warm["reached"],
            "network_kept": warm["used_net"],
            "speedup_epochs": round(plain["iterations"] / max(warm["iterations"], 1), 3),
            "speedup_time": round(plain["ms"] / max(warm["ms"], 1e-9), 3),
            "network": "models/warmstart_c24.xunt (128^3, 180 views, 480 pairs)"}


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
    ap.add_argument("--cpu-views", type=int, default=2)
    ap.add_argument("--no-warmstart", action="store_true",
                    help="skip the time-to-residual comparison (learned warm start vs plain MLEM)")
    ap.add_argument("--tau", type=float, default=1e-3)
    ap.add_argument("--subsets", type=int, default=20)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 1 GiB device-to-device copy (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (views would shard "
              "across GPUs with an all-reduce of the backprojection; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import torch

    import __graft_entry__ as entry

    xr = entry.load_package()
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

    # synthetic phantom (40 seeded ellipsoids plus a cylinder shell) and its
    # sinogram by the same operator (noise free, so the residual can reach tau)
    params = xr.ellipsoid_params(1234, 40, fov_radius=0.5 * N * vs)
    vref = torch.zeros(N3, device=dev)
    ctx.phantom_rasterise(params, N, vs, vref)
    ctx.phantom_cylinder_shell(N, vs, 0.45 * N * vs, 4 * vs, 0.02, vref)
    truth = torch.empty_like(vref)
    ctx.to_columns(vref, truth, N)
    del vref
    codes = plan.codes(range(P))
    sino = torch.zeros(lenD, device=dev)
    plan.project(truth, sino, codes, out_mode=xr.OUT_NORMALISED)
    del truth
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
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (40 seeded ellipsoids, sinogram from the projector)",
        "config": {"workload": "plain MLEM iteration (batch 1), %d^3 volume, %d views, %dx%d detector, "
                               "SOD 500 SDD 1000 det 0.4 mm" % (N, P, U, U),
                   "n_vox": N, "n_proj": P, "n_det": U, "parallelism": "single GPU"},
        "gups_forward": round(vv / (fwd_avg * 1e-3) / 1e9, 2), "gups_backward": round(vv / (bwd_avg * 1e-3) / 1e9, 2),
        "forward_ms": round(fwd_avg, 3), "backward_ms": round(bwd_avg, 3),
        "loss_first_last": [losses[0], losses[-1]],
        "roofline": roofline,
    }
    if not args.no_warmstart and os.path.exists(MODEL):
        out["time_to_residual"] = time_to_residual(ctx, xr, geom, sino, args.tau, args.subsets)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(N, U, P, sdd, sod, det, vs, args.cpu_views)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
