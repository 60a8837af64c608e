"""Learned warm start for plain MLEM (no counterpart in the reference).

A small 3-D U-Net (``UNet``, BF16 MFMA kernels in libxr.so) maps an early
MLEM volume (the plain method after a few iterations) to an estimate of a
late one.  It is trained only on pairs the plain method itself generates
(``generate_pairs``), and at run time it never replaces the method: MLEM
resumes from the network's output, and the projection-domain data residual
checks it — if the network's volume fits the data clearly worse (beyond a
relative tolerance) than the volume it was given, the driver falls back to
that volume (``warm_start``).

Data residual: r(x) = sum|y - A x| / sum|y| (relative L1 of the normalised
forward projection A, accumulated in FP64 by the forward kernel), the
reference's own MLEM loss (reconstruct.cu:131) divided by the data's L1 norm.

Everything here orchestrates device work through the C ABI; volumes stay in
HBM in the column layout.
"""
import time

import numpy as np

from . import BATCH_GRAPH, METHOD_MLEM, OUT_RATIO, Scan, UNet, XRError, make_geometry  # noqa: F401

__all__ = ["Scene", "make_phantom", "phantom_spec", "generate_pairs", "train", "warm_start", "plain_mlem", "positivity_floor",
           "apply_network", "trained_edge", "plain_curve", "warm_curve", "first_within"]


class Scene:
    """One scan geometry (single untilted condition) with a device sinogram.

    Wraps a one-problem batch (``Context.batch``) so MLEM runs on device
    buffers with the epoch loop captured as a HIP graph.  `subsets` > 1 gives
    ordered-subset MLEM (IR::reconstruct's batch, reconstruct.cu:61-65)."""

    def __init__(self, ctx, geom, sino_dev, subsets=1):
        import torch

        self.ctx, self.geom = ctx, geom
        self.batch = ctx.batch([geom], batch=[subsets], flags=BATCH_GRAPH)
        self.vol_addr, self.sino_addr = self.batch.buffers(0)
        self.plan = self.batch.plan(0)
        self.N, self.P, self.U, self.V = geom.n_vox, geom.n_proj, geom.n_det_u, geom.n_det_v
        ctx.copy(self.sino_addr, sino_dev)
        self.y_l1 = float(sino_dev.abs().double().sum())
        dev = sino_dev.device
        self.proj = torch.empty(self.P * self.U * self.V, dtype=torch.float32, device=dev)
        self.codes = self.plan.codes(range(self.P))
        self.sino = sino_dev
        self._subset = {}

    def close(self):
        self.batch.close()

    def set_volume(self, vol_cols):
        self.ctx.copy(self.vol_addr, vol_cols)

    def restart(self, vol_cols):
        """Start a new solve from vol_cols (subset order from the beginning)."""
        self.batch.restart()
        self.set_volume(vol_cols)

    def get_volume(self, out):
        self.ctx.copy(out, self.vol_addr, out.numel() * 4)

    def residual(self, vol_cols, every=1):
        """r(x) for a column-layout device volume (one forward projection).
        every > 1: the same ratio over views 0, every, 2*every, ... only (an
        estimate for comparing volumes at a fraction of the cost)."""
        if every <= 1:
            loss = self.plan.project(vol_cols, self.proj, self.codes, out_mode=OUT_RATIO, sino=self.sino)
            return loss / self.y_l1
        if every not in self._subset:
            import torch

            views = list(range(0, self.P, every))
            idx = torch.tensor(views, device=self.sino.device)
            y_l1 = float(self.sino.view(self.P, -1)[idx].abs().double().sum())
            self._subset[every] = (self.plan.codes(views), y_l1)
        codes, y_l1 = self._subset[every]
        # images are addressed by view in the full sinogram and ratio buffers
        loss = self.plan.project(vol_cols, self.proj, codes, out_mode=OUT_RATIO, sino=self.sino)
        return loss / y_l1

    def residual_l2(self, vol=None):
        """R(x) = ||y - A x||^2 / ||y||^2 over every view (FP64 sums), for a
        column-layout device volume or, by default, the scene's current one."""
        if not hasattr(self, "y_l2"):
            self.y_l2 = float((self.sino.double() ** 2).sum())
        _, l2 = self.plan.residual(self.vol_addr if vol is None else vol, self.proj, self.sino, self.codes)
        return l2 / self.y_l2

    def iterate(self, epochs):
        """MLEM epochs on the scene's volume; returns (residuals before each
        epoch, GPU milliseconds)."""
        losses = self.batch.run(epochs)[:, 0]
        return losses / self.y_l1, self.batch.elapsed_ms()


def make_phantom(ctx, N, vox_size, seed, n_ellipsoids=12, n_fibres=24, fibre_value=0.02, vol_cols=None,
                 ellipsoid_kw=None, shell=None):
    """Seeded ellipsoids plus fibre bundles (3 von Mises-Fisher axes, kappa
    20, radius 2-6 voxels) and optionally a z-axis cylinder shell, as a
    column-layout device volume.

    `ellipsoid_kw` overrides the ellipsoid draw (default: attenuation uniform
    in [0.005, 0.05]/mm, centres within half the field of view, semi-axes
    0.08-0.3 of its radius); `shell` = (outer radius, wall thickness, value)
    in mm, mm and 1/mm."""
    import torch

    from . import ellipsoid_params, fibre_params

    dev = torch.device("cuda")
    fov = 0.5 * N * vox_size
    ref = torch.zeros(N ** 3, dtype=torch.float32, device=dev)
    kw = dict(centre_frac=0.5, axis_lo=0.08, axis_hi=0.3, value_lo=0.005, value_hi=0.05)
    kw.update(ellipsoid_kw or {})
    ep = ellipsoid_params(seed, n_ellipsoids, fov_radius=fov, **kw)
    ctx.phantom_rasterise(ep, N, vox_size, ref)
    if shell is not None:
        ctx.phantom_cylinder_shell(N, vox_size, shell[0], shell[1], shell[2], ref)
    if n_fibres:
        fp, _ = fibre_params(seed + 7919, n_fibres, n_axes=3, kappa=20.0, fov_radius=fov, centre_frac=0.5,
                             r_lo=2.0 * vox_size, r_hi=6.0 * vox_size, half_length=0.6 * fov)
        ctx.phantom_fibres_scalar(fp, N, vox_size, fibre_value, ref)
    out = torch.empty_like(ref) if vol_cols is None else vol_cols
    ctx.to_columns(ref, out, N)
    return out


def phantom_spec(seed, N, vox_size, family="mixed"):
    """Keyword arguments of `make_phantom` for training phantom `seed`.

    family "fibres": 12 ellipsoids and 24 fibre bundles.  family "mixed":
    that with probability 0.4; otherwise a dense set of 20-48 overlapping
    ellipsoids (attenuation 0-0.05/mm each, centres within 0.6 of the field
    of view, semi-axes 0.1-0.35 of its radius), a cylinder shell with
    probability 0.75 (outer radius 0.35-0.48 N voxels, wall 1-5 voxels,
    0.01-0.04/mm) and 8-24 fibre bundles with probability 0.3.  The draw
    depends only on `seed`."""
    if family == "fibres":
        return {}
    rng = np.random.default_rng([seed, 0x5EED])
    if rng.random() < 0.4:
        return {}
    spec = {"n_ellipsoids": int(rng.integers(20, 49)), "n_fibres": 0,
            "ellipsoid_kw": dict(centre_frac=0.6, axis_lo=0.1, axis_hi=0.35, value_lo=0.0, value_hi=0.05)}
    if rng.random() < 0.75:
        spec["shell"] = (float(rng.uniform(0.35, 0.48)) * N * vox_size, float(rng.uniform(1.0, 5.0)) * vox_size,
                         float(rng.uniform(0.01, 0.04)))
    if rng.random() < 0.3:
        spec["n_fibres"] = int(rng.integers(8, 25))
    return spec


def project(ctx, geom, vol_cols):
    """Normalised forward projection of every view (the sinogram model)."""
    import torch

    plan = ctx.plan(geom, Scan())
    proj = torch.zeros(geom.n_proj * geom.n_det_u * geom.n_det_v, dtype=torch.float32, device=vol_cols.device)
    plan.project(vol_cols, proj, plan.codes(range(geom.n_proj)))
    ctx.synchronize()
    plan.close()
    return proj


def generate_pairs(ctx, count, N=128, P=180, U=192, sdd=1000.0, sod=500.0, det=0.4, pre=5, post=200, seed=0,
                   per_batch=48, init=0.01, subsets=1, family="fibres", photons=None, log=None):
    """(inputs, targets, geometry): the plain method's volume (MLEM with
    `subsets` ordered subsets, IR::reconstruct's batch) after `pre` and after
    `post` epochs for `count` seeded phantoms (`phantom_spec(seed + i,
    family)`), all from the batched solver (`per_batch` problems per
    persistent launch).  `pre` may be a list of epoch counts: each phantom
    then gives one input per count, all paired with the same target (the same
    tensor object).  `photons`: Poisson noise at that many photons per ray
    (seeded per phantom) and the driver's clamp of the noisy sinogram
    (main.cu:35); None leaves the sinogram noise free.  Lists of column-layout
    device tensors."""
    import torch

    pres = sorted(set([pre] if np.isscalar(pre) else [int(p) for p in pre]))
    if pres[0] < 0 or pres[-1] >= post:
        raise ValueError("pre epochs must lie in [0, post)")
    vs = sod * det / sdd  # one voxel per magnified detector pixel (geometry.h:13)
    geom = make_geometry(sdd, sod, det, N, U, P, vox_size=vs)
    inputs, targets = [], []

    def snapshot(batch, n):
        out = []
        for i in range(n):
            x = torch.empty(N ** 3, dtype=torch.float32, device="cuda")
            ctx.copy(x, batch.buffers(i)[0], x.numel() * 4)
            out.append(x)
        return out

    for b0 in range(0, count, per_batch):
        n = min(per_batch, count - b0)
        batch = ctx.batch([geom] * n, batch=[subsets] * n, flags=BATCH_GRAPH)
        for i in range(n):
            vol = make_phantom(ctx, N, vs, seed + b0 + i, **phantom_spec(seed + b0 + i, N, vs, family))
            sino = project(ctx, geom, vol)
            if photons:
                clean = sino
                sino = torch.empty_like(clean)
                ctx.noise_poisson(sino, clean, photons, seed + b0 + i)
                ctx.sino_clamp(sino)
                del clean
            va, sa = batch.buffers(i)
            ctx.copy(sa, sino)
            ctx.copy(va, torch.full((N ** 3,), init, dtype=torch.float32, device=sino.device))
        t0 = time.time()
        done, states = 0, []
        for p in pres:
            if p > done:
                batch.run(p - done)
                done = p
            states.append(snapshot(batch, n))
        batch.run(post - done)
        final = snapshot(batch, n)
        batch.close()
        for st in states:
            inputs.extend(st)
            targets.extend(final)
        if log:
            log("pairs %d-%d: %.1f s" % (b0, b0 + n, time.time() - t0))
    return inputs, targets, geom


def train(net, inputs, targets, dims, steps, lr=3e-4, seed=0, log=None, log_every=50, lr_final=None):
    """Adam on single-volume steps in a seeded random order, learning rate
    cosine-decayed from `lr` to `lr_final` (default lr/10); returns the loss
    history."""
    rng = np.random.default_rng(seed)
    lr_final = lr / 10 if lr_final is None else lr_final
    hist = []
    order = []
    for k in range(steps):
        if not order:
            order = list(rng.permutation(len(inputs)))
        i = order.pop()
        hist.append(net.loss_grad(inputs[i], targets[i], dims))
        t = k / max(steps - 1, 1)
        net.adam_step(lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * t)))
        if log and (k % log_every == 0 or k == steps - 1):
            log("step %d loss %.5f" % (k, hist[-1]))
    return hist


def positivity_floor(vol_cols, floor):
    """MLEM is multiplicative: a voxel at zero never moves again.  The
    network's output is clamped to at least `floor`."""
    vol_cols.clamp_(min=floor)
    return vol_cols


def trained_edge(model_path):
    """The volume edge a saved network was trained at (its training summary,
    model_path + ".json", written by train_warmstart.py), or None."""
    import json
    import os

    meta = model_path + ".json"
    if not os.path.exists(meta):
        return None
    with open(meta) as f:
        d = json.load(f)
    if "n_vox" in d:
        return int(d["n_vox"])
    stages = d.get("stages") or []
    return int(stages[-1]["n_vox"]) if stages else None


def apply_network(ctx, net, x, N, floor_frac=1e-3, train_n=None):
    """The network's estimate from the column-layout volume x (edge N), floored
    at floor_frac x the mean of its input.

    MLEM recovers an object's spatial frequencies in an order set by the
    field of view, not by the voxel size (the backprojection of a projection
    is a 1/r blur over the whole field), so the early iterate of an N^3 solve
    is, in field-of-view units, the early iterate of a coarser solve.  A
    network trained at edge train_n < N (N a multiple of it) is therefore
    applied at that scale: x is averaged over f^3 blocks (f = N / train_n),
    the network's correction y - x is computed there and added to x after
    trilinear interpolation back to N^3.  train_n None or N: applied as is."""
    import torch

    y = torch.empty_like(x)
    if train_n and N != train_n:
        if N % train_n:
            raise ValueError("volume edge %d is not a multiple of the network's %d" % (N, train_n))
        f = N // train_n
        xd = torch.empty(train_n ** 3, dtype=torch.float32, device=x.device)
        ctx.volume_downsample(x, N, f, xd)
        yd = torch.empty_like(xd)
        net.forward(xd, yd, (train_n, train_n, train_n))
        y.copy_(x)
        ctx.volume_upsample_add(yd, xd, train_n, f, y, floor_frac * float(net.last_scale()))
        return y
    net.forward(x, y, (N, N, N))
    positivity_floor(y, floor_frac * float(net.last_scale()))
    return y


def plain_mlem(scene, tau, max_iters, init=0.01):
    """Plain MLEM from a constant volume (main.cu:40-42) until r <= tau."""
    import torch

    scene.restart(torch.full((scene.N ** 3,), init, dtype=torch.float32, device=scene.sino.device))
    return _resume(scene, tau, max_iters, 0, 0.0, [])


def _resume(scene, tau, max_iters, iters, ms, hist):
    while iters < max_iters:
        r, t = scene.iterate(1)
        ms += t
        hist.append(float(r[0]))
        if r[0] <= tau:
            return {"iterations": iters, "ms": ms, "residuals": hist, "reached": True}
        iters += 1
    return {"iterations": iters, "ms": ms, "residuals": hist, "reached": False}


def warm_start(scene, net, tau, max_iters, pre=5, init=0.01, floor_frac=1e-3, use_net=True, check_every=4,
               accept_tol=0.1, train_n=None):
    """Plain MLEM for `pre` iterations, then the network (checked by the data
    residual, with fallback), then MLEM until r <= tau.  Returns iterations
    (MLEM iterations only), GPU milliseconds (everything: MLEM, network,
    residual checks), the residual history and whether the network's volume
    was kept.

    The check compares the residuals of the two volumes over the same views,
    every `check_every`-th one (both are sums over millions of detector
    pixels, so a quarter of the views ranks them as the full set does at a
    quarter of the cost); check_every=1 uses every view.  The network's volume
    is kept unless its residual exceeds the MLEM volume's by more than
    `accept_tol` (relative): a useful network changes the data residual little
    (it corrects the components MLEM converges on slowly, which the residual
    barely sees), while a harmful one raises it by far more.  The stopping
    test r <= tau always uses every view."""
    import torch

    dev = scene.sino.device
    N3 = scene.N ** 3
    scene.restart(torch.full((N3,), init, dtype=torch.float32, device=dev))
    r, ms = scene.iterate(pre)
    hist = [float(v) for v in r]
    used = False
    if use_net:
        x = torch.empty(N3, dtype=torch.float32, device=dev)
        scene.get_volume(x)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        scene.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        e0.record()
        y = apply_network(scene.ctx, net, x, scene.N, floor_frac, train_n)
        r_x = scene.residual(x, check_every)
        r_y = scene.residual(y, check_every)
        e1.record()
        torch.cuda.synchronize()
        scene.ctx.set_stream(0)
        ms += e0.elapsed_time(e1)
        used = r_y <= (1.0 + accept_tol) * r_x
        scene.set_volume(y if used else x)
    out = _resume(scene, tau, max_iters, pre, ms, hist)
    out["used_net"] = used
    return out


# ---- time to residual with the frozen relative criterion -------------------
#
# R(x) = ||y - A x||^2 / ||y||^2 (FP64), evaluated on the iterate after every
# epoch.  With noisy data R plateaus at the noise level, so the criterion is
# relative to the reduction a long plain run achieves:
#     c_k = (R_k - R_inf) / (R_0 - R_inf) <= tau,
# R_0 = R of the plain start (x0 = 0.01 everywhere, main.cu:40-42) and R_inf =
# R after `long_epochs` plain epochs.  Both methods are timed on the GPU
# (epochs, network, acceptance check); the per-epoch evaluations of R that
# decide when the criterion is met are instrumentation and are not timed
# (they cost one full forward projection each, the same for both methods).


def plain_curve(scene, epochs, init=0.01):
    """Plain MLEM from x0 = init for `epochs` epochs: R_0 .. R_epochs and the
    cumulative GPU milliseconds after each epoch (0 for R_0)."""
    import torch

    scene.restart(torch.full((scene.N ** 3,), init, dtype=torch.float32, device=scene.sino.device))
    R, ms = [scene.residual_l2()], [0.0]
    for _ in range(epochs):
        _, t = scene.iterate(1)
        ms.append(ms[-1] + t)
        R.append(scene.residual_l2())
    return R, ms


def first_within(R, R0, Rinf, tau):
    """First index k with (R_k - R_inf) / (R_0 - R_inf) <= tau, or None."""
    den = R0 - Rinf
    for k, r in enumerate(R):
        if (r - Rinf) <= tau * den:
            return k
    return None


def warm_curve(scene, net, R0, Rinf, tau, max_epochs, pre=5, init=0.01, floor_frac=1e-3, check_every=4,
               accept_tol=0.1, train_n=None):
    """Warm start against the frozen criterion: `pre` plain epochs from
    x0 = init, the network (its output clamped to a positivity floor), the
    acceptance check (the network's volume is kept unless its data residual
    over every `check_every`-th view exceeds the MLEM volume's by more than
    `accept_tol`), then plain epochs until the criterion holds.  Returns the
    epochs (plain epochs only), the GPU milliseconds (epochs + network +
    check), R after each epoch, and whether the network's volume was kept."""
    import torch

    dev = scene.sino.device
    N3 = scene.N ** 3
    scene.restart(torch.full((N3,), init, dtype=torch.float32, device=dev))
    _, ms = scene.iterate(pre) if pre > 0 else (None, 0.0)
    x = torch.empty(N3, dtype=torch.float32, device=dev)
    scene.get_volume(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    scene.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    e0.record()
    y = apply_network(scene.ctx, net, x, scene.N, floor_frac, train_n)
    r_x = scene.residual(x, check_every)
    r_y = scene.residual(y, check_every)
    e1.record()
    torch.cuda.synchronize()
    scene.ctx.set_stream(0)
    ms += e0.elapsed_time(e1)
    used = r_y <= (1.0 + accept_tol) * r_x
    scene.set_volume(y if used else x)
    del x, y
    R = [scene.residual_l2()]
    epochs = pre
    while (R[-1] - Rinf) > tau * (R0 - Rinf) and epochs < max_epochs:
        _, t = scene.iterate(1)
        ms += t
        epochs += 1
        R.append(scene.residual_l2())
    reached = (R[-1] - Rinf) <= tau * (R0 - Rinf)
    return {"epochs": epochs, "ms": ms, "R": R, "used_net": bool(used), "reached": bool(reached)}
