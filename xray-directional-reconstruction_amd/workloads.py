"""Synthetic workloads of the measured configurations (no reference counterpart).

Each constructor returns the geometry, the scan tables and device data built
by the library's own phantom, projector and noise kernels:

* ``xtt_scan`` / ``xtt_workload`` — directional (XTT) reconstruction input:
  a 256^3 volume of 13 scattering-direction coefficients per voxel, fibre
  bundles whose axes follow von Mises-Fisher distributions (kappa 20) around
  3 random axes with radius 2-6 voxels, 4 grating orientations (0/45/90/135
  degrees in the detector plane) x 360 views on a 512^2 detector, dark-field
  data with 1 % Gaussian noise on the visibility.
* ``batch_problems`` — 48 independent MLEM problems with volumes drawn from
  {64, 96, 128, 192, 256}^3, {90, 180, 360} views, detector 1.5x the volume
  edge.

Direction set: "13 icosahedral half-sphere vectors" has no icosahedral
realisation (those give 6, 10 or 15 directions); the 13 used here are the
cubic set: the 3 axes, the 4 body diagonals and the 6 face diagonals, each
taken in the half space z >= 0 (or y > 0 where z = 0), normalised.  The
grating vector of condition c lies in the detector plane spanned by R x and
R z (ir.cu:1874-1876), at angle phi_c from R x (the reference fixes it at R z,
ir.cu:1901; phi = 90 degrees reproduces that).
"""
import numpy as np

XTT_BASIS13 = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [1, -1, 1], [-1, 1, 1], [-1, -1, 1],
                        [1, 1, 0], [1, -1, 0], [1, 0, 1], [-1, 0, 1], [0, 1, 1], [0, -1, 1]], dtype=np.float64)
XTT_BASIS13 /= np.linalg.norm(XTT_BASIS13, axis=1, keepdims=True)


def grating(phis_deg):
    """Per-condition grating vectors (cos phi, 0, sin phi) in the detector frame."""
    return np.concatenate([[np.cos(np.deg2rad(p)), 0.0, np.sin(np.deg2rad(p))] for p in phis_deg])


def xtt_scan(xr, phis_deg=(0.0, 45.0, 90.0, 135.0), basis=XTT_BASIS13):
    """Untilted conditions (identity rotation), one grating angle each."""
    C = len(phis_deg)
    return xr.Scan(cond_R=np.tile(np.eye(3).reshape(-1), C), basis=basis, grating=grating(phis_deg))


def xtt_workload(xr, ctx, N=256, U=512, P=360, phis_deg=(0.0, 45.0, 90.0, 135.0), sdd=1000.0, sod=500.0, det=0.8,
                 seed=3000, n_fibres=120, sigma=0.01, noise=True):
    """Geometry, scan, ground truth (K channels, column layout) and the noisy
    dark-field sinogram (n_cond x P x V x U) on the device.

    Data: the XTT forward projection p (normalised), visibility exp(-p) times
    (1 + sigma N(0, 1)) (``xr_noise_gaussian``), then -ln.  Voxel size is the
    reference's sod * det / sdd (geometry.h:13)."""
    import torch

    geom = xr.make_geometry(sdd, sod, det, N, U, P)
    vs = geom.vox_size
    scan = xtt_scan(xr, phis_deg)
    K = scan.n_basis
    dev = torch.device("cuda")
    fov = 0.5 * N * vs
    fp, axes = xr.fibre_params(seed, n_fibres, n_axes=3, kappa=20.0, fov_radius=fov, centre_frac=0.55,
                               r_lo=2.0 * vs, r_hi=6.0 * vs, half_length=0.7 * fov)
    ref = torch.zeros(K * N ** 3, dtype=torch.float32, device=dev)
    ctx.phantom_fibres(fp, N, vs, scan.basis.reshape(-1, 3), ref)
    ref.mul_(0.02)
    truth = torch.empty_like(ref)
    ctx.to_columns(ref, truth, N, K)
    del ref
    plan = ctx.plan(geom, scan)
    C = scan.n_cond
    sino = torch.zeros(C * P * U * U, dtype=torch.float32, device=dev)
    codes = np.arange(C * P, dtype=np.int32)
    plan.project(truth, sino, codes, out_mode=xr.OUT_NORMALISED, xtt=True)
    if noise:
        vis = torch.exp(-sino)
        ctx.noise_gaussian(vis, vis, sigma, seed + 1)
        sino = -torch.log(vis.clamp_min(1e-30))
        del vis
    ctx.synchronize()
    return geom, scan, plan, truth, sino, axes


C4_SIZES = (64, 96, 128, 192, 256)
C4_VIEWS = (90, 180, 360)


def batch_problems(xr, n=48, seed=4, sdd=1000.0, sod=500.0, det=0.4):
    """(N, P) draws and geometries of the batched workload: U = 1.5 N, the
    voxel size of the largest volume for all (so every volume fills the same
    fraction of its detector)."""
    rng = np.random.default_rng(seed)
    Ns = [int(v) for v in rng.choice(C4_SIZES, n)]
    Ps = [int(v) for v in rng.choice(C4_VIEWS, n)]
    vs = xr.default_vox_size(sdd, sod, det)
    geoms = [xr.make_geometry(sdd, sod, det, N, 3 * N // 2, P, vox_size=vs) for N, P in zip(Ns, Ps)]
    return Ns, Ps, geoms
