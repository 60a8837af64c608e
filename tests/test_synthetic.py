"""Synthetic inputs (phantoms, noise): CPU pins and GPU parity.

Pins: the Philox4x32-10 block function against the published known-answer
vectors (Salmon et al. 2011, Random123 kat_vectors); Poisson / Gaussian noise
moments against their closed forms; vMF fibre directions against the mean
resultant length of the von Mises-Fisher distribution; GPU samples equal to the
oracle's CPU restatement of the same counter-based samplers."""
import numpy as np
import pytest

from conftest import rel_l2

KAT = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
       ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
       ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(xr, ctr, key, want):
    assert xr.philox4x32(ctr, key) == want


@pytest.mark.parametrize("lam", [3.0, 25.0, 1e4, 1e5])
def test_poisson_moments(orc, lam):
    n = 200000
    I0 = 1e5
    p = np.full(n, -np.log(lam / I0), np.float32)
    y = orc.noise_poisson(p, I0, seed=7)
    m = np.round(I0 * np.exp(-y.astype(np.float64)))  # = max(c, 1)
    lam_eff = I0 * np.exp(-np.float64(p[0]))
    p0 = np.exp(-lam_eff)  # max(c, 1) folds c = 0 into 1
    mean = lam_eff + p0
    var = lam_eff + lam_eff ** 2 + p0 - mean ** 2
    assert abs(m.mean() - mean) < 5 * np.sqrt(var / n)
    assert abs(m.var() / var - 1) < 5 * np.sqrt(2.0 / n) + 1e-3


def test_gaussian_moments(orc):
    n = 400000
    x = np.full(n, 2.0, np.float32)
    y = orc.noise_gaussian(x, 0.01, seed=9)
    g = (y.astype(np.float64) / 2.0 - 1.0) / 0.01
    assert abs(g.mean()) < 5 / np.sqrt(n) and abs(g.std() - 1) < 5 / np.sqrt(2 * n)


def test_fibre_directions_follow_vmf(xr):
    kappa = 20.0
    params, axes = xr.fibre_params(5, 3000, n_axes=3, kappa=kappa, fov_radius=10.0)
    f = params[:, 3:6]
    assert np.allclose(np.linalg.norm(f, axis=1), 1.0) and np.allclose(np.linalg.norm(axes, axis=1), 1.0)
    # mean resultant length of vMF on S^2: coth(kappa) - 1/kappa
    want = 1.0 / np.tanh(kappa) - 1.0 / kappa
    for a in range(3):
        cos = f[a::3] @ axes[a]
        assert abs(cos.mean() - want) < 5 * cos.std() / np.sqrt(len(cos))


def test_cylinder_shell_closed_form(orc):
    N, vs = 40, 0.5
    vol = orc.phantom_cylinder_shell(N, vs, 8.0, 1.5, 0.04)
    c = (N + 1) / 2
    z, y, x = np.meshgrid(*(np.arange(N) - c,) * 3, indexing="ij")
    r = np.sqrt((x * vs) ** 2 + (y * vs) ** 2)
    assert np.array_equal(vol > 0, (r <= 8.0) & (r > 6.5))


@pytest.mark.gpu
def test_noise_gpu_matches_oracle(orc, ctx, torch_dev):
    import torch

    rng = np.random.default_rng(1)
    p = (rng.random(300000) * 5).astype(np.float32)
    dp = torch.from_numpy(p).to(torch_dev)
    y = torch.empty_like(dp)
    ctx.noise_poisson(y, dp, 1e4, 11)
    got = y.cpu().numpy()
    want = orc.noise_poisson(p, 1e4, 11)
    # counter-based: identical samples except where FP64 libm and device log/exp/lgamma
    # round a rejection test differently
    assert np.mean(got == want) > 0.999
    x = (rng.random(300000) + 0.5).astype(np.float32)
    dx = torch.from_numpy(x).to(torch_dev)
    ctx.noise_gaussian(y, dx, 0.01, 12)
    assert rel_l2(y.cpu().numpy(), orc.noise_gaussian(x, 0.01, 12)) < 1e-6


@pytest.mark.gpu
def test_phantoms_gpu_match_oracle(orc, xr, ctx, torch_dev):
    import torch

    N, vs = 48, 0.5
    basis = np.eye(3)
    params, _ = xr.fibre_params(3, 20, fov_radius=0.5 * N * vs, r_lo=1.0, r_hi=3.0)
    out = torch.zeros(3 * N ** 3, device=torch_dev)
    ctx.phantom_fibres(params, N, vs, basis, out)
    want = orc.phantom_fibres(N, vs, params, basis)
    assert np.count_nonzero(np.abs(out.cpu().numpy().reshape(want.shape) - want) > 1e-6) <= 3
    shell = torch.zeros(N ** 3, device=torch_dev)
    ctx.phantom_cylinder_shell(N, vs, 10.0, 1.5, 0.04, shell)
    assert np.count_nonzero(shell.cpu().numpy().reshape(N, N, N) != orc.phantom_cylinder_shell(N, vs, 10.0, 1.5, 0.04)) <= 2


def test_training_phantom_draw(xr):
    """The warm-start training phantoms: the draw depends only on the seed,
    "fibres" is always the ellipsoid-and-fibre default, and "mixed" keeps its
    stated proportions (0.4 default; 0.75 of the rest with a cylinder shell)
    and parameter ranges."""
    import importlib

    ws = importlib.import_module(xr.__name__ + ".warmstart")
    N, vs = 128, 0.2
    specs = [ws.phantom_spec(s, N, vs) for s in range(2000)]
    assert specs == [ws.phantom_spec(s, N, vs) for s in range(2000)]
    assert all(ws.phantom_spec(s, N, vs, "fibres") == {} for s in range(50))
    dense = [d for d in specs if d]
    n = len(specs)
    assert abs(len(dense) / n - 0.6) < 5 * np.sqrt(0.24 / n)
    shells = [d["shell"] for d in dense if "shell" in d]
    assert abs(len(shells) / len(dense) - 0.75) < 5 * np.sqrt(0.1875 / len(dense))
    for d in dense:
        assert 20 <= d["n_ellipsoids"] <= 48
        assert d["n_fibres"] == 0 or 8 <= d["n_fibres"] <= 24
    r, t, v = np.array(shells).T
    assert r.min() >= 0.35 * N * vs and r.max() <= 0.48 * N * vs
    assert t.min() >= vs and t.max() <= 5 * vs
    assert v.min() >= 0.01 and v.max() <= 0.04
