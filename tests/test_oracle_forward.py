"""Pins the oracle's forward/back projectors (ir.cu:1480-1603, 1720-1850).

- the chord length L against an independent slab intersection,
- a constant volume against L * c * ratio * f/(f+1e-5) (the normalisation),
- adjointness of the raw splat and the raw gather (same (u, v), same weights),
- linearity, and the 90-degree rotation symmetry of the whole operator,
- the XTT operator against the plain one for a single basis vector.
"""
import numpy as np
import pytest

from conftest import rel_l2

SDD, SOD, DET = 600.0, 300.0, 0.8


def f32(a):
    """The oracle's inputs are FP32 values; compare against their exact FP64 images."""
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def ratio_of(vs):
    vs, det = f32(vs), f32(DET)
    return (vs * vs) / (det * (SOD / SDD) * det * (SOD / SDD))


def slab_chord(src, pix, lo, hi):
    d = pix - src
    t0, t1 = -np.inf, np.inf
    for i in range(3):
        if abs(d[i]) < 1e-300:
            if not (lo[i] < src[i] < hi[i]):
                return 0.0
            continue
        a, b = (lo[i] - src[i]) / d[i], (hi[i] - src[i]) / d[i]
        t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
    return max(0.0, t1 - t0) * np.linalg.norm(d)


def test_chord_length_matches_slab_method(orc):
    N, U, P, vs = 24, 48, 10, 0.5
    case = orc.Case(SDD, SOD, DET, N, U, P, vox_size=vs, cond_offset=[0.3, -0.2, 0.1], cond_t=[0.5, 0.25, -0.4])
    half = (N + 1) / 2 * f32(vs)
    det = f32(DET)
    for n in (0, 3, 7):
        L = case.chord_length(n)
        th = float(orc.theta_f32(n, P))
        R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
        off = R @ f32([0.3, -0.2, 0.1])
        t = f32([0.5, 0.25, -0.4])
        lo, hi = -half - off - t, half - off - t
        src = R @ np.array([0, SOD, 0])
        for (u, v) in [(0, 0), (24, 24), (10, 30), (47, 12), (23, 5)]:
            pix = R @ np.array([(u - (U - 1) / 2) * det, -(SDD - SOD), (v - (U - 1) / 2) * det])
            assert abs(L[v, u] - slab_chord(src, pix, lo, hi)) < 1e-9


@pytest.mark.parametrize("prec", [0, 1])
def test_constant_volume_normalisation(orc, prec):
    N, U, P, vs = 20, 40, 6, 0.6
    case = orc.Case(SDD, SOD, DET, N, U, P, vox_size=vs)
    c = f32(0.02)
    vol = np.full((N, N, N), c, dtype=np.float32)
    ratio = ratio_of(vs)
    for n in (0, 2, 5):
        p, f = case.forward_view(vol, n, prec=prec)
        L = case.chord_length(n)
        want = L * c * ratio * f / (f + 1e-5)
        tol = 1e-12 if prec == 1 else 2e-6
        assert rel_l2(p, want) < tol


def test_splat_gather_adjoint(orc):
    N, U, P, vs = 18, 30, 8, 0.7
    rng = np.random.default_rng(1)
    case = orc.Case(SDD, SOD, DET, N, U, P, vox_size=vs, cond_offset=[0.2, 0.1, -0.3])
    ratio = ratio_of(vs)
    x = rng.random((N, N, N)).astype(np.float32)
    y = rng.random((U, U)).astype(np.float32)
    for n in (0, 1, 5, -3):
        Sx, _ = case.forward_view(x, n, normalise=False)
        tmp = np.zeros(N ** 3)
        fac = np.zeros(N ** 3)
        case.backward_view(y, n, tmp, fac)
        lhs = float(np.sum(Sx * y.astype(np.float64)))
        rhs = ratio * float(np.sum(x.reshape(-1).astype(np.float64) * tmp))
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


def test_forward_linearity(orc):
    N, U, P, vs = 16, 28, 6, 0.7
    rng = np.random.default_rng(2)
    case = orc.Case(SDD, SOD, DET, N, U, P, vox_size=vs)
    a = rng.random((N, N, N)).astype(np.float32)
    b = rng.random((N, N, N)).astype(np.float32)
    pa, _ = case.forward_view(a, 3)
    pb, _ = case.forward_view(b, 3)
    pab, _ = case.forward_view((2.0 * a + 3.0 * b).astype(np.float32), 3)
    assert rel_l2(pab, 2.0 * pa + 3.0 * pb) < 1e-6  # the float32 rounding of 2a+3b


def test_quarter_turn_symmetry(orc, xr):
    # Rotating the object by +90 deg about z equals viewing it P/4 views earlier.
    # Checked on the raw splat: the normalisation's factor counts every grid voxel,
    # and the grid (centre (N+1)/2) is not symmetric under a quarter turn.
    N, U, P, vs = 33, 56, 16, 0.5
    case = orc.Case(SDD, SOD, DET, N, U, P, vox_size=vs)
    c = (N + 1) // 2
    params = xr.ellipsoid_params(11, 5, fov_radius=0.3 * N * vs, centre_frac=0.4, axis_lo=0.1, axis_hi=0.3,
                                 value_lo=0.01, value_hi=0.05)
    vol = orc.phantom_ellipsoids(N, vs, params)  # [z, y, x]
    assert not vol[:, :, :3].any() and not vol[:, :3, :].any()  # supported away from the asymmetric edge
    rot = np.zeros_like(vol)
    # (X, Y) -> (-Y, X): rot[z, y', x'] = vol[z, y, x] with x' - c = -(y - c), y' - c = x - c
    for y in range(N):
        for x in range(N):
            xp, yp = c - (y - c), c + (x - c)
            if 0 <= xp < N and 0 <= yp < N:
                rot[:, yp, xp] = vol[:, y, x]
    for n in (5, 9):
        a, _ = case.forward_view(rot, n, normalise=False)
        b, _ = case.forward_view(vol, n - P // 4, normalise=False)
        assert rel_l2(a, b) < 1e-5
        c_, _ = case.forward_view(vol, n + P // 4, normalise=False)  # the opposite sense must fail
        assert rel_l2(a, c_) > 1e-2


def test_xtt_single_basis_reduces_to_plain(orc):
    # K = 1: XTT splat = plain splat with each voxel weighted by |B x S|^2 (S.G)^2 vs.
    N, U, P, vs = 12, 24, 4, 0.7
    S = np.array([0.6, 0.0, 0.8])
    case = orc.Case(SDD, SOD, DET, N, U, P, vox_size=vs, basis=S)
    rng = np.random.default_rng(3)
    vol = rng.random((N, N, N)).astype(np.float32)
    n = 1
    w = np.zeros((N, N, N))
    for z in range(N):
        for y in range(N):
            for x in range(N):
                out = case.raycast(x, y, z, n)
                B, G = out[2:5], out[5:8]
                w[z, y, x] = (np.linalg.norm(np.cross(B, f32(S))) * abs(f32(S) @ G)) ** 2 * f32(vs)
    px, _ = case.forward_view(vol, n, xtt=True, normalise=False)
    pp, _ = case.forward_view((vol * w), n, xtt=False, normalise=False)
    # (float32 rounding of vol*w in the plain path)
    assert rel_l2(px, pp) < 1e-6


def test_xtt_weight_closed_forms(orc):
    N, U, P, vs = 9, 20, 4, 0.7
    case = orc.Case(SDD, SOD, DET, N, U, P, vox_size=vs, basis=[0, 0, 1])
    out = case.raycast(5, 5, 5, 0)  # origin voxel: B = -y (towards -y from source at +y), G = z
    B, G = out[2:5], out[5:8]
    assert np.allclose(B, [0, -1, 0]) and np.allclose(G, [0, 0, 1])
    # S = z: |B x S| = 1, |S.G| = 1 -> v = 1;  S = y (parallel to B): v = 0
    assert abs(np.linalg.norm(np.cross(B, [0, 0, 1])) * abs(np.dot([0, 0, 1], G)) - 1.0) < 1e-12
    assert np.linalg.norm(np.cross(B, [0, 1, 0])) < 1e-12
