"""Pins the oracle's orientation-aware XTT operator (forwardOrth / backwardOrth,
ir.cu:295-423) and rodriguesRotation (ir.cu:597-621).

- the rotation against scipy's rotation-vector form, orthonormality and the
  fixed axis;
- cos theta = 1 (no rotation) reduces forwardOrth to the plain XTT operator;
- a uniform rotation R equals the plain XTT operator with the rotated basis
  R S_k (so S' = R S, not R^T S);
- adjointness of the raw forwardOrth splat and backwardOrth gather under a
  random per-voxel orientation field.
"""
import numpy as np
import pytest

from conftest import rel_l2

SDD, SOD, DET = 600.0, 300.0, 0.8
BASIS3 = np.array([[0.666667, -0.666667, -0.333333], [-0.333333, -0.666667, 0.666667],
                   [0.666667, 0.333333, 0.666667]], np.float32)


def test_rodrigues_matches_rotation_vector(orc):
    from scipy.spatial.transform import Rotation

    rng = np.random.default_rng(3)
    for _ in range(20):
        phi = rng.uniform(-np.pi, np.pi)
        theta = rng.uniform(0, np.pi)
        axis = np.array([np.cos(phi), np.sin(phi), 0.0])
        R = orc.rodrigues(axis[0], axis[1], axis[2], np.cos(theta), np.sin(theta))
        want = Rotation.from_rotvec(axis * theta).as_matrix()
        assert np.allclose(R, want, atol=1e-12)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12) and np.allclose(R @ axis, axis, atol=1e-12)
    assert np.array_equal(orc.rodrigues(0, 0, 0, 0.3, 0.9), np.eye(3))  # zero axis: identity


def orth_case(N=14, U=26, P=8, vs=0.7, basis=BASIS3):
    import numpy as np

    grating = np.array([0.0, 0.0, 1.0, 1.0, 0.0, 0.0], np.float32)
    return lambda orc: orc.Case(SDD, SOD, DET, N, U, P, vox_size=vs, cond_R=np.tile(np.eye(3).reshape(-1), 2),
                                basis=basis, grating=grating)


def test_unrotated_orth_is_the_xtt_operator(orc):
    N = 14
    case = orth_case(N)(orc)
    rng = np.random.default_rng(4)
    vol = rng.random((3, N, N, N)).astype(np.float32)
    coef = np.stack([rng.uniform(-3, 3, (N, N, N)), np.ones((N, N, N))]).astype(np.float32)
    for cond, n in ((0, 0), (1, 3), (0, -5)):
        want, _ = case.forward_view(vol, n, cond=cond, xtt=True)
        got, _ = case.forward_view_orth(vol, coef, n, cond=cond)
        assert rel_l2(got, want) <= 1e-14


def test_uniform_rotation_rotates_the_basis(orc):
    N = 14
    phi, ct = 0.7, 0.35
    R = orc.rodrigues(np.cos(phi), np.sin(phi), 0.0, ct, np.sqrt(1 - ct * ct))
    rotated = (BASIS3.astype(np.float64) @ R.T).astype(np.float32)  # rows R S_k
    case = orth_case(N)(orc)
    ref = orth_case(N, basis=rotated)(orc)
    rng = np.random.default_rng(5)
    vol = rng.random((3, N, N, N)).astype(np.float32)
    coef = np.stack([np.full((N, N, N), phi), np.full((N, N, N), ct)]).astype(np.float32)
    for cond, n in ((0, 1), (1, 4)):
        got, _ = case.forward_view_orth(vol, coef, n, cond=cond)
        want, _ = ref.forward_view(vol, n, cond=cond, xtt=True)
        assert rel_l2(got, want) <= 1e-6  # the rotated basis rounded to FP32
        wrong = orth_case(N, basis=(BASIS3.astype(np.float64) @ R).astype(np.float32))(orc)
        bad, _ = wrong.forward_view(vol, n, cond=cond, xtt=True)
        assert rel_l2(got, bad) > 1e-2  # R^T S would be a different operator


def test_orth_splat_gather_adjoint(orc):
    N, U, P, vs = 12, 24, 6, 0.7
    case = orth_case(N, U, P, vs)(orc)
    rng = np.random.default_rng(6)
    x = rng.random((3, N, N, N)).astype(np.float32)
    coef = np.stack([rng.uniform(-np.pi, np.pi, (N, N, N)), rng.uniform(-1, 1, (N, N, N))]).astype(np.float32)
    y = rng.random((U, U)).astype(np.float32)
    vsf = float(np.float32(vs))
    ratio = (vsf * vsf) / (float(np.float32(DET)) * (SOD / SDD) * float(np.float32(DET)) * (SOD / SDD))
    for cond, n in ((0, 0), (1, 2), (0, -3)):
        Sx, _ = case.forward_view_orth(x, coef, n, cond=cond, normalise=False)
        tmp = np.zeros(3 * N ** 3)
        fac = np.zeros(3 * N ** 3)
        case.backward_view_orth(y, coef, n, tmp, fac, cond=cond)
        lhs = float(np.sum(Sx * y.astype(np.float64)))
        rhs = vsf * ratio * float(np.sum(x.reshape(-1).astype(np.float64) * tmp))
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_three_direction_normal_closed_form(orc):
    # basis e_x, e_y, e_z and no rotation: the normal of the plane through
    # mu0 e_x, mu1 e_y, mu2 e_z is (mu1 mu2, mu0 mu2, mu0 mu1) normalised
    # (ir.cu:912-922); coef <- (atan2(n_x, -n_y), n_z) and loss = |e_z x n|.
    case = orc.Case(SDD, SOD, DET, 4, 8, 4, basis=np.eye(3, dtype=np.float32))
    rng = np.random.default_rng(8)
    mu = rng.uniform(0.1, 1.0, (3, 4, 4, 4)).astype(np.float32)
    coef0 = np.stack([rng.uniform(-3, 3, (4, 4, 4)), np.ones((4, 4, 4))]).astype(np.float32)  # R = I
    m = mu.astype(np.float64)
    for rot, s1, s2 in ((0, 1, 1), (1, -1, 1), (2, -1, -1), (3, 1, -1)):
        n = np.stack([s1 * s2 * m[1] * m[2], s2 * m[0] * m[2], s1 * m[0] * m[1]])
        n *= np.where(n[2] < 0, -1.0, 1.0)
        n /= np.linalg.norm(n, axis=0)
        coef, loss = case.orth_normal(mu, coef0, rot)
        assert np.allclose(coef[1], n[2], atol=1e-12)
        assert np.allclose(coef[0], np.arctan2(n[0], -n[1]), atol=1e-6)
        assert np.allclose(loss.reshape(4, 4, 4), np.hypot(n[0], n[1]), atol=1e-6)
        # the signs from an estimate channel: rot r <-> est value r
        est = np.stack([np.zeros((4, 4, 4)), np.full((4, 4, 4), float(rot))]).astype(np.float32)
        c2, _ = case.orth_normal(mu, coef0, 0, est=est)
        assert np.array_equal(c2, coef)


def test_normal_round_trips_through_md(orc):
    # convertNormVector (ir.cu:1448-1480) of the coefficients calcNormalVectorThreeDirec
    # writes gives mean(mu) times the normal: R(phi, cos) e_z = (sin t sin phi, -sin t cos phi, cos t)
    case = orc.Case(SDD, SOD, DET, 4, 8, 4, basis=BASIS3)
    rng = np.random.default_rng(9)
    mu = rng.uniform(0.1, 1.0, (3, 4, 4, 4)).astype(np.float32)
    coef0 = np.stack([np.zeros((4, 4, 4)), np.ones((4, 4, 4))]).astype(np.float32)
    coef, _ = case.orth_normal(mu, coef0, 0)
    md = orc.orth_md(mu, coef)
    c = coef.astype(np.float64)
    st = np.sqrt(1 - c[1] ** 2)
    n = np.stack([st * np.sin(c[0]), -st * np.cos(c[0]), c[1]])
    assert np.allclose(md, mu.astype(np.float64).mean(axis=0) * n, atol=1e-6)
    # and that unit vector is the normal of the three basis-weighted directions
    S = BASIS3.astype(np.float64)
    v1 = mu[0, None] * S[0][:, None, None, None] - mu[1, None] * S[1][:, None, None, None]
    v2 = mu[0, None] * S[0][:, None, None, None] - mu[2, None] * S[2][:, None, None, None]
    w = np.cross(v1, v2, axis=0)
    w *= np.where(w[2] < 0, -1.0, 1.0)
    w /= np.linalg.norm(w, axis=0)
    assert np.allclose(n, w, atol=1e-6)


def test_update_estimate_keeps_the_best(orc):
    case = orc.Case(SDD, SOD, DET, 4, 8, 4, basis=np.eye(3, dtype=np.float32))
    rng = np.random.default_rng(10)
    mu = rng.uniform(0.1, 1.0, (3, 4, 4, 4)).astype(np.float32)
    m = mu.astype(np.float64)
    n = np.stack([m[1] * m[2], m[0] * m[2], m[0] * m[1]])
    cos = n[2] / np.linalg.norm(n, axis=0)
    est = np.stack([np.full((4, 4, 4), 0.5), np.full((4, 4, 4), 7.0)]).astype(np.float32)
    out = case.orth_update_estimate(mu, est, 3)
    better = cos > 0.5
    assert np.allclose(out[0][better], cos[better], atol=1e-12) and np.all(out[1][better] == 3)
    assert np.all(out[0][~better] == 0.5) and np.all(out[1][~better] == 7)


def test_mean_filter_uniform_field_is_fixed(orc):
    # meanFiltFiber (ir.cu:1231-1303): equal orientations everywhere sum to the
    # same orientation; weight 0 keeps the centre's own
    N = 6
    rng = np.random.default_rng(11)
    phi, ct = 0.4, 0.8
    src = np.stack([np.full((N, N, N), phi), np.full((N, N, N), ct)]).astype(np.float32)
    vol = rng.uniform(0.5, 1.0, (3, N, N, N)).astype(np.float32)
    out = orc.orth_mean_filter(src, vol, 1.0)
    assert np.allclose(out, src, atol=1e-6)
    src2 = np.stack([rng.uniform(-3, 3, (N, N, N)), rng.uniform(0.1, 1, (N, N, N))]).astype(np.float32)
    out2 = orc.orth_mean_filter(src2, vol, 0.0)
    # phi is stored shifted by pi/2 (the filter reads phi - pi/2 and writes the axis angle)
    inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    assert np.allclose(out2[1][inner[1:]], src2[1][inner[1:]], atol=1e-6)
    d = np.angle(np.exp(1j * (out2[0][inner[1:]].astype(np.float64) - src2[0][inner[1:]])))
    assert np.allclose(d, 0.0, atol=1e-5)
