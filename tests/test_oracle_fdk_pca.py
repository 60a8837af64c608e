"""Pins the oracle's FDK pieces (fdk.cu:13-170, reconstruct.cu:1651-1724) and PCA
(direction/src/pca.cu:12-65) against textbook routines and closed forms."""
import os

import numpy as np
import pytest

from conftest import rel_l2

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _d(case):
    g = case.g
    return float(np.float32(np.float32(g.det_size) * np.float32(np.float32(g.sod) / np.float32(g.sdd))))


@pytest.mark.parametrize("kind", [0, 1])
def test_filter_frequency_response(orc, kind):
    # reference kernel 1/(pi^2 d (1-4m^2)) = (d/2) x Shepp-Logan -> sin(pi f d)/(2 pi d);
    # Ram-Lak in the same convention -> |f|/2 (sample spacing d, 1/2 of a full turn).
    case = orc.Case(1000.0, 500.0, 0.4, 16, 512, 8)
    d = _d(case)
    h = case.fdk_taps(kind)
    m = np.arange(1, len(h))
    for fr in (0.05, 0.1, 0.2, 0.3, 0.45):
        f = fr / d
        H = h[0] + 2 * np.sum(h[1:] * np.cos(2 * np.pi * f * m * d))
        want = np.sin(np.pi * f * d) / (2 * np.pi * d) if kind == 0 else f / 2
        assert abs(H / want - 1) < 1e-4


def test_weight_is_cosine_of_ray_angle(orc):
    case = orc.Case(600.0, 300.0, 0.8, 8, 40, 4, V=30)
    w = case.fdk_weight()
    sdd, det = 600.0, float(np.float32(0.8))
    for (u, v) in [(0, 0), (19, 14), (20, 15), (39, 29), (7, 22)]:
        ur, vr = (u + 0.5 - 20) * det, (v + 0.5 - 15) * det
        cosang = sdd / np.linalg.norm([ur, sdd, vr])
        assert abs(w[v, u] - cosang) < 1e-12
    assert np.allclose(w, w[::-1, ::-1])  # symmetric about the detector centre


@pytest.mark.parametrize("kind", [0, 1])
def test_row_filter_is_a_convolution(orc, kind):
    case = orc.Case(600.0, 300.0, 0.8, 8, 48, 4, V=6)
    rng = np.random.default_rng(7)
    view = rng.random((6, 48)).astype(np.float32)
    got = case.fdk_filter_view(view, kind)
    h = case.fdk_taps(kind)
    w = case.fdk_weight()
    U = 48
    hs = np.concatenate([h[:0:-1], h])  # h[|k|], k = -(U-1) .. U-1
    for v in range(6):
        want = np.convolve(view[v].astype(np.float64) * w[v], hs)[U - 1: 2 * U - 1]
        assert rel_l2(got[v], want) < 1e-12


def test_fdk_uniform_cylinder(orc):
    # FDK of a projected uniform cylinder recovers its attenuation in the interior
    # (reference preset voxel size, so the projector's ratio is 1).
    N, U, P = 40, 80, 120
    case = orc.Case(600.0, 300.0, 0.8, N, U, P)
    vs = case.g.vox_size
    c = (N + 1) / 2
    z, y, x = np.meshgrid(*(np.arange(N) - c,) * 3, indexing="ij")
    mu = 0.03
    vol = np.where((x * x + y * y <= (0.3 * N) ** 2) & (np.abs(z) <= 0.25 * N), mu, 0.0).astype(np.float32)
    sino = np.stack([case.forward_view(vol, n)[0] for n in range(P)]).astype(np.float32)
    for kind in (0, 1):
        rec, _ = case.fdk_reconstruct(sino[None], kind=kind)
        # FDK uses the (N-1)/2 centre: the object sits one voxel further along each axis.
        core = rec[0][int(c) - 4 + 1:int(c) + 4 + 1, int(c) - 6 + 1:int(c) + 6 + 1, int(c) - 6 + 1:int(c) + 6 + 1]
        assert abs(core.mean() / mu - 1) < 0.05, (kind, core.mean())


def test_eig3_matches_numpy_golden(orc):
    data = np.load(os.path.join(GOLDEN, "eig3_numpy.npz"))
    for A, w, V, degenerate in zip(data["A"], data["w"], data["V"], data["degenerate"]):
        gw, gV = orc.eig3(A)
        assert np.allclose(gw, w, rtol=0, atol=1e-12 * max(1.0, np.abs(w).max()))
        assert np.allclose(gV.T @ gV, np.eye(3), atol=1e-12)
        assert np.allclose(A @ gV, gV * gw, atol=1e-11 * max(1.0, np.abs(w).max()))
        if not degenerate:
            for i in range(3):
                assert abs(abs(gV[:, i] @ V[:, i]) - 1) < 1e-10


def test_pca_axis_basis_known_answer(orc):
    # basis = the three axes: C = diag(a^2, b^2, c^2)/3, the smallest-eigenvalue axis
    # is the one with the smallest |mu|; md = mean(mu) * that axis (z >= 0).
    basis = np.eye(3)
    mu = np.array([[0.3, 0.1, 0.5], [0.2, 0.4, 0.05], [0.6, 0.7, 0.3]], dtype=np.float32)  # (K, nvox)
    md, ev = orc.pca(mu, basis)
    for i in range(3):
        k = int(np.argmin(np.abs(mu[:, i])))
        want = np.zeros(3)
        want[k] = mu[:, i].astype(np.float64).mean()
        assert np.allclose(md[:, i], want, atol=1e-12)
        sq = np.sort(mu[:, i].astype(np.float64) ** 2)
        assert np.allclose(ev[:, i], sq / sq.sum(), atol=1e-12)


def test_pca_random_against_numpy(orc):
    rng = np.random.default_rng(8)
    K, nv = 13, 200
    basis = rng.normal(size=(K, 3))
    basis /= np.linalg.norm(basis, axis=1, keepdims=True)
    basis = basis.astype(np.float32)
    mu = rng.random((K, nv)).astype(np.float32)
    md, ev = orc.pca(mu, basis)
    b = basis.astype(np.float64)
    for i in range(nv):
        m = mu[:, i].astype(np.float64)
        s = m[:, None] * b
        C = s.T @ s / K
        w, V = np.linalg.eigh(C)
        e = V[:, 0] * (1 if V[2, 0] >= 0 else -1)
        assert np.allclose(md[:, i], m.mean() * e, atol=1e-9)
        assert np.allclose(ev[:, i], w / w.sum(), atol=1e-12)


def test_fdk_distance_weight_closed_form(orc):
    # FDK backprojection of a constant filtered image c over one view n adds
    # c * (sum of bilinear weights = 1) * Ud^2 * 2 pi / P to every accepted
    # voxel, Ud = sod / (sod + y_v) with y_v the voxel's UN-rotated y
    # coordinate, (y - (N-1)/2) vs - offset_y - t_y (fdk.cu:129-135, 160-169).
    # At 90 degrees the rotated y would be the voxel's x: a variant using it,
    # a sign slip (sod - y) or a missing weight all fail below.
    N, U, P = 12, 64, 8
    ty = 0.37
    case = orc.Case(600.0, 300.0, 0.8, N, U, P, cond_t=[0.0, ty, 0.0])
    vs = case.g.vox_size
    sod = case.g.sod
    c = 0.75
    img = np.full((U, U), c)
    idx = np.arange(N) - (N - 1) / 2.0
    yv = idx[None, :, None] * vs - ty  # [z][y][x] -> depends on y only
    xv = idx[None, None, :] * vs
    for n in (0, 2):  # 0 and 90 degrees
        vol = np.zeros(N ** 3)
        case.fdk_backproject_view(img, n, vol)
        vol = vol.reshape(N, N, N)
        want = c * (sod / (sod + yv)) ** 2 * 2 * np.pi / P * np.ones((N, N, N))
        assert np.allclose(vol, want, rtol=1e-6), n  # FP32 geometry inputs
        if n == 2:
            rotated = c * (sod / (sod + xv - ty)) ** 2 * 2 * np.pi / P * np.ones((N, N, N))
            assert np.abs(vol - rotated).max() > 1e-3 * np.abs(want).max()
        flipped = c * (sod / (sod - yv)) ** 2 * 2 * np.pi / P
        assert np.abs(vol - flipped).max() > 1e-3 * np.abs(want).max()
        assert np.abs(vol - c * 2 * np.pi / P).max() > 1e-3 * np.abs(want).max()


def hilbert_numpy(rows):
    """The reference's Hilbert chain with numpy's FFT: rfft, every half-spectrum
    bin times -i/(U/2+1), irfft (which drops the imaginary parts of the DC and
    Nyquist bins as cuFFT's C2R does) times U (cuFFT's inverse is unnormalised)."""
    U = rows.shape[-1]
    X = np.fft.rfft(rows.astype(np.float64), axis=-1)
    return np.fft.irfft(X * (-1j / (U // 2 + 1)), n=U, axis=-1) * U


@pytest.mark.parametrize("U", [64, 63, 96])
def test_hilbert_filter_matches_numpy_fft(orc, U):
    # the oracle's explicit-DFT restatement (fdk.cu:298-375) against numpy's FFT
    V = 5
    case = orc.Case(600.0, 300.0, 0.8, 8, U, 4, V=V)
    rng = np.random.default_rng(U)
    view = rng.standard_normal((V, U)).astype(np.float32)
    got = case.fdk_filter_view(view, kind=2).reshape(V, U)
    want = hilbert_numpy(view)
    assert rel_l2(got, want) <= 1e-12
    got32 = case.fdk_filter_view(view, kind=2, prec=0).reshape(V, U)
    assert rel_l2(got32, want) <= 1e-5


def test_hilbert_filter_closed_forms(orc):
    # the Hilbert transform of a discrete cosine is the sine (times U/(U/2+1) here);
    # DC and the Nyquist row go to zero
    U, V = 64, 4
    case = orc.Case(600.0, 300.0, 0.8, 8, U, 4, V=V)
    u = np.arange(U)
    rows = np.stack([np.cos(2 * np.pi * 3 * u / U), np.cos(2 * np.pi * 11 * u / U), np.ones(U),
                     np.cos(np.pi * u)]).astype(np.float32)
    got = case.fdk_filter_view(rows, kind=2).reshape(V, U)
    s = U / (U // 2 + 1)
    assert np.allclose(got[0], s * np.sin(2 * np.pi * 3 * u / U), atol=1e-6)
    assert np.allclose(got[1], s * np.sin(2 * np.pi * 11 * u / U), atol=1e-6)
    assert np.abs(got[2]).max() < 1e-6 and np.abs(got[3]).max() < 1e-5
    # a sign slip (+i) or the 1/U normalisation would fail the first two rows
