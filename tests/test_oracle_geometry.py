"""Pins the oracle's ray casting (ir.cu:1852-1902, fdk.cu:111-148) with closed forms.

Each check is chosen so that a plausible slip fails it: the source side, the
voxel-centre convention, the rotation sign, the u/v axes and the magnification.
"""
import numpy as np
import pytest

SDD, SOD, DET = 600.0, 300.0, 0.8


@pytest.mark.parametrize("prec", [0, 1])
def test_theta_is_reference_fp32(orc, prec):
    for n, P in [(0, 90), (1, 90), (37, 90), (-5, 720), (719, 720)]:
        want = np.float32(np.float32(np.float32(2.0) * np.float32(np.pi)) * np.float32(n)) / np.float32(P)
        assert orc.theta_f32(n, P) == want


@pytest.mark.parametrize("prec", [0, 1])
def test_origin_voxel_hits_detector_centre(orc, prec):
    # odd N: voxel index (N+1)/2 is the physical origin in the IR convention.
    N, U, P = 31, 40, 36
    case = orc.Case(SDD, SOD, DET, N, U, P, vox_size=0.5)
    c = (N + 1) // 2
    for n in range(0, P, 5):
        u, v = case.raycast(c, c, c, n, prec=prec)[:2]
        assert abs(u - U / 2) < 1e-4 and abs(v - U / 2) < 1e-4


@pytest.mark.parametrize("prec", [0, 1])
def test_view0_magnification_and_axes(orc, prec):
    # theta = 0: source at +sod on y, u along x, v along z.
    N, U, P, vs = 21, 64, 8, 0.7
    case = orc.Case(SDD, SOD, DET, N, U, P, vox_size=vs)
    c = (N + 1) / 2
    for (x, y, z) in [(3, 4, 5), (17, 2, 9), (11, 18, 1), (0, 0, 20)]:
        X, Y, Z = (x - c) * vs, (y - c) * vs, (z - c) * vs
        u, v = case.raycast(x, y, z, 0, prec=prec)[:2]
        mag = SDD / (SOD - Y)
        assert abs(u - (U / 2 + X * mag / DET)) < 2e-4
        assert abs(v - (U / 2 + Z * mag / DET)) < 2e-4


@pytest.mark.parametrize("prec", [0, 1])
def test_quarter_turn_rotates_source_counter_clockwise(orc, prec):
    # view P/4 (theta = 90 deg, CW sign +1): source at Rz(90)(0,sod,0) = (-sod,0,0), u along +y.
    N, U, P, vs = 21, 64, 8, 0.7
    case = orc.Case(SDD, SOD, DET, N, U, P, vox_size=vs)
    c = (N + 1) / 2
    for (x, y, z) in [(3, 4, 5), (17, 2, 9)]:
        X, Y, Z = (x - c) * vs, (y - c) * vs, (z - c) * vs
        u, v = case.raycast(x, y, z, P // 4, prec=prec)[:2]
        mag = SDD / (SOD + X)
        assert abs(u - (U / 2 + Y * mag / DET)) < 2e-3
        assert abs(v - (U / 2 + Z * mag / DET)) < 2e-3


@pytest.mark.parametrize("prec", [0, 1])
def test_fdk_convention_source_side_and_centre(orc, prec):
    # FDK: source at -sod on y at theta = 0, centre (N-1)/2 (fdk.cu:125-135).
    N, U, P, vs = 20, 64, 8, 0.7
    case = orc.Case(SDD, SOD, DET, N, U, P, vox_size=vs)
    c = (N - 1) / 2
    for (x, y, z) in [(3, 4, 5), (17, 2, 9)]:
        X, Y, Z = (x - c) * vs, (y - c) * vs, (z - c) * vs
        u, v = case.raycast(x, y, z, 0, conv=1, prec=prec)[:2]
        mag = SDD / (SOD + Y)
        assert abs(u - (U / 2 + X * mag / DET)) < 2e-4
        assert abs(v - (U / 2 + Z * mag / DET)) < 2e-4


def test_beam_direction_and_grating(orc):
    N, U, P, vs = 21, 64, 12, 0.7
    ang = np.deg2rad(30.0)
    cR = np.array([[1, 0, 0], [0, np.cos(ang), -np.sin(ang)], [0, np.sin(ang), np.cos(ang)]])
    g = np.array([1.0, 0.0, 1.0]) / np.sqrt(2)
    case = orc.Case(SDD, SOD, DET, N, U, P, vox_size=vs, cond_R=cR.reshape(-1), grating=g)
    c = (N + 1) / 2
    for n in (0, 3, 7):
        th = float(orc.theta_f32(n, P))
        Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
        R = cR @ Rz
        out = case.raycast(4, 9, 13, n)
        B, G = out[2:5], out[5:8]
        src = R @ np.array([0.0, SOD, 0.0])
        X = (np.array([4, 9, 13]) - c) * vs
        d = X - src
        assert np.allclose(B, d / np.linalg.norm(d), atol=1e-12)
        assert np.allclose(G, R @ g, atol=1e-12)


def test_default_voxel_size_is_reference_formula(orc):
    case = orc.Case(1000.0, 500.0, 0.4, 8, 16, 4)
    assert case.g.vox_size == np.float32(np.float32(500.0) * np.float32(0.4)) / np.float32(1000.0)
