"""CPU checks of the C ABI: the library loads, exports every declared entry
point, and its host-only functions (settings reader, filter taps, phantom
parameters, error paths) behave as documented.  No GPU needed."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "xr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(xr):
    lib = ctypes.CDLL(xr.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_context_create_fails_cleanly_without_gpu(xr):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = xr.lib().xr_context_create(ctypes.byref(h), 0)
    assert rc in (-2, -1) and not h.value


def test_filter_taps_match_oracle(xr, orc):
    g = xr.make_geometry(1000.0, 500.0, 0.4, 16, 96, 8)
    case = orc.Case(1000.0, 500.0, 0.4, 16, 96, 8)
    assert np.array_equal(xr.fdk_taps(g, 0).astype(np.float64), case.fdk_taps(0, prec=0))
    assert np.allclose(xr.fdk_taps(g, 1), case.fdk_taps(1), rtol=1e-6, atol=0)


@pytest.mark.parametrize("U", [96, 95])
def test_hilbert_taps_reproduce_the_oracle_filter(xr, orc, U):
    # the library's closed-form circular taps (a sine sum) applied as a circular
    # convolution equal the oracle's DFT restatement of the FFT chain
    g = xr.make_geometry(1000.0, 500.0, 0.4, 16, U, 8)
    h = xr.fdk_taps(g, xr.FILTER_HILBERT).astype(np.float64)
    case = orc.Case(1000.0, 500.0, 0.4, 16, U, 8, V=3)
    rows = np.random.default_rng(7).standard_normal((3, U)).astype(np.float32)
    want = case.fdk_filter_view(rows, kind=2).reshape(3, U)
    m = (np.arange(U)[:, None] - np.arange(U)[None, :]) % U
    got = rows.astype(np.float64) @ h[m].T
    assert np.allclose(got, want, rtol=0, atol=1e-6 * np.abs(want).max())


def test_settings_sorting_and_area_translation(xr):
    st, g, scan = xr.load_settings(os.path.join(GOLDEN, "settings_fixture.json"), "bench_tilted")
    assert (st.n_cond, st.n_basis, st.n_vox, st.n_proj, st.n_det_u, st.n_det_v) == (2, 3, 32, 60, 64, 48)
    assert list(st.load_index[:2]) == [3, 1]  # index [2, 0] -> 1-based file indices
    R = scan.cond_R.reshape(2, 3, 3)
    assert np.array_equal(R[0], [[0, 0, 1], [0, 1, 0], [-1, 0, 0]])
    assert np.array_equal(R[1], np.eye(3))
    t = scan.cond_t.reshape(2, 3)
    at = np.float32([0.1, 0.2, -0.3])
    assert np.array_equal(t[0], np.float32([-0.75, 0.5, 1.5]) + at)
    assert np.array_equal(t[1], np.float32([0, 0, 0]) + at)
    assert np.array_equal(scan.cond_offset.reshape(2, 3)[0], np.float32([0, 0.2, 0]))
    assert abs(g.vox_size - np.float32(400.0) * np.float32(0.5) / np.float32(800.0)) == 0
    assert st.proj_path == b"proj/bench"


def test_settings_without_sorting_key(xr):
    st, g, scan = xr.load_settings(os.path.join(GOLDEN, "settings_fixture.json"), "no_sorting")
    assert st.n_cond == 2 and list(st.load_index[:2]) == [1, 2]
    assert np.array_equal(scan.cond_R.reshape(2, 3, 3)[1], [[0, -1, 0], [1, 0, 0], [0, 0, 1]])


def test_settings_errors(xr):
    lib = xr.lib()
    st = xr.Settings()
    assert lib.xr_load_settings(b"/nonexistent/settings.json", b"x", ctypes.byref(st)) == -4
    path = os.path.join(GOLDEN, "settings_fixture.json").encode()
    assert lib.xr_load_settings(path, b"no_such_tag", ctypes.byref(st)) == -5


def test_reference_settings_file_parses(xr):
    path = "/root/reference/utility/settings.json"
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    st, g, scan = xr.load_settings(path, "oilpan")
    assert (st.n_cond, st.n_basis, st.n_vox, st.n_proj) == (7, 7, 300, 360)
    R = scan.cond_R.reshape(-1, 3, 3).astype(np.float64)
    for r in R:  # calibrated rotations are orthonormal to ~1e-5
        assert np.allclose(r @ r.T, np.eye(3), atol=1e-4)


def test_ellipsoid_parameters(xr):
    a = xr.ellipsoid_params(42, 50, fov_radius=10.0, centre_frac=0.6, axis_lo=0.1, axis_hi=0.35)
    b = xr.ellipsoid_params(42, 50, fov_radius=10.0, centre_frac=0.6, axis_lo=0.1, axis_hi=0.35)
    c = xr.ellipsoid_params(43, 50, fov_radius=10.0, centre_frac=0.6, axis_lo=0.1, axis_hi=0.35)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert (np.linalg.norm(a[:, :3], axis=1) <= 6.0 + 1e-12).all()
    assert ((a[:, 3:6] >= 1.0) & (a[:, 3:6] <= 3.5)).all()
    for q in a:
        R = q[6:15].reshape(3, 3)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12) and abs(np.linalg.det(R) - 1) < 1e-12
    assert ((a[:, 15] >= 0.0) & (a[:, 15] <= 0.05)).all()


REF_SETTINGS = "/root/reference/utility/settings.json"


def restate_init_params(data, tag):
    """params.cu:42-104 restated with Python's json reader (independent of the
    library's own reader): FP32 vecTrans += areaTrans for every rotation, then
    the per-condition tables selected by `index` when `sorting` is true (a
    missing key reads as false; the reference's lookup would throw there)."""
    t, rv = data[tag], data["recon_variable"]
    C, K = rv["condition"], rv["vector"]
    vT = np.array(t["vecTrans"], np.float32).reshape(-1, 3)
    vT[: t["rot"]] += np.array(t["areaTrans"], np.float32)
    mR = np.array(t["matrixRot"], np.float32).reshape(-1, 9)
    off = np.array(t["offset"], np.float32).reshape(-1, 3)
    idx = [int(i) for i in t["index"][:C]] if t.get("sorting", False) else list(range(C))
    return dict(n_cond=C, n_basis=K, base=np.array(rv["base"][: 3 * K], np.float32), load_index=[i + 1 for i in idx],
                R=mR[idx].reshape(-1), T=vT[idx].reshape(-1), O=off[idx].reshape(-1),
                geom=(t["sod"], t["sdd"], t["proj"], t["num_det_u"], t["num_det_v"], t["det_size"], t["num_voxel"]),
                paths=(t["proj_path"], t["vol_path"], t["direc_path"]))


@pytest.mark.parametrize("tag", ["oilpan", "phaseCT", "cfrp_7d_13rot", "cfrp_7d_6rot", "simulation"])
def test_reference_settings_every_tag(xr, tag, tmp_path):
    """Every tag of the reference's utility/settings.json loads through
    xr_load_settings exactly as init_params would set it up."""
    import json

    if not os.path.exists(REF_SETTINGS):
        pytest.skip("reference tree not present")
    with open(REF_SETTINGS) as f:
        data = json.load(f)
    path = REF_SETTINGS
    t = data[tag]
    if not t.get("sorting", False) and data["recon_variable"]["condition"] > t["rot"]:
        # recon_variable.condition (7) exceeds this tag's rotations without an
        # index: the reference would run the extra conditions on zero-filled
        # constant tables (params.cu:88-94); the library refuses instead ...
        lib = xr.lib()
        assert lib.xr_load_settings(path.encode(), tag.encode(), ctypes.byref(xr.Settings())) == -1
        # ... and loads the tag once the condition count matches its rotations
        data["recon_variable"]["condition"] = t["rot"]
        path = os.path.join(tmp_path, "settings.json")
        with open(path, "w") as f:
            json.dump(data, f)
    want = restate_init_params(data, tag)
    st, g, scan = xr.load_settings(path, tag)
    assert (st.n_cond, st.n_basis) == (want["n_cond"], want["n_basis"])
    assert list(st.load_index[: st.n_cond]) == want["load_index"]
    assert np.array_equal(scan.cond_R, want["R"])
    assert np.array_equal(scan.cond_t, want["T"])
    assert np.array_equal(scan.cond_offset, want["O"])
    assert np.array_equal(np.ctypeslib.as_array(st.basis)[: 3 * st.n_basis], want["base"])
    sod, sdd, P, U, V, det, N = want["geom"]
    assert (st.n_proj, st.n_det_u, st.n_det_v, st.n_vox) == (P, U, V, N)
    assert (st.sod, st.sdd, st.det_size) == (np.float32(sod), np.float32(sdd), np.float32(det))
    assert (g.n_vox, g.n_det_u, g.n_det_v, g.n_proj) == (N, U, V, P)
    assert g.vox_size == np.float32(np.float32(sod) * np.float32(det)) / np.float32(sdd)
    assert (st.proj_path.decode(), st.vol_path.decode(), st.direc_path.decode()) == want["paths"]
