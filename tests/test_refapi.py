"""The reference-compatible C++ API (include/xr/reconstruct.hpp) as a drop-in:
a program written against it (tests/cpp/refapi_demo.cpp, the calls of the
reference's main.cu) builds, links libxr.so, and on the GPU produces exactly
what the same calls through the Python binding produce."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "settings_fixture.json")


def test_refapi_demo_builds():
    import __graft_entry__ as entry

    out = entry.build_refapi_demo()
    assert os.path.exists(out)


def synthetic_sino(U, V, P, c):
    """The demo's data, in its FP32 arithmetic (0.05f + 0.01f * k)."""
    n, v, u = np.meshgrid(np.arange(P), np.arange(V), np.arange(U), indexing="ij")
    k = ((u * 7 + v * 3 + n + c) % 11).astype(np.float32)
    return np.float32(0.05) + np.float32(0.01) * k


@pytest.mark.gpu
def test_refapi_demo_matches_python_binding(xr, ctx, tmp_path):
    exe = os.path.join(ROOT, "tests", "cpp", "refapi_demo")
    assert os.path.exists(exe), "build() compiles tests/cpp/refapi_demo"
    env = dict(os.environ, XR_LOSS_CSV="")  # no loss file from this run
    vol_path = str(tmp_path / "ct")
    r = subprocess.run([exe, FIXTURE, "bench_tilted", vol_path], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    st, g, scan = xr.load_settings(FIXTURE, "bench_tilted")
    C, K, N = st.n_cond, st.n_basis, st.n_vox
    sino = [synthetic_sino(st.n_det_u, st.n_det_v, st.n_proj, c).reshape(-1) for c in range(C)]
    vol = [np.full(N ** 3, 0.01, np.float32) for _ in range(C)]
    ctx.ir_reconstruct(g, scan, [s.copy() for s in sino], vol, 2, 4)
    ct = [np.full(N ** 3, 0.01, np.float32) for _ in range(K)]
    md = [np.zeros(N ** 3, np.float32) for _ in range(3)]
    s2 = [s.copy() for s in sino]
    ctx.xtt_reconstruct(g, scan, s2, ct, md, 1, 4)
    fv = [np.zeros(N ** 3, np.float32) for _ in range(C)]
    ctx.fdk_reconstruct(g, scan, s2, fv)  # FDK reads the sinogram XTT overwrote, as the demo does
    # gradReconstruct reads sinogram[0] as FDK left it; blockSize 32 = N, so
    # the x = N and z = N layers are skipped as in the reference's launches
    gv = np.zeros(N ** 3, np.float32)
    grads = [np.zeros((N + 1) ** 3, np.float32) for _ in range(3 * C)]
    ctx.grad_reconstruct(g, scan, s2[0], gv, 1, 4, method=xr.METHOD_ART, lam=1e-2, block_size=32, grads=grads)
    assert got["voxel_after"] == N + 1  # geom.voxel + 1, as the reference leaves it
    # 3*min(NUM_BASIS_VECTOR, NUM_PROJ_COND) gradient volumes saved
    for i in range(3 * min(K, C)):
        path = vol_path + "grad%d_%dx%dx%d.raw" % (i + 1, N + 1, N + 1, N + 1)
        saved = np.fromfile(path, dtype=np.float32)
        assert saved.size == (N + 1) ** 3
        assert np.linalg.norm(saved - grads[i]) <= 1e-6 * np.linalg.norm(grads[i])
    assert not os.path.exists(vol_path + "grad%d_%dx%dx%d.raw" % (3 * min(K, C) + 1, N + 1, N + 1, N + 1))
    want = {"grad": float(gv.astype(np.float64).sum()),
            "ir": sum(float(v.astype(np.float64).sum()) for v in vol),
            "xtt": sum(float(v.astype(np.float64).sum()) for v in ct),
            "md": sum(float(np.abs(v).astype(np.float64).sum()) for v in md),
            "fdk": sum(float(v.astype(np.float64).sum()) for v in fv)}
    for k in want:
        assert abs(got[k] - want[k]) <= 1e-6 * max(abs(want[k]), 1e-12), (k, got[k], want[k])


def flow_settings(tmp, N=24, U=40, P=16):
    """A settings.json tag for the driver flow: paths into `tmp`, two untilted
    conditions (sorting with index [1, 0]), the 7 basis vectors of the
    reference's recon_variable."""
    base = [0.666667, -0.666667, -0.333333, -0.333333, -0.666667, 0.666667, 0.666667, 0.333333, 0.666667,
            0.57735, -0.57735, 0.57735, 0.19245, 0.96225, 0.19245, -0.19245, 0.19245, 0.96225,
            0.96225, 0.19245, -0.19245]
    data = {"recon_variable": {"blockSize": 32, "vector": 7, "condition": 2, "base": base},
            "flow": {"proj_path": os.path.join(tmp, "proj", "scan"), "vol_path": os.path.join(tmp, "vol", "ct"),
                     "direc_path": os.path.join(tmp, "vol", "md"), "sod": 300.0, "sdd": 600.0, "proj": P,
                     "num_det_u": U, "num_det_v": U, "det_size": 0.8, "num_voxel": N, "rot": 2,
                     "matrixRot": [1, 0, 0, 0, 1, 0, 0, 0, 1] * 2, "vecTrans": [0, 0, 0, 0, 0, 0],
                     "areaTrans": [0, 0, 0], "offset": [0, 0, 0, 0, 0, 0], "index": [1, 0], "sorting": True}}
    path = os.path.join(tmp, "settings.json")
    with open(path, "w") as f:
        json.dump(data, f)
    return path, data


def test_main_flow_file_names(tmp_path):
    """The driver's raw file names (main.cu:31-32, 83-107), printed by the
    C++ helpers without touching the GPU, against the names main.cu builds."""
    exe = os.path.join(ROOT, "tests", "cpp", "main_flow")
    if not os.path.exists(exe):
        import __graft_entry__ as entry

        entry.build_refapi_demo()
    path, data = flow_settings(str(tmp_path))
    env = dict(os.environ)
    env.pop("XR_LOSS_CSV", None)
    r = subprocess.run([exe, path, "flow", "--paths"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=60, env=env)
    assert r.returncode == 0, r.stderr
    got = [line.split(" ", 1) for line in r.stdout.strip().splitlines()]
    t = data["flow"]
    U, V, P, N = t["num_det_u"], t["num_det_v"], t["proj"], t["num_voxel"]
    sino_dims = "_%dx%dx%d.raw" % (U, V, P)
    vol_dims = "_%dx%dx%d.raw" % (N, N, N)
    want = [("load", t["proj_path"] + "2" + sino_dims), ("load", t["proj_path"] + "1" + sino_dims),
            ("proj", t["vol_path"] + "_proj1" + sino_dims), ("proj", t["vol_path"] + "_proj2" + sino_dims)]
    want += [("volume", t["vol_path"] + "%d" % (k + 1) + vol_dims) for k in range(7)]
    want += [("direction", t["direc_path"] + "%d" % (i + 1) + vol_dims) for i in range(3)]
    want += [("angle", t["direc_path"] + "_phi_%dx%dx%d.raw" % (N, N, N)),
             ("angle", t["direc_path"] + "_theta_%dx%dx%d.raw" % (N, N, N)), ("loss", "../python/loss.csv")]
    assert [tuple(g) for g in got] == want


@pytest.mark.gpu
def test_main_flow_end_to_end(orc, xr, ctx, tmp_path):
    """main.cu's flow through the C++ API with files on disk: sinograms saved
    under the reference's names, XTT::reconstruct, the volumes, directions and
    angles saved under the reference's names, loss.csv in ../python.  The
    results equal the same solve through the Python binding."""
    tmp = str(tmp_path)
    for d in ("proj", "vol", "python", "run"):
        os.makedirs(os.path.join(tmp, d))
    path, data = flow_settings(tmp)
    st, g, scan = xr.load_settings(path, "flow")
    N, U, P, C, K = st.n_vox, st.n_det_u, st.n_proj, st.n_cond, st.n_basis
    basis = np.array(data["recon_variable"]["base"], np.float32).reshape(K, 3)
    fp, _ = xr.fibre_params(5, 12, fov_radius=0.5 * N * g.vox_size, r_lo=2 * g.vox_size, r_hi=4 * g.vox_size)
    truth = (orc.phantom_fibres(N, g.vox_size, fp, basis) * 0.02).astype(np.float32)
    case = orc.Case(g.sdd, g.sod, g.det_size, N, U, P, vox_size=g.vox_size, n_cond=C, basis=basis)
    sinos = [np.stack([case.forward_view(truth, n, cond=c, xtt=True)[0] for n in range(P)]).astype(np.float32)
             for c in range(C)]
    sinos[1][0, 0, :3] = -0.5  # negative values: the driver clamps them to 1e-8 (main.cu:35)
    # condition i loads file LOAD_INDEX[i] = index[i] + 1
    for i, idx in enumerate(data["flow"]["index"]):
        sinos[i].tofile(data["flow"]["proj_path"] + "%d_%dx%dx%d.raw" % (idx + 1, U, U, P))
    exe = os.path.join(ROOT, "tests", "cpp", "main_flow")
    env = dict(os.environ)
    env.pop("XR_LOSS_CSV", None)
    epochs, batch = 2, 4
    r = subprocess.run([exe, path, "flow", str(epochs), str(batch)], cwd=os.path.join(tmp, "run"),
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr
    # the same solve through the binding
    s = [np.where(x < 0, np.float32(1e-8), x).reshape(-1).copy() for x in sinos]
    ct = [np.full(N ** 3, 0.01, np.float32) for _ in range(K)]
    md = [np.zeros(N ** 3, np.float32) for _ in range(3)]
    losses = ctx.xtt_reconstruct(g, scan, s, ct, md, epochs, batch, lam=1e-3)
    vol_dims = "_%dx%dx%d.raw" % (N, N, N)
    t = data["flow"]
    for k in range(K):
        got = np.fromfile(t["vol_path"] + "%d" % (k + 1) + vol_dims, np.float32)
        assert np.array_equal(got, ct[k]), k
    for i in range(3):
        got = np.fromfile(t["direc_path"] + "%d" % (i + 1) + vol_dims, np.float32)
        assert np.array_equal(got, md[i]), i
    for i in range(C):  # the sinogram overwritten with the final ratios (reconstruct.cu:924-925)
        got = np.fromfile(t["vol_path"] + "_proj%d_%dx%dx%d.raw" % (i + 1, U, U, P), np.float32)
        assert np.array_equal(got, s[i]), i
    phi = np.fromfile(t["direc_path"] + "_phi_%dx%dx%d.raw" % (N, N, N), np.float32)
    theta = np.fromfile(t["direc_path"] + "_theta_%dx%dx%d.raw" % (N, N, N), np.float32)
    want = orc.md_angles(np.stack(md))
    assert np.abs(phi - want[0]).max() <= 1e-6 and np.abs(theta - want[1]).max() <= 1e-6
    with open(os.path.join(tmp, "python", "loss.csv")) as f:
        vals = [float(v) for v in f.read().strip().rstrip(",").split(",")]
    assert np.allclose(vals, np.float32(losses / (U * U * P)), rtol=1e-6)
    # directionNormalization: |md| <= 0.65 everywhere at these attenuations, so all zero,
    # as the oracle's restatement of pca.cu:85-106 gives
    unit = json.loads(r.stdout.strip().splitlines()[-1])["unit_directions"]
    assert unit == float((orc.md_normalise(np.stack(md)).astype(np.float64) ** 2).sum())
