"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the reference results to compare against or the timed CPU baseline.  ``prec`` selects the
reference's FP32 arithmetic (REF32) or the same operator in FP64 (F64).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF32, F64 = 0, 1

_f = ctypes.POINTER(ctypes.c_float)
_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)


class OrcGeometry(ctypes.Structure):
    _fields_ = [("sdd", ctypes.c_float), ("sod", ctypes.c_float), ("vox_size", ctypes.c_float),
                ("det_size", ctypes.c_float), ("n_vox", ctypes.c_int32), ("n_det_u", ctypes.c_int32),
                ("n_det_v", ctypes.c_int32), ("n_proj", ctypes.c_int32)]


class OrcScan(ctypes.Structure):
    _fields_ = [("n_cond", ctypes.c_int32), ("n_basis", ctypes.c_int32), ("cond_R", _f), ("cond_t", _f),
                ("cond_offset", _f), ("basis", _f), ("grating", _f)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    G, S, I, F = ctypes.POINTER(OrcGeometry), ctypes.POINTER(OrcScan), ctypes.c_int, ctypes.c_float
    sig = {
        "orc_default_vox_size": (F, [F, F, F]),
        "orc_theta_f32": (F, [I, I]),
        "orc_raycast": (None, [I, G, S, I, I, I, I, I, I, _d]),
        "orc_forward_view": (None, [I, G, S, I, I, _f, I, I, _d, _d]),
        "orc_chord_length": (None, [I, G, S, I, I, _d]),
        "orc_backward_view": (None, [I, G, S, I, I, _f, I, _d, _d]),
        "orc_forward_view_orth": (None, [I, G, S, I, I, _f, _f, I, _d, _d]),
        "orc_backward_view_orth": (None, [I, G, S, I, I, _f, _f, _d, _d]),
        "orc_rodrigues": (None, [ctypes.c_double] * 5 + [_d]),
        "orc_orth_normal": (None, [I, S, ctypes.c_int64, _f, _f, _f, I, _f]),
        "orc_orth_update_estimate": (None, [I, S, ctypes.c_int64, _f, _f, I]),
        "orc_orth_md": (None, [I, ctypes.c_int64, _f, _f, _f]),
        "orc_orth_mean_filter": (None, [I, I, _f, _f, F, _f]),
        "orc_orth_reconstruct": (None, [I, G, S, _f, _f, I, I, I, I, I, F, ctypes.c_uint32, _d, _d]),
        "orc_subset_orders": (None, [I, I, I, _i32]),
        "orc_ratio_step": (ctypes.c_double, [I, ctypes.c_int64, _f, _d, I]),
        "orc_ir_reconstruct": (None, [I, G, S, _f, _f, I, I, I, I, F, _d, _d]),
        "orc_grad_forward_view": (None, [I, G, S, I, I, _f, I, _d]),
        "orc_grad_backward_view": (None, [I, G, S, I, I, _f, I, _d, _d]),
        "orc_grad_coef": (ctypes.c_double, [I, I, I, I]),
        "orc_grad_reconstruct": (None, [I, G, S, _f, I, I, I, I, F, I, _d, _d]),
        "orc_unet_num_params": (ctypes.c_int64, [I]),
        "orc_conv3": (None, [_f, I, I, I, I, _f, _f, I, I, _d]),
        "orc_unet_forward": (None, [_f, I, _f, I, I, I, F, I, _f]),
        "orc_unet_loss_grad": (ctypes.c_double, [_f, I, _f, _f, I, I, I, F, I, _d]),
        "orc_xtt_reconstruct": (None, [I, G, S, _f, _f, I, I, I, I, F, _d, _d, _d]),
        "orc_fdk_weight": (None, [I, G, _d]),
        "orc_fdk_taps": (None, [I, G, I, _d]),
        "orc_fdk_filter_view": (None, [I, G, I, _f, _d]),
        "orc_fdk_backproject_view": (None, [I, G, S, I, I, _d, _d]),
        "orc_fdk_reconstruct": (None, [I, G, S, _f, I, I, _d, _d]),
        "orc_pca": (None, [_f, I, _f, ctypes.c_int64, _d, _d]),
        "orc_eig3": (None, [_d, _d, _d]),
        "orc_md_angles": (None, [_f, _f, ctypes.c_int64]),
        "orc_md_normalise": (None, [_f, ctypes.c_int64, ctypes.c_float]),
        "orc_phantom_ellipsoids": (None, [I, ctypes.c_double, _d, I, _f]),
        "orc_num_threads": (I, []),
        "orc_noise_poisson": (None, [_f, _f, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64]),
        "orc_noise_gaussian": (None, [_f, _f, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64]),
        "orc_phantom_cylinder_shell": (None, [I, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _f]),
        "orc_phantom_fibres": (None, [I, ctypes.c_double, _d, I, _f, I, _f]),
    }
    for n, (r, a) in sig.items():
        fn = getattr(lib, n)
        fn.restype, fn.argtypes = r, a
    return lib


_lib = _load()


def _fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_f)


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_d)


class Case:
    """Geometry + scan tables for the oracle, from plain Python values."""

    def __init__(self, sdd, sod, det, N, U, P, vox_size=None, V=None, cond_R=None, cond_t=None, cond_offset=None,
                 basis=None, grating=None, n_cond=1):
        self.g = OrcGeometry()
        self.g.sdd, self.g.sod, self.g.det_size = sdd, sod, det
        self.g.vox_size = _lib.orc_default_vox_size(sdd, sod, det) if vox_size is None else vox_size
        self.g.n_vox, self.g.n_det_u = N, U
        self.g.n_det_v = U if V is None else V
        self.g.n_proj = P
        if cond_R is not None:
            n_cond = len(np.asarray(cond_R).reshape(-1)) // 9
        self.n_cond = n_cond
        self.cond_R = np.ascontiguousarray(np.tile(np.eye(3).reshape(-1), n_cond) if cond_R is None else cond_R,
                                           dtype=np.float32).reshape(-1)
        self.cond_t = np.ascontiguousarray(np.zeros(3 * n_cond) if cond_t is None else cond_t,
                                           dtype=np.float32).reshape(-1)
        self.cond_offset = np.ascontiguousarray(np.zeros(3 * n_cond) if cond_offset is None else cond_offset,
                                                dtype=np.float32).reshape(-1)
        self.basis = None if basis is None else np.ascontiguousarray(basis, dtype=np.float32).reshape(-1)
        self.grating = None if grating is None else np.ascontiguousarray(grating, dtype=np.float32).reshape(-1)
        self.s = OrcScan()
        self.s.n_cond = n_cond
        self.s.n_basis = 0 if self.basis is None else len(self.basis) // 3
        self.s.cond_R, self.s.cond_t, self.s.cond_offset = _fp(self.cond_R), _fp(self.cond_t), _fp(self.cond_offset)
        self.s.basis = _fp(self.basis) if self.basis is not None else None
        self.s.grating = _fp(self.grating) if self.grating is not None else None

    @property
    def N(self):
        return self.g.n_vox

    @property
    def U(self):
        return self.g.n_det_u

    @property
    def V(self):
        return self.g.n_det_v

    @property
    def P(self):
        return self.g.n_proj

    def raycast(self, x, y, z, n, cond=0, conv=0, prec=F64):
        out = np.zeros(8)
        _lib.orc_raycast(prec, self.g, self.s, cond, x, y, z, n, conv, _dp(out))
        return out

    def forward_view(self, vol, n, cond=0, xtt=False, normalise=True, prec=F64):
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        p = np.zeros(self.U * self.V)
        f = np.zeros(self.U * self.V)
        _lib.orc_forward_view(prec, self.g, self.s, cond, n, _fp(vol), int(xtt), int(normalise), _dp(p), _dp(f))
        return p.reshape(self.V, self.U), f.reshape(self.V, self.U)

    def chord_length(self, n, cond=0, prec=F64):
        out = np.zeros(self.U * self.V)
        _lib.orc_chord_length(prec, self.g, self.s, cond, n, _dp(out))
        return out.reshape(self.V, self.U)

    def backward_view(self, proj_view, n, tmp, factor, cond=0, xtt=False, prec=F64):
        pv = np.ascontiguousarray(proj_view, dtype=np.float32).reshape(-1)
        _lib.orc_backward_view(prec, self.g, self.s, cond, n, _fp(pv), int(xtt), _dp(tmp), _dp(factor))

    def forward_view_orth(self, vol, coef, n, cond=0, normalise=True, prec=F64):
        """forwardOrth: the XTT operator with the basis rotated per voxel by the
        Rodrigues rotation of (phi, cos theta) = coef[0], coef[1] (reference layout)."""
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        coef = np.ascontiguousarray(coef, dtype=np.float32)
        p = np.zeros(self.U * self.V)
        f = np.zeros(self.U * self.V)
        _lib.orc_forward_view_orth(prec, self.g, self.s, cond, n, _fp(vol), _fp(coef), int(normalise), _dp(p), _dp(f))
        return p.reshape(self.V, self.U), f.reshape(self.V, self.U)

    def backward_view_orth(self, proj_view, coef, n, tmp, factor, cond=0, prec=F64):
        pv = np.ascontiguousarray(proj_view, dtype=np.float32).reshape(-1)
        coef = np.ascontiguousarray(coef, dtype=np.float32)
        _lib.orc_backward_view_orth(prec, self.g, self.s, cond, n, _fp(pv), _fp(coef), _dp(tmp), _dp(factor))

    def orth_normal(self, vol, coef, rot=0, est=None, prec=F64):
        """calcNormalVectorThreeDirec (rot) / ...WithEst (est): returns the new
        coef (2 channels) and the per-voxel loss; vol 3 channels."""
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        c = np.array(coef, dtype=np.float32, copy=True)
        n = vol.size // 3
        loss = np.zeros(n, np.float32)
        e = None if est is None else np.ascontiguousarray(est, dtype=np.float32)
        _lib.orc_orth_normal(prec, self.s, n, _fp(vol), _fp(c), _fp(loss), rot, None if e is None else _fp(e))
        return c, loss

    def orth_update_estimate(self, vol, est, it, prec=F64):
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        e = np.array(est, dtype=np.float32, copy=True)
        _lib.orc_orth_update_estimate(prec, self.s, vol.size // 3, _fp(vol), _fp(e), it)
        return e

    def orth_reconstruct(self, sino, vol0, iter1, iter2, batch, direction=0, art=False, lam=1e-2, seed=0, prec=F64):
        """XTT::orthTwiceReconstruct: (kept volume, main directions), 3 channels each."""
        sino = np.ascontiguousarray(sino, dtype=np.float32)
        vol0 = np.ascontiguousarray(vol0, dtype=np.float32)
        vol = np.zeros(vol0.size)
        md = np.zeros(vol0.size)
        _lib.orc_orth_reconstruct(prec, self.g, self.s, _fp(sino), _fp(vol0), iter1, iter2, batch, direction, int(art),
                                  lam, seed, _dp(vol), _dp(md))
        return vol.reshape(vol0.shape), md.reshape(vol0.shape)

    def ir_reconstruct(self, sino, vol0, epochs, batch, direction=0, art=False, lam=1e-2, prec=F64):
        sino = np.ascontiguousarray(sino, dtype=np.float32)
        vol0 = np.ascontiguousarray(vol0, dtype=np.float32)
        out = np.zeros(vol0.size)
        losses = np.zeros(max(1, self.n_cond * epochs))
        _lib.orc_ir_reconstruct(prec, self.g, self.s, _fp(sino), _fp(vol0), epochs, batch, direction, int(art), lam,
                                _dp(out), _dp(losses))
        return out.reshape(vol0.shape), losses[: self.n_cond * epochs]

    def _grad_geometry(self):
        g = OrcGeometry()
        for name, _ in OrcGeometry._fields_:
            setattr(g, name, getattr(self.g, name))
        g.n_vox = self.g.n_vox + 1
        return g

    def grad_extent(self, block_size):
        """x and z extent of the gradient volume that the reference's launches
        cover: min(N+1, ceil(N/blockSize)*blockSize)."""
        N = self.N
        return min(N + 1, -(-N // block_size) * block_size)

    def grad_forward_view(self, vol, n, ext=None, cond=0, prec=F64):
        """forwardProjGrad of one view: vol is the (N+1)^3 gradient volume
        ([z][y][x]); the result is the un-normalised splat of vs*ratio*voxel."""
        g1 = self._grad_geometry()
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        p = np.zeros(self.U * self.V)
        _lib.orc_grad_forward_view(prec, g1, self.s, cond, n, _fp(vol), g1.n_vox if ext is None else ext, _dp(p))
        return p.reshape(self.V, self.U)

    def grad_backward_view(self, proj_view, n, tmp, factor, ext=None, cond=0, prec=F64):
        g1 = self._grad_geometry()
        pv = np.ascontiguousarray(proj_view, dtype=np.float32).reshape(-1)
        _lib.orc_grad_backward_view(prec, g1, self.s, cond, n, _fp(pv), g1.n_vox if ext is None else ext, _dp(tmp),
                                    _dp(factor))

    def grad_reconstruct(self, sino0, epochs, batch, direction=0, art=True, lam=1e-2, block_size=32, prec=F64):
        """IR::gradReconstruct up to the Poisson solve: (3*n_cond gradient
        volumes of side N+1, slot 3*cond + diff; per-epoch loss sums)."""
        sino0 = np.ascontiguousarray(sino0, dtype=np.float32)
        N1 = self.N + 1
        grads = np.zeros(3 * self.n_cond * N1 ** 3)
        losses = np.zeros(max(1, epochs))
        _lib.orc_grad_reconstruct(prec, self.g, self.s, _fp(sino0), epochs, batch, direction, int(art), lam,
                                  block_size, _dp(grads), _dp(losses))
        return grads.reshape(3 * self.n_cond, N1, N1, N1), losses[:epochs]

    def xtt_reconstruct(self, sino, vol0, epochs, batch, direction=0, art=False, lam=1e-2, prec=F64):
        sino = np.ascontiguousarray(sino, dtype=np.float32)
        vol0 = np.ascontiguousarray(vol0, dtype=np.float32)
        out = np.zeros(vol0.size)
        proj = np.zeros(sino.size)
        losses = np.zeros(max(1, epochs))
        _lib.orc_xtt_reconstruct(prec, self.g, self.s, _fp(sino), _fp(vol0), epochs, batch, direction, int(art), lam,
                                 _dp(out), _dp(proj), _dp(losses))
        return out.reshape(vol0.shape), proj.reshape(sino.shape), losses[:epochs]

    def fdk_weight(self, prec=F64):
        out = np.zeros(self.U * self.V)
        _lib.orc_fdk_weight(prec, self.g, _dp(out))
        return out.reshape(self.V, self.U)

    def fdk_taps(self, kind=0, prec=F64):
        out = np.zeros(self.U)
        _lib.orc_fdk_taps(prec, self.g, kind, _dp(out))
        return out

    def fdk_filter_view(self, view, kind=0, prec=F64):
        v = np.ascontiguousarray(view, dtype=np.float32).reshape(-1)
        out = np.zeros(self.U * self.V)
        _lib.orc_fdk_filter_view(prec, self.g, kind, _fp(v), _dp(out))
        return out.reshape(self.V, self.U)

    def fdk_backproject_view(self, filt_view, n, vol, cond=0, prec=F64):
        fv = np.ascontiguousarray(filt_view, dtype=np.float64).reshape(-1)
        _lib.orc_fdk_backproject_view(prec, self.g, self.s, cond, n, _dp(fv), _dp(vol))

    def fdk_reconstruct(self, sino, direction=0, kind=0, prec=F64):
        sino = np.ascontiguousarray(sino, dtype=np.float32)
        vol = np.zeros(self.n_cond * self.N ** 3)
        filt = np.zeros(sino.size)
        _lib.orc_fdk_reconstruct(prec, self.g, self.s, _fp(sino), direction, kind, _dp(vol), _dp(filt))
        return vol.reshape(self.n_cond, self.N, self.N, self.N), filt.reshape(sino.shape)


def rodrigues(x, y, z, c, sn):
    """rodriguesRotation (ir.cu:597-621) in FP64: 3x3 rotation about axis
    (x, y, z) with cosine c and sine sn."""
    R = np.zeros(9)
    _lib.orc_rodrigues(float(x), float(y), float(z), float(c), float(sn), _dp(R))
    return R.reshape(3, 3)


def orth_md(vol, coef, prec=F64):
    """convertNormVector: md (3 channels) = mean(mu) * R e_z."""
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    coef = np.ascontiguousarray(coef, dtype=np.float32)
    md = np.zeros(vol.size, np.float32)
    _lib.orc_orth_md(prec, vol.size // 3, _fp(vol), _fp(coef), _fp(md))
    return md.reshape(vol.shape)


def orth_mean_filter(src, vol, w, dst=None, prec=F64):
    """meanFiltFiber over the interior voxels (reference layout); dst's boundary
    voxels are kept (default: a copy of src)."""
    src = np.ascontiguousarray(src, dtype=np.float32)
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    N = round((src.size // 2) ** (1.0 / 3.0))
    out = np.array(src if dst is None else dst, dtype=np.float32, copy=True)
    _lib.orc_orth_mean_filter(prec, N, _fp(src), _fp(vol), w, _fp(out))
    return out


def theta_f32(n, P):
    return _lib.orc_theta_f32(n, P)


def subset_orders(batch, epochs, per_epoch_engine):
    out = np.zeros(batch * epochs, dtype=np.int32)
    _lib.orc_subset_orders(batch, epochs, int(per_epoch_engine), out.ctypes.data_as(_i32))
    return out.reshape(epochs, batch)


def ratio_step(sino, proj, art=False, prec=F64):
    """projRatio / projSubtract (ir.cu:1616-1666) on arrays of pixels: returns
    (ratio or residual image, sum |y - p|)."""
    y = np.ascontiguousarray(sino, dtype=np.float32).reshape(-1)
    p = np.array(proj, dtype=np.float64).reshape(-1)
    assert y.size == p.size
    loss = _lib.orc_ratio_step(prec, y.size, _fp(y), _dp(p), int(art))
    return p.reshape(np.shape(proj)), float(loss)


def pca(mu, basis):
    """mu: (K, nvox) float32; returns md (3, nvox), evalue (3, nvox) in FP64."""
    mu = np.ascontiguousarray(mu, dtype=np.float32)
    K, nvox = mu.shape
    b = np.ascontiguousarray(basis, dtype=np.float32).reshape(-1)
    md = np.zeros(3 * nvox)
    ev = np.zeros(3 * nvox)
    _lib.orc_pca(_fp(mu), K, _fp(b), nvox, _dp(md), _dp(ev))
    return md.reshape(3, nvox), ev.reshape(3, nvox)


def md_angles(md):
    """calcAngleFromMD (pca.cu:67-83): md (3, n) float32 -> (phi, theta) (2, n)."""
    md = np.ascontiguousarray(md, dtype=np.float32)
    n = md.shape[1]
    out = np.zeros((2, n), np.float32)
    _lib.orc_md_angles(_fp(md), _fp(out), n)
    return out


def md_normalise(md, eps=0.65):
    """directionNormalization (pca.cu:85-106): returns the normalised copy."""
    out = np.array(md, dtype=np.float32, order="C")
    _lib.orc_md_normalise(_fp(out), out.shape[1], eps)
    return out


def eig3(A):
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(9)
    w = np.zeros(3)
    v = np.zeros(9)
    _lib.orc_eig3(_dp(A), _dp(w), _dp(v))
    return w, v.reshape(3, 3)


def phantom_ellipsoids(N, vs, params):
    p = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
    vol = np.zeros(N ** 3, dtype=np.float32)
    _lib.orc_phantom_ellipsoids(N, vs, _dp(p), len(p) // 16, _fp(vol))
    return vol.reshape(N, N, N)


def noise_poisson(p, I0, seed):
    p = np.ascontiguousarray(p, dtype=np.float32).reshape(-1)
    y = np.zeros_like(p)
    _lib.orc_noise_poisson(_fp(y), _fp(p), p.size, I0, seed)
    return y


def noise_gaussian(x, sigma, seed):
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
    y = np.zeros_like(x)
    _lib.orc_noise_gaussian(_fp(y), _fp(x), x.size, sigma, seed)
    return y


def phantom_cylinder_shell(N, vs, r_out, thick, value, vol=None):
    vol = np.zeros(N ** 3, np.float32) if vol is None else np.ascontiguousarray(vol, np.float32).reshape(-1)
    _lib.orc_phantom_cylinder_shell(N, vs, r_out, thick, value, _fp(vol))
    return vol.reshape(N, N, N)


def phantom_fibres(N, vs, params, basis):
    p = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
    b = np.ascontiguousarray(basis, dtype=np.float32).reshape(-1)
    K = len(b) // 3
    vol = np.zeros(K * N ** 3, np.float32)
    _lib.orc_phantom_fibres(N, vs, _dp(p), len(p) // 8, _fp(b), K, _fp(vol))
    return vol.reshape(K, N, N, N)


def num_threads():
    return _lib.orc_num_threads()


# ---- learned warm-start network (restatement of the product's U-Net) ----

def unet_num_params(C):
    return int(_lib.orc_unet_num_params(C))


def unet_forward(params, C, x, scale, emulate_bf16=True):
    """x: (D0, D1, D2) float32 (D2 fastest); returns y = x + s*net(x/s) (float32)."""
    p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
    x = np.ascontiguousarray(x, dtype=np.float32)
    d0, d1, d2 = x.shape
    y = np.zeros_like(x)
    _lib.orc_unet_forward(_fp(p), C, _fp(x), d0, d1, d2, scale, int(emulate_bf16), _fp(y))
    return y


def unet_loss_grad(params, C, x, t, scale, emulate_bf16=True):
    """Returns (loss, grads (FP64, same layout as params))."""
    p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    d0, d1, d2 = x.shape
    g = np.zeros(p.size, dtype=np.float64)
    loss = _lib.orc_unet_loss_grad(_fp(p), C, _fp(x), _fp(t), d0, d1, d2, scale, int(emulate_bf16), _dp(g))
    return float(loss), g


def conv3(x, W, b, relu=False):
    """x: (D0, D1, D2, Cin); W: (Cout, 27, Cin); returns (D0, D1, D2, Cout) FP64."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    W = np.ascontiguousarray(W, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    d0, d1, d2, cin = x.shape
    cout = W.shape[0]
    out = np.zeros((d0, d1, d2, cout))
    _lib.orc_conv3(_fp(x), cin, d0, d1, d2, _fp(W), _fp(b), cout, int(relu), _dp(out))
    return out


def grad_coef(n, P, diff, prec=F64):
    """sinogramGradientCoef's factor {cos theta, sin theta, 1}[diff] (ir.cu:585-594)."""
    return _lib.orc_grad_coef(prec, n, P, diff)


def poisson_matrix(n):
    """The n^2 x n^2 system of poissonSolveLDLT (poisson_cpu.cpp:77-90): identity
    rows on the slice border, the 5-point Laplacian (4, four -1) inside, row and
    column index x + n*y."""
    import scipy.sparse as sp

    rows, cols, vals = [], [], []
    for x in range(n):
        for y in range(n):
            r = x + n * y
            if x == 0 or x == n - 1 or y == 0 or y == n - 1:
                rows.append(r), cols.append(r), vals.append(1.0)
                continue
            for dx, dy, v in ((0, 0, 4.0), (1, 0, -1.0), (-1, 0, -1.0), (0, 1, -1.0), (0, -1, -1.0)):
                rows.append(r), cols.append(x + dx + n * (y + dy)), vals.append(v)
    return sp.csc_matrix((vals, (rows, cols)), shape=(n * n, n * n))


def poisson_rhs(src0, src1, n):
    """Right sides of poissonSolveLDLT (poisson_cpu.cpp:98-105), in FP32 as the
    reference: src0(x,y,z) - src0(x+1,y,z) + src1(x,y,z) - src1(x,y+1,z) for
    1 <= x, y <= n-2, zero on the border.  src0/src1: (n+1)^3 float32 [z][y][x];
    returns [z][y][x] float32 of side n."""
    s0 = np.asarray(src0, dtype=np.float32)
    s1 = np.asarray(src1, dtype=np.float32)
    b = np.zeros((n, n, n), np.float32)
    i = slice(1, n - 1)
    b[:, i, i] = ((s0[:n, 1:n - 1, 1:n - 1] - s0[:n, 1:n - 1, 2:n]) + s1[:n, 1:n - 1, 1:n - 1]) - s1[:n, 2:n, 1:n - 1]
    return b


def poisson_solve(src0, src1, n):
    """poissonSolveLDLT (poisson_cpu.cpp:15-112): per z slice, solve the system
    of poisson_matrix(n) for poisson_rhs by sparse LU — Eigen's SparseLU in FP32
    there, SciPy's SuperLU in FP64 here.  Returns [z][y][x] float64 of side n."""
    import scipy.sparse.linalg as spl

    lu = spl.splu(poisson_matrix(n))
    b = poisson_rhs(src0, src1, n).astype(np.float64)
    out = np.zeros((n, n, n))
    for z in range(n):
        out[z] = lu.solve(b[z].reshape(-1)).reshape(n, n)
    return out

