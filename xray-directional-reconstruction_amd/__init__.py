"""Python binding of libxr.so, the MI355X cone-beam reconstruction library.

ctypes over the C ABI declared in include/xr.h.  Device buffers are torch
tensors on the HIP device (torch is only used for memory and streams); host
buffers are numpy arrays.  The functions mirror the reference's host API
(reconstruct/include/reconstruct.cuh:14-73): ``ir_reconstruct``,
``xtt_reconstruct``, ``fdk_reconstruct``, plus the device-level operators the
solvers are built from.  There is no CPU fallback: if libxr.so is missing or
fails to load, importing this package raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxr.so")

XR_OK = 0
CONV_IR, CONV_FDK, CONV_GRAD = 0, 1, 2
CW, CCW = 0, 1
METHOD_XTT, METHOD_MLEM, METHOD_ART, METHOD_FDK = 0, 1, 2, 3
FILTER_REFERENCE, FILTER_RAMLAK, FILTER_HILBERT = 0, 1, 2
OUT_RAW, OUT_NORMALISED, OUT_RATIO, OUT_RESIDUAL = 0, 1, 2, 3
TIMING_FORWARD, TIMING_BACKWARD, TIMING_FILTER, TIMING_OTHER = 0, 1, 2, 3
BATCH_GRAPH, BATCH_AUTOTUNE = 1, 2

_c_float_p = ctypes.POINTER(ctypes.c_float)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int32_p = ctypes.POINTER(ctypes.c_int32)
_c_uint8_p = ctypes.POINTER(ctypes.c_uint8)


class Geometry(ctypes.Structure):
    """xr_geometry: the reference Geometry (geometry.h:8-25) with U and V split."""

    _fields_ = [("sdd", ctypes.c_float), ("sod", ctypes.c_float), ("vox_size", ctypes.c_float),
                ("det_size", ctypes.c_float), ("n_vox", ctypes.c_int32), ("n_det_u", ctypes.c_int32),
                ("n_det_v", ctypes.c_int32), ("n_proj", ctypes.c_int32)]


class ScanStruct(ctypes.Structure):
    _fields_ = [("n_cond", ctypes.c_int32), ("n_basis", ctypes.c_int32), ("cond_R", _c_float_p),
                ("cond_t", _c_float_p), ("cond_offset", _c_float_p), ("basis", _c_float_p),
                ("grating", _c_float_p)]


class Settings(ctypes.Structure):
    _fields_ = [("proj_path", ctypes.c_char * 512), ("vol_path", ctypes.c_char * 512),
                ("direc_path", ctypes.c_char * 512), ("block_size", ctypes.c_int32), ("n_basis", ctypes.c_int32),
                ("n_cond", ctypes.c_int32), ("sod", ctypes.c_float), ("sdd", ctypes.c_float),
                ("det_size", ctypes.c_float), ("n_proj", ctypes.c_int32), ("n_det_u", ctypes.c_int32),
                ("n_det_v", ctypes.c_int32), ("n_vox", ctypes.c_int32), ("load_index", ctypes.c_int32 * 64),
                ("cond_R", ctypes.c_float * (9 * 64)), ("cond_t", ctypes.c_float * (3 * 64)),
                ("cond_offset", ctypes.c_float * (3 * 64)), ("basis", ctypes.c_float * (3 * 16))]


class Stats(ctypes.Structure):
    _fields_ = [("losses", _c_double_p)]


class Tiles(ctypes.Structure):
    """xr_tiles: projector tile shape (zeros select the defaults)."""

    _fields_ = [("bwd_zpl", ctypes.c_int32), ("bwd_gx", ctypes.c_int32), ("fwd_rows", ctypes.c_int32),
                ("fwd_scan", ctypes.c_int32), ("fwd_split", ctypes.c_int32)]

    def as_tuple(self):
        return (self.bwd_zpl, self.bwd_gx, self.fwd_rows)


class XRError(RuntimeError):
    pass


def _load():
    # torch ships its own HIP runtime (libamdhip64).  Both it and libxr.so need
    # the same runtime instance in the process, so torch (when present) is
    # imported first and libxr.so binds to the runtime torch already loaded.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libxr.so not built (%s); run __graft_entry__.build() or "
            "python xray-directional-reconstruction_amd/build_lib.py" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    P, I, F, D, V = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double, None
    sig = {
        "xr_context_create": (I, [ctypes.POINTER(ctypes.c_void_p), I]),
        "xr_context_destroy": (V, [P]),
        "xr_context_set_stream": (I, [P, P]),
        "xr_context_synchronize": (I, [P]),
        "xr_memcpy": (I, [P, P, P, ctypes.c_size_t]),
        "xr_last_error": (ctypes.c_char_p, [P]),
        "xr_timing_enable": (I, [P, I]),
        "xr_timing_reset": (I, [P]),
        "xr_timing_read": (I, [P, I, _c_double_p, ctypes.POINTER(ctypes.c_int64)]),
        "xr_plan_create": (I, [P, ctypes.POINTER(Geometry), ctypes.POINTER(ScanStruct), I, I,
                               ctypes.POINTER(ctypes.c_void_p)]),
        "xr_plan_destroy": (V, [P]),
        "xr_default_vox_size": (F, [F, F, F]),
        "xr_volume_to_columns": (I, [P, P, P, I, I]),
        "xr_volume_from_columns": (I, [P, P, P, I, I]),
        "xr_project": (I, [P, P, P, P, P, _c_int32_p, I, _c_uint8_p, I, I, _c_double_p]),
        "xr_project_noalloc": (I, [P, P, P, P, P, _c_int32_p, I, _c_uint8_p, I, I, _c_double_p]),
        "xr_residual": (I, [P, P, P, P, P, _c_int32_p, I, I, _c_double_p, _c_double_p]),
        "xr_backproject": (I, [P, P, P, _c_int32_p, I, I, P, P]),
        "xr_backproject_update": (I, [P, P, P, _c_int32_p, I, I, I, F, I, P]),
        "xr_fdk_filter": (I, [P, P, I, P, P, _c_int32_p, I]),
        "xr_fdk_backproject": (I, [P, P, P, _c_int32_p, I, I, P]),
        "xr_fdk_taps": (I, [ctypes.POINTER(Geometry), I, _c_float_p]),
        "xr_volume_sqrt": (I, [P, P, ctypes.c_int64]),
        "xr_volume_downsample": (I, [P, P, I, I, P]),
        "xr_orth_normal": (I, [P, _c_float_p, P, ctypes.c_int64, P, P, I, P]),
        "xr_orth_reconstruct": (I, [P, ctypes.POINTER(Geometry), ctypes.POINTER(ScanStruct),
                                    ctypes.POINTER(_c_float_p), ctypes.POINTER(_c_float_p), ctypes.POINTER(_c_float_p),
                                    I, I, I, I, I, ctypes.c_float, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_char_p]),
        "xr_orth_update_estimate": (I, [P, _c_float_p, P, ctypes.c_int64, P, I]),
        "xr_orth_md": (I, [P, P, P, ctypes.c_int64, P]),
        "xr_orth_mean_filter": (I, [P, P, P, I, ctypes.c_float, P]),
        "xr_volume_upsample_add": (I, [P, P, P, I, I, P, ctypes.c_float]),
        "xr_pca": (I, [P, P, I, _c_float_p, ctypes.c_int64, ctypes.c_int64, P, P]),
        "xr_ir_reconstruct": (I, [P, ctypes.POINTER(Geometry), ctypes.POINTER(ScanStruct),
                                  ctypes.POINTER(_c_float_p), ctypes.POINTER(_c_float_p), I, I, I, I, F,
                                  ctypes.POINTER(Stats)]),
        "xr_xtt_reconstruct": (I, [P, ctypes.POINTER(Geometry), ctypes.POINTER(ScanStruct),
                                   ctypes.POINTER(_c_float_p), ctypes.POINTER(_c_float_p),
                                   ctypes.POINTER(_c_float_p), ctypes.POINTER(_c_float_p), I, I, I, I, F,
                                   ctypes.POINTER(Stats)]),
        "xr_fdk_reconstruct": (I, [P, ctypes.POINTER(Geometry), ctypes.POINTER(ScanStruct),
                                   ctypes.POINTER(_c_float_p), ctypes.POINTER(_c_float_p), I, I]),
        "xr_grad_reconstruct": (I, [P, ctypes.POINTER(Geometry), ctypes.POINTER(ScanStruct), _c_float_p,
                                    _c_float_p, ctypes.POINTER(_c_float_p), I, I, I, I, F, I,
                                    ctypes.POINTER(Stats)]),
        "xr_poisson_solve": (I, [P, I, P, P, P]),
        "xr_load_settings": (I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Settings)]),
        "xr_settings_geometry": (I, [ctypes.POINTER(Settings), ctypes.POINTER(Geometry),
                                     ctypes.POINTER(ScanStruct)]),
        "xr_phantom_ellipsoid_params": (I, [ctypes.c_uint64, I, D, D, D, D, D, D, _c_double_p]),
        "xr_phantom_rasterise": (I, [P, _c_double_p, I, I, F, P]),
        "xr_pca_host": (I, [P, _c_float_p, I, _c_float_p, ctypes.c_int64, _c_float_p, _c_float_p]),
        "xr_phantom_cylinder_shell": (I, [P, I, F, D, D, D, P]),
        "xr_phantom_fibre_params": (I, [ctypes.c_uint64, I, I, D, D, D, D, D, D, _c_double_p, _c_double_p]),
        "xr_phantom_fibres": (I, [P, _c_double_p, I, I, F, _c_float_p, I, P]),
        "xr_phantom_fibres_scalar": (I, [P, _c_double_p, I, I, F, F, P]),
        "xr_noise_poisson": (I, [P, P, P, ctypes.c_int64, D, ctypes.c_uint64]),
        "xr_noise_gaussian": (I, [P, P, P, ctypes.c_int64, D, ctypes.c_uint64]),
        "xr_sino_clamp": (I, [P, P, ctypes.c_int64]),
        "xr_md_angles": (I, [P, P, P, ctypes.c_int64]),
        "xr_md_normalise": (I, [P, P, ctypes.c_int64, F]),
        "xr_md_angles_host": (I, [P, _c_float_p, _c_float_p, ctypes.c_int64]),
        "xr_md_normalise_host": (I, [P, _c_float_p, ctypes.c_int64, F]),
        "xr_philox4x32": (None, [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                 ctypes.POINTER(ctypes.c_uint32)]),
        "xr_plan_set_tiles": (I, [P, ctypes.POINTER(Tiles)]),
        "xr_plan_set_orientation": (I, [P, P]),
        "xr_test_fail_allocations": (V, [I]),
        "xr_plan_get_tiles": (I, [P, ctypes.POINTER(Tiles)]),
        "xr_autotune": (I, [P, P, P, P, ctypes.POINTER(Tiles), _c_double_p]),
        "xr_batch_create": (I, [P, I, ctypes.POINTER(Geometry), ctypes.POINTER(ScanStruct), _c_int32_p, I, I, F,
                                I, ctypes.POINTER(ctypes.c_void_p)]),
        "xr_batch_destroy": (V, [P]),
        "xr_batch_buffers": (I, [P, I, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]),
        "xr_batch_plan": (P, [P, I]),
        "xr_batch_autotune": (I, [P]),
        "xr_batch_run": (I, [P, I, _c_double_p]),
        "xr_batch_restart": (I, [P]),
        "xr_batch_elapsed_ms": (I, [P, _c_double_p]),
        "xr_unet_create": (I, [P, I, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
        "xr_unet_destroy": (V, [P]),
        "xr_unet_num_params": (ctypes.c_int64, [P]),
        "xr_unet_get_params": (I, [P, _c_float_p]),
        "xr_unet_set_params": (I, [P, _c_float_p]),
        "xr_unet_get_grads": (I, [P, _c_float_p]),
        "xr_unet_forward": (I, [P, P, P, I, I, I, F]),
        "xr_unet_loss_grad": (I, [P, P, P, I, I, I, F, _c_double_p]),
        "xr_unet_adam_step": (I, [P, F, F, F, F]),
        "xr_unet_last_scale": (F, [P]),
        "xr_unet_set_fused_upsampling": (I, [P, I]),
        "xr_unet_save": (I, [P, ctypes.c_char_p]),
        "xr_unet_load": (I, [P, ctypes.c_char_p]),
        "xr_batch_reconstruct": (I, [P, I, ctypes.POINTER(Geometry), ctypes.POINTER(ScanStruct), _c_int32_p,
                                     ctypes.POINTER(_c_float_p), ctypes.POINTER(_c_float_p), I, I, I, F, I,
                                     _c_double_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = _load()
EXPORTED = tuple(sorted(n for n in dir(_lib) if n.startswith("xr_")))


def lib():
    return _lib


# ---------------------------------------------------------------- helpers --

def _fptr(a):
    return a.ctypes.data_as(_c_float_p)


def _dptr(x):
    """Device pointer of a torch tensor (or an int address)."""
    if x is None:
        return None
    if isinstance(x, int):
        return ctypes.c_void_p(x)
    if not x.is_cuda:
        raise XRError("expected a device tensor")
    if not x.is_contiguous():
        raise XRError("expected a contiguous tensor")
    return ctypes.c_void_p(x.data_ptr())


def default_vox_size(sdd, sod, det_size):
    """voxSize = sod * detSize / sdd (geometry.h:13)."""
    return float(_lib.xr_default_vox_size(sdd, sod, det_size))


def make_geometry(sdd, sod, det_size, n_vox, n_det, n_proj, vox_size=None, n_det_v=None):
    g = Geometry()
    g.sdd, g.sod, g.det_size = sdd, sod, det_size
    g.vox_size = default_vox_size(sdd, sod, det_size) if vox_size is None else vox_size
    g.n_vox, g.n_det_u = n_vox, n_det
    g.n_det_v = n_det if n_det_v is None else n_det_v
    g.n_proj = n_proj
    return g


class Scan:
    """Per-condition tables (rotation, translation, offset, grating) and basis vectors."""

    def __init__(self, cond_R=None, cond_t=None, cond_offset=None, basis=None, grating=None, n_cond=1):
        n_cond = len(cond_R) // 9 if cond_R is not None else n_cond
        eye = np.tile(np.eye(3, dtype=np.float32).reshape(-1), n_cond)
        self.cond_R = np.ascontiguousarray(eye if cond_R is None else cond_R, dtype=np.float32).reshape(-1)
        self.cond_t = np.ascontiguousarray(np.zeros(3 * n_cond) if cond_t is None else cond_t,
                                           dtype=np.float32).reshape(-1)
        self.cond_offset = np.ascontiguousarray(np.zeros(3 * n_cond) if cond_offset is None else cond_offset,
                                                dtype=np.float32).reshape(-1)
        self.basis = None if basis is None else np.ascontiguousarray(basis, dtype=np.float32).reshape(-1)
        self.grating = None if grating is None else np.ascontiguousarray(grating, dtype=np.float32).reshape(-1)
        self.n_cond = n_cond
        self.n_basis = 0 if self.basis is None else len(self.basis) // 3
        s = ScanStruct()
        s.n_cond, s.n_basis = self.n_cond, self.n_basis
        s.cond_R, s.cond_t, s.cond_offset = _fptr(self.cond_R), _fptr(self.cond_t), _fptr(self.cond_offset)
        s.basis = _fptr(self.basis) if self.basis is not None else None
        s.grating = _fptr(self.grating) if self.grating is not None else None
        self.struct = s


class Context:
    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = _lib.xr_context_create(ctypes.byref(h), device)
        if rc != XR_OK:
            raise XRError("xr_context_create failed (%d)" % rc)
        self.h = h

    def close(self):
        if self.h:
            _lib.xr_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != XR_OK:
            raise XRError("libxr error %d: %s" % (rc, _lib.xr_last_error(self.h).decode()))

    def set_stream(self, stream_ptr):
        self.check(_lib.xr_context_set_stream(self.h, ctypes.c_void_p(stream_ptr)))

    def synchronize(self):
        self.check(_lib.xr_context_synchronize(self.h))

    def copy(self, dst, src, nbytes=None):
        """Synchronous copy between device tensors / raw device addresses / numpy arrays."""
        def addr(x):
            if isinstance(x, int):
                return ctypes.c_void_p(x)
            if isinstance(x, np.ndarray):
                if not x.flags.c_contiguous:
                    raise XRError("expected a contiguous array")
                return ctypes.c_void_p(x.ctypes.data)
            return _dptr(x)

        if nbytes is None:
            ref = src if not isinstance(src, int) else dst
            nbytes = ref.nbytes if isinstance(ref, np.ndarray) else ref.numel() * ref.element_size()
        self.check(_lib.xr_memcpy(self.h, addr(dst), addr(src), nbytes))

    def timing(self, enable=True):
        self.check(_lib.xr_timing_enable(self.h, int(enable)))

    def timing_reset(self):
        self.check(_lib.xr_timing_reset(self.h))

    def timing_read(self, kind):
        t, n = ctypes.c_double(), ctypes.c_int64()
        self.check(_lib.xr_timing_read(self.h, kind, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    # ---- plans and device operators ----
    def plan(self, geom, scan, convention=CONV_IR, direction=CW):
        return Plan(self, geom, scan, convention, direction)

    def to_columns(self, ref, col, n_vox, channels=1):
        self.check(_lib.xr_volume_to_columns(self.h, _dptr(ref), _dptr(col), n_vox, channels))

    def from_columns(self, col, ref, n_vox, channels=1):
        self.check(_lib.xr_volume_from_columns(self.h, _dptr(col), _dptr(ref), n_vox, channels))

    def orth_normal(self, basis, vol, nvox, coef, loss, rot=0, est=None):
        """calcNormalVectorThreeDirec (rot) / ...WithEst (est): coef updated in place."""
        b = np.ascontiguousarray(basis, dtype=np.float32).reshape(-1)
        self.check(_lib.xr_orth_normal(self.h, _fptr(b), _dptr(vol), nvox, _dptr(coef), _dptr(loss), rot,
                                       _dptr(est)))

    def orth_update_estimate(self, basis, vol, nvox, est, iteration):
        b = np.ascontiguousarray(basis, dtype=np.float32).reshape(-1)
        self.check(_lib.xr_orth_update_estimate(self.h, _fptr(b), _dptr(vol), nvox, _dptr(est), iteration))

    def orth_md(self, vol, coef, nvox, md):
        self.check(_lib.xr_orth_md(self.h, _dptr(vol), _dptr(coef), nvox, _dptr(md)))

    def orth_mean_filter(self, coef_src, vol, n_vox, weight, coef_dst):
        self.check(_lib.xr_orth_mean_filter(self.h, _dptr(coef_src), _dptr(vol), n_vox, weight, _dptr(coef_dst)))

    def volume_downsample(self, fine, n_vox, factor, coarse):
        """coarse = mean of each factor^3 block of fine (column layout)."""
        self.check(_lib.xr_volume_downsample(self.h, _dptr(fine), n_vox, factor, _dptr(coarse)))

    def volume_upsample_add(self, coarse_a, coarse_b, n_coarse, factor, fine, floor=-3.0e38):
        """fine = max(fine + trilinear(coarse_a - coarse_b), floor) (column layout)."""
        self.check(_lib.xr_volume_upsample_add(self.h, _dptr(coarse_a), _dptr(coarse_b), n_coarse, factor,
                                               _dptr(fine), floor))

    def volume_sqrt(self, vol):
        self.check(_lib.xr_volume_sqrt(self.h, _dptr(vol), vol.numel()))

    def pca(self, mu, basis, nvox, chan_stride, md, evalue):
        b = np.ascontiguousarray(basis, dtype=np.float32).reshape(-1)
        self.check(_lib.xr_pca(self.h, _dptr(mu), len(b) // 3, _fptr(b), nvox, chan_stride, _dptr(md), _dptr(evalue)))

    def phantom_cylinder_shell(self, n_vox, vox_size, r_out, thickness, value, vol_ref):
        self.check(_lib.xr_phantom_cylinder_shell(self.h, n_vox, vox_size, r_out, thickness, value, _dptr(vol_ref)))

    def phantom_fibres(self, params, n_vox, vox_size, basis, vol_ref):
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(basis, dtype=np.float32).reshape(-1)
        self.check(_lib.xr_phantom_fibres(self.h, p.ctypes.data_as(_c_double_p), len(p) // 8, n_vox, vox_size,
                                          _fptr(b), len(b) // 3, _dptr(vol_ref)))

    def phantom_fibres_scalar(self, params, n_vox, vox_size, value, vol_ref):
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        self.check(_lib.xr_phantom_fibres_scalar(self.h, p.ctypes.data_as(_c_double_p), len(p) // 8, n_vox,
                                                 vox_size, value, _dptr(vol_ref)))

    def noise_poisson(self, y, p, I0, seed):
        self.check(_lib.xr_noise_poisson(self.h, _dptr(y), _dptr(p), p.numel(), I0, seed))

    def noise_gaussian(self, y, x, sigma, seed):
        self.check(_lib.xr_noise_gaussian(self.h, _dptr(y), _dptr(x), x.numel(), sigma, seed))

    def md_angles(self, md, angle, nvox):
        """calcAngleFromMD on device buffers: md 3 x nvox, angle 2 x nvox."""
        self.check(_lib.xr_md_angles(self.h, _dptr(md), _dptr(angle), nvox))

    def md_normalise(self, md, nvox, eps=0.65):
        """directionNormalization on a device buffer of 3 x nvox (in place)."""
        self.check(_lib.xr_md_normalise(self.h, _dptr(md), nvox, eps))

    def md_angles_host(self, md):
        md = np.ascontiguousarray(md, dtype=np.float32)
        out = np.zeros((2, md.shape[1]), np.float32)
        self.check(_lib.xr_md_angles_host(self.h, _fptr(md), _fptr(out), md.shape[1]))
        return out

    def md_normalise_host(self, md, eps=0.65):
        out = np.array(md, dtype=np.float32, order="C")
        self.check(_lib.xr_md_normalise_host(self.h, _fptr(out), out.shape[1], eps))
        return out

    def sino_clamp(self, sino):
        """main.cu:35: sinogram values below 0 become 1e-8 (in place)."""
        self.check(_lib.xr_sino_clamp(self.h, _dptr(sino), sino.numel()))

    def pca_host(self, mu, basis):
        mu = np.ascontiguousarray(mu, dtype=np.float32)
        K, nv = mu.shape
        b = np.ascontiguousarray(basis, dtype=np.float32).reshape(-1)
        md = np.zeros(3 * nv, np.float32)
        ev = np.zeros(3 * nv, np.float32)
        self.check(_lib.xr_pca_host(self.h, _fptr(mu), K, _fptr(b), nv, _fptr(md), _fptr(ev)))
        return md.reshape(3, nv), ev.reshape(3, nv)

    def phantom_rasterise(self, params, n_vox, vox_size, vol_ref):
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        self.check(_lib.xr_phantom_rasterise(self.h, p.ctypes.data_as(_c_double_p), len(p) // 16, n_vox,
                                             vox_size, _dptr(vol_ref)))

    # ---- host-buffer solvers (reference API) ----
    def ir_reconstruct(self, geom, scan, sino, vol, epochs, batch, direction=CW, method=METHOD_MLEM, lam=1e-2):
        """IR::reconstruct: sino/vol are lists (one per condition) of float32 numpy arrays; vol in/out."""
        sp, vp = _ptr_list(sino), _ptr_list(vol)
        losses = np.zeros(max(1, scan.n_cond * epochs), dtype=np.float64)
        st = Stats(losses.ctypes.data_as(_c_double_p))
        self.check(_lib.xr_ir_reconstruct(self.h, ctypes.byref(geom), ctypes.byref(scan.struct), sp, vp, epochs,
                                          batch, direction, method, lam, ctypes.byref(st)))
        return losses[: scan.n_cond * epochs]

    def xtt_reconstruct(self, geom, scan, sino, vol, md, epochs, batch, direction=CW, method=METHOD_MLEM,
                        lam=1e-2, evalue=None):
        sp, vp, mp = _ptr_list(sino), _ptr_list(vol), _ptr_list(md)
        ep = _ptr_list(evalue) if evalue is not None else None
        losses = np.zeros(max(1, epochs), dtype=np.float64)
        st = Stats(losses.ctypes.data_as(_c_double_p))
        self.check(_lib.xr_xtt_reconstruct(self.h, ctypes.byref(geom), ctypes.byref(scan.struct), sp, vp, mp, ep,
                                           epochs, batch, direction, method, lam, ctypes.byref(st)))
        return losses[:epochs]

    def orth_reconstruct(self, geom, scan, sino, vol, md, iter1, iter2, batch, direction=CW, method=METHOD_MLEM,
                         lam=1e-2, seed=0, volume_path=None, direction_path=None):
        """XTT::orthTwiceReconstruct: sino[c], vol[3] (in/out), md[3] (out) float32 numpy arrays."""
        enc = lambda p: None if p is None else p.encode()  # noqa: E731
        self.check(_lib.xr_orth_reconstruct(self.h, ctypes.byref(geom), ctypes.byref(scan.struct), _ptr_list(sino),
                                            _ptr_list(vol), _ptr_list(md), iter1, iter2, batch, direction, method, lam,
                                            seed, enc(volume_path), enc(direction_path)))

    def batch(self, geoms, batch=None, direction=CW, method=METHOD_MLEM, lam=1e-2, flags=BATCH_GRAPH, scans=None):
        return Batch(self, geoms, batch, direction, method, lam, flags, scans)

    def batch_reconstruct(self, geoms, sino, vol, epochs, batch=None, direction=CW, method=METHOD_MLEM, lam=1e-2,
                          flags=BATCH_GRAPH, scans=None):
        """IR::reconstruct for many problems at once: sino[i], vol[i] float32 numpy arrays (vol in/out).
        Returns the [epoch][problem] raw loss sums."""
        n = len(geoms)
        ga = (Geometry * n)(*geoms)
        sa = None if scans is None else (ScanStruct * n)(*[s.struct for s in scans])
        ba = None if batch is None else np.ascontiguousarray(batch, dtype=np.int32)
        losses = np.zeros(max(1, epochs * n), dtype=np.float64)
        self.check(_lib.xr_batch_reconstruct(self.h, n, ga, sa, None if ba is None else ba.ctypes.data_as(_c_int32_p),
                                             _ptr_list(sino), _ptr_list(vol), epochs, direction, method, lam, flags,
                                             losses.ctypes.data_as(_c_double_p)))
        return losses[: epochs * n].reshape(epochs, n)

    def fdk_reconstruct(self, geom, scan, sino, vol, direction=CW, filter=FILTER_REFERENCE):
        sp, vp = _ptr_list(sino), _ptr_list(vol)
        self.check(_lib.xr_fdk_reconstruct(self.h, ctypes.byref(geom), ctypes.byref(scan.struct), sp, vp, direction,
                                           filter))


    def grad_reconstruct(self, geom, scan, sino, vol, epochs, batch, direction=CW, method=METHOD_ART, lam=1e-2,
                         block_size=32, grads=None):
        """IR::gradReconstruct: sino one float32 U*V*P sinogram, vol the N^3
        float32 result (out); grads: None or 3*n_cond float32 arrays of
        (N+1)^3 that receive the gradient volumes (slot 3*cond + diff).
        Returns the [slot][epoch] raw loss sums."""
        for a in (sino, vol):
            if a.dtype != np.float32 or not a.flags.c_contiguous:
                raise XRError("expected contiguous float32 arrays")
        slots = 3 * scan.n_cond
        gp = None if grads is None else _ptr_list(grads)
        if grads is not None and len(grads) != slots:
            raise XRError("grads needs %d arrays" % slots)
        losses = np.zeros(max(1, slots * epochs), dtype=np.float64)
        st = Stats(losses.ctypes.data_as(_c_double_p))
        self.check(_lib.xr_grad_reconstruct(self.h, ctypes.byref(geom), ctypes.byref(scan.struct), _fptr(sino),
                                            _fptr(vol), gp, epochs, batch, direction, method, lam, block_size,
                                            ctypes.byref(st)))
        return losses[: slots * epochs].reshape(slots, epochs)

    def poisson_solve(self, n, gx, gy, vol):
        """poissonSolveLDLT on device tensors: gx, gy (n+1)^3, vol n^3 (reference layouts)."""
        self.check(_lib.xr_poisson_solve(self.h, n, _dptr(gx), _dptr(gy), _dptr(vol)))


def _ptr_list(arrs):
    for a in arrs:
        if a.dtype != np.float32 or not a.flags.c_contiguous:
            raise XRError("expected contiguous float32 arrays")
    return (_c_float_p * len(arrs))(*[_fptr(a) for a in arrs])


class Plan:
    def __init__(self, ctx, geom, scan, convention=CONV_IR, direction=CW):
        self.ctx, self.geom, self.scan = ctx, geom, scan
        h = ctypes.c_void_p()
        ctx.check(_lib.xr_plan_create(ctx.h, ctypes.byref(geom), ctypes.byref(scan.struct), convention, direction,
                                      ctypes.byref(h)))
        self.h = h

    def close(self):
        # objects of a context must go before it; once the context is gone
        # the handle is dropped (its device memory stays until the process
        # ends) rather than destroyed through a dangling context
        if self.h and self.ctx.h:
            _lib.xr_plan_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def codes(self, views, cond=0):
        return np.ascontiguousarray(np.asarray(views, dtype=np.int64) + cond * self.geom.n_proj, dtype=np.int32)

    def project(self, vol, proj, codes, out_mode=OUT_NORMALISED, sino=None, accumulate=None, xtt=False):
        c = np.ascontiguousarray(codes, dtype=np.int32)
        acc = None if accumulate is None else np.ascontiguousarray(accumulate, dtype=np.uint8)
        loss = ctypes.c_double(0.0)
        self.ctx.check(_lib.xr_project(self.ctx.h, self.h, _dptr(vol), _dptr(proj), _dptr(sino),
                                       c.ctypes.data_as(_c_int32_p), len(c),
                                       None if acc is None else acc.ctypes.data_as(_c_uint8_p), int(xtt), out_mode,
                                       ctypes.byref(loss)))
        return loss.value

    def residual(self, vol, resid, sino, codes, xtt=False):
        """(sum |y - p|, sum (y - p)^2) over the listed views; resid gets y - p."""
        c = np.ascontiguousarray(codes, dtype=np.int32)
        l1, l2 = ctypes.c_double(), ctypes.c_double()
        self.ctx.check(_lib.xr_residual(self.ctx.h, self.h, _dptr(vol), _dptr(resid), _dptr(sino),
                                        c.ctypes.data_as(_c_int32_p), len(c), int(xtt), ctypes.byref(l1),
                                        ctypes.byref(l2)))
        return l1.value, l2.value

    def project_noalloc(self, vol, proj, codes, out_mode=OUT_NORMALISED, sino=None, accumulate=None, xtt=False):
        """project() through xr_project_noalloc (no allocation; the plan's
        tables must exist, i.e. after a first project() call)."""
        c = np.ascontiguousarray(codes, dtype=np.int32)
        acc = None if accumulate is None else np.ascontiguousarray(accumulate, dtype=np.uint8)
        loss = ctypes.c_double(0.0)
        self.ctx.check(_lib.xr_project_noalloc(self.ctx.h, self.h, _dptr(vol), _dptr(proj), _dptr(sino),
                                               c.ctypes.data_as(_c_int32_p), len(c),
                                               None if acc is None else acc.ctypes.data_as(_c_uint8_p), int(xtt),
                                               out_mode, ctypes.byref(loss)))
        return loss.value

    def backproject(self, proj, codes, tmp, factor, xtt=False):
        c = np.ascontiguousarray(codes, dtype=np.int32)
        self.ctx.check(_lib.xr_backproject(self.ctx.h, self.h, _dptr(proj), c.ctypes.data_as(_c_int32_p), len(c),
                                           int(xtt), _dptr(tmp), _dptr(factor)))

    def backproject_update(self, proj, codes, vol, method=METHOD_MLEM, lam=0.0, subset_size=1, xtt=False):
        c = np.ascontiguousarray(codes, dtype=np.int32)
        self.ctx.check(_lib.xr_backproject_update(self.ctx.h, self.h, _dptr(proj), c.ctypes.data_as(_c_int32_p),
                                                  len(c), int(xtt), method, lam, subset_size, _dptr(vol)))

    def set_orientation(self, coef):
        """Orientation field of the XTT projections (forwardOrth / backwardOrth):
        a device tensor of 2*N^3 floats, phi then cos theta, column layout; None
        returns to the fixed basis.  The plan keeps a reference to the tensor."""
        self._orient = coef
        self.ctx.check(_lib.xr_plan_set_orientation(self.h, None if coef is None else _dptr(coef)))

    def set_tiles(self, bwd_zpl=0, bwd_gx=0, fwd_rows=0, fwd_scan=0, fwd_split=0):
        t = Tiles(bwd_zpl, bwd_gx, fwd_rows, fwd_scan, fwd_split)
        self.ctx.check(_lib.xr_plan_set_tiles(self.h, ctypes.byref(t)))

    def tiles(self):
        """Effective tile shape (zpl, gx, rows)."""
        t = Tiles()
        self.ctx.check(_lib.xr_plan_get_tiles(self.h, ctypes.byref(t)))
        return t.as_tuple()

    def autotune(self, vol, sino):
        t = Tiles()
        ms = ctypes.c_double()
        self.ctx.check(_lib.xr_autotune(self.ctx.h, self.h, _dptr(vol), _dptr(sino), ctypes.byref(t),
                                        ctypes.byref(ms)))
        return t.as_tuple(), ms.value

    def fdk_filter(self, sino, filt, codes, filter=FILTER_REFERENCE):
        c = np.ascontiguousarray(codes, dtype=np.int32)
        self.ctx.check(_lib.xr_fdk_filter(self.ctx.h, self.h, filter, _dptr(sino), _dptr(filt),
                                          c.ctypes.data_as(_c_int32_p), len(c)))

    def fdk_backproject(self, filt, codes, vol, overwrite=True):
        c = np.ascontiguousarray(codes, dtype=np.int32)
        self.ctx.check(_lib.xr_fdk_backproject(self.ctx.h, self.h, _dptr(filt), c.ctypes.data_as(_c_int32_p),
                                               len(c), int(overwrite), _dptr(vol)))


class _BorrowedPlan(Plan):
    """A plan owned by a batch (not destroyed from Python)."""

    def __init__(self, ctx, geom, handle):
        self.ctx, self.geom, self.scan = ctx, geom, None
        self.h = handle

    def close(self):
        self.h = None


class Batch:
    """Batched multi-problem IR solver (xr_batch_*): persistent launches over all problems,
    epoch loop captured as a HIP graph with BATCH_GRAPH."""

    def __init__(self, ctx, geoms, batch=None, direction=CW, method=METHOD_MLEM, lam=1e-2, flags=BATCH_GRAPH,
                 scans=None):
        self.ctx, self.geoms = ctx, list(geoms)
        n = len(self.geoms)
        ga = (Geometry * n)(*self.geoms)
        sa = None if scans is None else (ScanStruct * n)(*[s.struct for s in scans])
        self._scans = scans
        ba = None if batch is None else np.ascontiguousarray(batch, dtype=np.int32)
        h = ctypes.c_void_p()
        ctx.check(_lib.xr_batch_create(ctx.h, n, ga, sa, None if ba is None else ba.ctypes.data_as(_c_int32_p),
                                       direction, method, lam, flags, ctypes.byref(h)))
        self.h = h
        self.n = n

    def close(self):
        # objects of a context must go before it; once the context is gone
        # the handle is dropped (its device memory stays until the process
        # ends) rather than destroyed through a dangling context
        if self.h and self.ctx.h:
            _lib.xr_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def buffers(self, i):
        """(vol_cols, sino) device addresses of problem i."""
        v, s = ctypes.c_void_p(), ctypes.c_void_p()
        self.ctx.check(_lib.xr_batch_buffers(self.h, i, ctypes.byref(v), ctypes.byref(s)))
        return v.value, s.value

    def plan(self, i):
        return _BorrowedPlan(self.ctx, self.geoms[i], ctypes.c_void_p(_lib.xr_batch_plan(self.h, i)))

    def autotune(self):
        self.ctx.check(_lib.xr_batch_autotune(self.h))

    def run(self, epochs):
        losses = np.zeros(max(1, epochs * self.n), dtype=np.float64)
        self.ctx.check(_lib.xr_batch_run(self.h, epochs, losses.ctypes.data_as(_c_double_p)))
        return losses[: epochs * self.n].reshape(epochs, self.n)

    def restart(self):
        """Restart the subset order (as a new IR::reconstruct call)."""
        self.ctx.check(_lib.xr_batch_restart(self.h))

    def elapsed_ms(self):
        ms = ctypes.c_double()
        self.ctx.check(_lib.xr_batch_elapsed_ms(self.h, ctypes.byref(ms)))
        return ms.value


class UNet:
    """The learned warm-start network (xr_unet_*): y = x + s*net(x/s) on device
    volumes in the column layout (dims (D0, D1, D2), D2 fastest)."""

    def __init__(self, ctx, base_channels=24, seed=0):
        self.ctx, self.C = ctx, base_channels
        h = ctypes.c_void_p()
        ctx.check(_lib.xr_unet_create(ctx.h, base_channels, seed, ctypes.byref(h)))
        self.h = h
        self.n_params = int(_lib.xr_unet_num_params(h))

    def close(self):
        # objects of a context must go before it; once the context is gone
        # the handle is dropped (its device memory stays until the process
        # ends) rather than destroyed through a dangling context
        if self.h and self.ctx.h:
            _lib.xr_unet_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        self.ctx.check(_lib.xr_unet_get_params(self.h, _fptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise XRError("expected %d parameters" % self.n_params)
        self.ctx.check(_lib.xr_unet_set_params(self.h, _fptr(p)))

    def grads(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        self.ctx.check(_lib.xr_unet_get_grads(self.h, _fptr(out)))
        return out

    def forward(self, x, y, dims, scale=0.0):
        self.ctx.check(_lib.xr_unet_forward(self.h, _dptr(x), _dptr(y), dims[0], dims[1], dims[2], scale))

    def loss_grad(self, x, t, dims, scale=0.0):
        loss = ctypes.c_double()
        self.ctx.check(_lib.xr_unet_loss_grad(self.h, _dptr(x), _dptr(t), dims[0], dims[1], dims[2], scale,
                                              ctypes.byref(loss)))
        return loss.value

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.ctx.check(_lib.xr_unet_adam_step(self.h, lr, beta1, beta2, eps))

    def last_scale(self):
        return float(_lib.xr_unet_last_scale(self.h))

    def set_fused_upsampling(self, enable):
        """Inference with the decoder's upsampling fused into the next
        convolution's fetch (True, the default) or materialised (False)."""
        self.ctx.check(_lib.xr_unet_set_fused_upsampling(self.h, 1 if enable else 0))

    def save(self, path):
        self.ctx.check(_lib.xr_unet_save(self.h, path.encode()))

    def load(self, path):
        self.ctx.check(_lib.xr_unet_load(self.h, path.encode()))


def fdk_taps(geom, filter=FILTER_REFERENCE):
    out = np.zeros(geom.n_det_u, dtype=np.float32)
    rc = _lib.xr_fdk_taps(ctypes.byref(geom), filter, _fptr(out))
    if rc != XR_OK:
        raise XRError("xr_fdk_taps failed (%d)" % rc)
    return out


def load_settings(path, tag):
    """init_params(tag) equivalent: returns (Settings, Geometry, Scan)."""
    st = Settings()
    rc = _lib.xr_load_settings(path.encode(), tag.encode(), ctypes.byref(st))
    if rc != XR_OK:
        raise XRError("xr_load_settings(%s, %s) failed (%d)" % (path, tag, rc))
    g = Geometry()
    s = ScanStruct()
    _lib.xr_settings_geometry(ctypes.byref(st), ctypes.byref(g), ctypes.byref(s))
    scan = Scan(cond_R=np.ctypeslib.as_array(st.cond_R)[: 9 * st.n_cond].copy(),
                cond_t=np.ctypeslib.as_array(st.cond_t)[: 3 * st.n_cond].copy(),
                cond_offset=np.ctypeslib.as_array(st.cond_offset)[: 3 * st.n_cond].copy(),
                basis=np.ctypeslib.as_array(st.basis)[: 3 * st.n_basis].copy() if st.n_basis else None)
    return st, g, scan


def fibre_params(seed, count, n_axes=3, kappa=20.0, fov_radius=1.0, centre_frac=0.6, r_lo=0.4, r_hi=1.2,
                 half_length=None):
    """Fibre bundles: returns (params (count, 8), axes (n_axes, 3))."""
    out = np.zeros(8 * max(count, 1), dtype=np.float64)
    axes = np.zeros(3 * n_axes, dtype=np.float64)
    hl = fov_radius if half_length is None else half_length
    rc = _lib.xr_phantom_fibre_params(seed, count, n_axes, kappa, fov_radius, centre_frac, r_lo, r_hi, hl,
                                      out.ctypes.data_as(_c_double_p), axes.ctypes.data_as(_c_double_p))
    if rc != XR_OK:
        raise XRError("xr_phantom_fibre_params failed (%d)" % rc)
    return out[: 8 * count].reshape(count, 8), axes.reshape(n_axes, 3)


def philox4x32(counter, key):
    c = (ctypes.c_uint32 * 4)(*counter)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    _lib.xr_philox4x32(c, k, o)
    return list(o)


def ellipsoid_params(seed, count, fov_radius, centre_frac=0.6, axis_lo=0.1, axis_hi=0.35, value_lo=0.0,
                     value_hi=0.05):
    out = np.zeros(16 * max(count, 1), dtype=np.float64)
    rc = _lib.xr_phantom_ellipsoid_params(seed, count, fov_radius, centre_frac, axis_lo, axis_hi, value_lo, value_hi,
                                          out.ctypes.data_as(_c_double_p))
    if rc != XR_OK:
        raise XRError("xr_phantom_ellipsoid_params failed (%d)" % rc)
    return out[: 16 * count].reshape(count, 16)


def test_fail_allocations(after):
    """Test hook: device allocations of the library fail after `after` more
    (negative: off)."""
    _lib.xr_test_fail_allocations(int(after))
