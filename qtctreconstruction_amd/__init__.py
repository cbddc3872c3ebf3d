"""MI355X-native FBP backend: Python host side over the C ABI of libct_hip.so.

Mirrors the reference's plugin interface (QtCTReconstruction/IReconstructionBackend.h:10-64,
CTTypes.h) with the same method names, argument meaning and error behaviour:
empty inputs give empty results (SerialBackend.cpp:64,148), failures raise.

The compute path is the HIP library only. If libct_hip.so is missing or no
device is visible, construction raises; nothing here falls back to the CPU.
"""
from __future__ import annotations

import ctypes
import enum
import os
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libct_hip.so")


class FilterType(enum.IntEnum):
    """ReconstructionParams::FilterType (CTTypes.h:151), same order."""
    Ramp = 0
    SheppLogan = 1
    Hamming = 2
    Cosine = 3
    Hann = 4
    Bartlett = 5


class FilterBackend(enum.IntEnum):
    """Row-filter backends (CT_FILTER_BACKEND_* in include/ct_hip.h)."""
    FFT_F32 = 0
    GEMM_BF16 = 1
    GEMM_SPLIT_BF16 = 2
    GEMM_FP8 = 3


class BackendError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[ct_hip {code}] {msg}")
        self.code = code


(CT_OK, CT_ERR_INVALID, CT_ERR_HIP, CT_ERR_NOMEM, CT_ERR_BUDGET, CT_ERR_UNSUPPORTED, CT_ERR_NODEVICE,
 CT_ERR_IO) = (0, -1, -2, -3, -4, -5, -6, -7)


class Ellipsoid(ctypes.Structure):
    """ct_ellipsoid (include/ct_hip.h)."""
    _fields_ = [(n, ctypes.c_float) for n in
                ("cx", "cy", "cz", "ax", "ay", "az", "phi", "theta", "psi", "rho")]


_f32p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


class FbpJob(ctypes.Structure):
    """ct_fbp_job (include/ct_hip.h): one volume of a batch, device pointers."""
    _fields_ = [("d_sinos", ctypes.c_void_p), ("d_out", ctypes.c_void_p),
                ("angles_deg", ctypes.POINTER(ctypes.c_float)), ("nz", ctypes.c_int),
                ("na", ctypes.c_int), ("bins", ctypes.c_int), ("filter", ctypes.c_int)]


class Metrics(ctypes.Structure):
    """ct_metrics (include/ct_hip.h): the controller's reconstruction metrics."""
    _fields_ = [("mse", ctypes.c_double), ("mae", ctypes.c_double), ("psnr", ctypes.c_double),
                ("max_abs", ctypes.c_float), ("vmin", ctypes.c_float), ("vmax", ctypes.c_float),
                ("count", ctypes.c_uint64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class SinoInfo(ctypes.Structure):
    """ct_sino_info (include/ct_hip.h): shape and sidecar fields of a sinogram file."""
    _fields_ = [("nz", ctypes.c_size_t), ("na", ctypes.c_size_t), ("bins", ctypes.c_size_t),
                ("detector_spacing_mm", ctypes.c_float), ("min_hu", ctypes.c_float),
                ("max_hu", ctypes.c_float), ("meta_found", ctypes.c_int),
                ("meta_minmax", ctypes.c_int), ("meta_angles", ctypes.c_int)]


CT_BATCH_GRAPH, CT_BATCH_AUTOTUNE = 1, 2
CT_STAGE_RADON, CT_STAGE_FBP = 0, 1
SERVE_FIFO, SERVE_SIZE_AGED = 0, 1
_lib = None


def load_library(path: str = LIB_PATH):
    """Load libct_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m qtctreconstruction_amd.build`")
    L = ctypes.CDLL(path)
    ctx = _vp
    sz = ctypes.c_size_t
    i = ctypes.c_int
    sig = {
        "ct_hip_device_count": (i, []),
        "ct_hip_create": (i, [i, sz, ctypes.POINTER(_vp)]),
        "ct_hip_destroy": (None, [ctx]),
        "ct_hip_last_error": (ctypes.c_char_p, [ctx]),
        "ct_hip_bytes_in_use": (sz, [ctx]),
        "ct_hip_clear_workspace": (i, [ctx]),
        "ct_hip_alloc": (i, [ctx, sz, ctypes.POINTER(_vp)]),
        "ct_hip_free": (i, [ctx, _vp]),
        "ct_hip_upload": (i, [ctx, _vp, _vp, sz]),
        "ct_hip_download": (i, [ctx, _vp, _vp, sz]),
        "ct_hip_synchronize": (i, [ctx]),
        "ct_phantom_params": (i, [ctypes.c_uint64, i, ctypes.POINTER(Ellipsoid)]),
        "ct_hip_phantom": (i, [ctx, ctypes.POINTER(Ellipsoid), i, i, i, i, _f32p]),
        "ct_hip_phantom_dev": (i, [ctx, ctypes.POINTER(Ellipsoid), i, i, i, i, _vp]),
        "ct_hip_compute_sinogram": (i, [ctx, _f32p, i, i, i, i, _f32p, _f32p]),
        "ct_hip_reconstruct_slice": (i, [ctx, _f32p, i, i, _f32p, i, i, _f32p]),
        "ct_hip_reconstruct_volume": (i, [ctx, _f32p, i, i, i, i, i, _f32p, _f32p]),
        "ct_hip_reconstruct_volume_dev": (i, [ctx, _vp, i, i, i, i, i, _vp, _vp]),
        "ct_hip_reconstruct_from_sinograms": (i, [ctx, _f32p, i, i, i, _f32p, i, _f32p]),
        "ct_hip_reconstruct_from_sinograms_dev": (i, [ctx, _vp, i, i, i, _f32p, i, _vp]),
        "ct_hip_reconstruct_from_sinograms_dev_noalloc": (i, [ctx, _vp, i, i, i, _f32p, i, _vp]),
        "ct_hip_extract_point_cloud": (i, [ctx, _f32p, i, i, i, _f32p, _f32p, _f32p, ctypes.c_float,
                                           _f32p, sz, ctypes.POINTER(sz)]),
        "ct_hip_extract_point_cloud_dev": (i, [ctx, _vp, i, i, i, _f32p, _f32p, _f32p,
                                               ctypes.c_float, _f32p, sz, ctypes.POINTER(sz)]),
        "ct_hip_design_filter": (i, [ctx, i, i, _f32p]),
        "ct_hip_set_filter_backend": (i, [ctx, i]),
        "ct_hip_fbp_batch_dev": (i, [ctx, ctypes.POINTER(FbpJob), i, i]),
        "ct_hip_bp_config_count": (i, []),
        "ct_hip_bp_config": (i, [i, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]),
        "ct_hip_autotune": (i, [ctx, i, i, ctypes.POINTER(i), ctypes.POINTER(ctypes.c_double), i]),
        "ct_hip_set_bp_config": (i, [ctx, i]),
        "ct_hip_server_create": (i, [i, sz, i, i, ctypes.POINTER(_vp)]),
        "ct_hip_server_destroy": (None, [_vp]),
        "ct_hip_server_set_flags": (i, [_vp, i]),
        "ct_hip_server_set_filter_backend": (i, [_vp, i]),
        "ct_hip_server_set_policy": (i, [_vp, i, ctypes.c_double, ctypes.c_double]),
        "ct_hip_server_submit": (i, [_vp, ctypes.POINTER(FbpJob), ctypes.POINTER(ctypes.c_uint64)]),
        "ct_hip_server_wait": (i, [_vp, ctypes.c_double, ctypes.POINTER(ctypes.c_uint64),
                                   ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_double)]),
        "ct_hip_server_download": (i, [_vp, ctypes.c_uint64, _f32p, sz]),
        "ct_hip_server_release": (i, [_vp, ctypes.c_uint64]),
        "ct_hip_server_stats": (i, [_vp, ctypes.POINTER(ctypes.c_double), i]),
        "ct_hip_server_last_error": (ctypes.c_char_p, [_vp]),
        "ct_hip_radon_dev": (i, [ctx, _vp, i, i, i, i, i, _vp, _f32p]),
        "ct_hip_noisy_log_dev": (i, [ctx, _vp, sz, ctypes.c_float, ctypes.c_float, ctypes.c_uint64]),
        "ct_hip_filter_sinograms": (i, [ctx, _f32p, i, i, i, i, _f32p]),
        "ct_hip_last_timings": (i, [ctx, ctypes.POINTER(ctypes.c_double)]),
        "ct_hip_check_shape": (i, [i, i, i, i, i]),
        "ct_hip_last_kernel_timings": (i, [ctx, ctypes.POINTER(ctypes.c_double), i]),
        "ct_hip_last_radon_timings": (i, [ctx, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_int)]),
        "ct_hip_volume_metrics": (i, [ctx, _f32p, _f32p, sz, ctypes.POINTER(Metrics)]),
        "ct_hip_volume_metrics_dev": (i, [ctx, _vp, _vp, sz, ctypes.POINTER(Metrics)]),
        "ct_hip_save_volume_npy": (i, [_f32p, sz, sz, sz, ctypes.c_char_p]),
        "ct_hip_load_volume_npy": (i, [ctypes.c_char_p, i, _f32p, sz, ctypes.POINTER(sz)]),
        "ct_hip_save_point_cloud_npy": (i, [_f32p, sz, ctypes.c_char_p]),
        "ct_hip_load_point_cloud_npy": (i, [ctypes.c_char_p, _f32p, sz, ctypes.POINTER(sz)]),
        "ct_hip_export_sinograms": (i, [ctx, _f32p, i, i, i, i, ctypes.c_char_p]),
        "ct_hip_load_sinograms": (i, [ctypes.c_char_p, _f32p, sz, _f32p, ctypes.POINTER(SinoInfo)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols() -> List[str]:
    """Names this wrapper binds (the test suite checks them against include/ct_hip.h)."""
    return [
        "ct_hip_device_count", "ct_hip_create", "ct_hip_destroy", "ct_hip_last_error",
        "ct_hip_bytes_in_use", "ct_hip_clear_workspace", "ct_hip_alloc", "ct_hip_free",
        "ct_hip_upload", "ct_hip_download", "ct_hip_synchronize", "ct_phantom_params",
        "ct_hip_phantom", "ct_hip_phantom_dev", "ct_hip_compute_sinogram",
        "ct_hip_reconstruct_slice", "ct_hip_reconstruct_volume", "ct_hip_reconstruct_volume_dev",
        "ct_hip_reconstruct_from_sinograms", "ct_hip_reconstruct_from_sinograms_dev",
        "ct_hip_reconstruct_from_sinograms_dev_noalloc",
        "ct_hip_extract_point_cloud", "ct_hip_extract_point_cloud_dev",
        "ct_hip_design_filter", "ct_hip_set_filter_backend", "ct_hip_filter_sinograms",
        "ct_hip_fbp_batch_dev", "ct_hip_bp_config_count", "ct_hip_bp_config", "ct_hip_autotune",
        "ct_hip_set_bp_config", "ct_hip_server_create", "ct_hip_server_destroy",
        "ct_hip_server_set_flags", "ct_hip_server_set_filter_backend",
        "ct_hip_server_set_policy", "ct_hip_server_submit",
        "ct_hip_server_wait",
        "ct_hip_server_download", "ct_hip_server_release", "ct_hip_server_stats",
        "ct_hip_server_last_error", "ct_hip_radon_dev", "ct_hip_noisy_log_dev",
        "ct_hip_last_timings", "ct_hip_check_shape", "ct_hip_last_kernel_timings",
        "ct_hip_last_radon_timings", "ct_hip_volume_metrics", "ct_hip_volume_metrics_dev",
        "ct_hip_save_volume_npy", "ct_hip_load_volume_npy", "ct_hip_save_point_cloud_npy",
        "ct_hip_load_point_cloud_npy", "ct_hip_export_sinograms", "ct_hip_load_sinograms",
    ]


def _p(a: np.ndarray):
    return a.ctypes.data_as(_f32p)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def check_shape(stage: int, w: int, h: int, bins: int, filter_backend: int = 0) -> int:
    """ct_hip_check_shape: CT_OK, CT_ERR_INVALID or CT_ERR_UNSUPPORTED for a
    stage's shape (host only, no device needed)."""
    return int(load_library().ct_hip_check_shape(int(stage), int(w), int(h), int(bins),
                                                 int(filter_backend)))


def phantom_params(seed: int, n: int):
    """Seeded ellipsoid parameters (host only, no device needed)."""
    L = load_library()
    arr = (Ellipsoid * max(n, 1))()
    rc = L.ct_phantom_params(ctypes.c_uint64(seed), n, arr)
    if rc != CT_OK:
        raise BackendError(rc, "ct_phantom_params failed")
    return arr


@dataclass
class Sinogram:
    """Sinogram (CTTypes.h:131-137): data [num_angles][detector_bins]."""
    data: np.ndarray
    angles_deg: np.ndarray
    detector_spacing_mm: float = 1.0
    original_min_hu: float = 0.0
    original_max_hu: float = 0.0


@dataclass
class ReconstructionParams:
    """ReconstructionParams (CTTypes.h:150-156)."""
    filter: FilterType = FilterType.SheppLogan
    num_angles: int = 180
    zero_padding: bool = True


class DeviceBuffer:
    """Device allocation owned by a context (freed with it or by .free())."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _vp()
        ctx._call("ct_hip_alloc", ctx.handle, self.nbytes, ctypes.byref(p))
        self.ptr = p.value

    def upload(self, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("upload larger than buffer")
        self.ctx._call("ct_hip_upload", self.ctx.handle, self.ptr, a.ctypes.data, a.nbytes)
        return self

    def download(self, shape, dtype=np.float32) -> np.ndarray:
        out = np.empty(shape, dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("download larger than buffer")
        self.ctx._call("ct_hip_download", self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes)
        return out

    def free(self):
        if self.ptr:
            self.ctx._call("ct_hip_free", self.ctx.handle, self.ptr)
            self.ptr = None


class Context:
    """One device context (ct_hip_ctx)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self.lib = load_library()
        h = _vp()
        rc = self.lib.ct_hip_create(device, hbm_budget_bytes, ctypes.byref(h))
        if rc != CT_OK:
            raise BackendError(rc, "ct_hip_create failed (no HIP device?)")
        self.handle = h.value

    def _call(self, name: str, *args):
        rc = getattr(self.lib, name)(*args)
        if rc != CT_OK:
            msg = self.lib.ct_hip_last_error(self.handle) or b""
            raise BackendError(rc, f"{name}: {msg.decode(errors='replace')}")
        return rc

    def close(self):
        if getattr(self, "handle", None):
            self.lib.ct_hip_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bytes_in_use(self) -> int:
        return int(self.lib.ct_hip_bytes_in_use(self.handle))

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def synchronize(self):
        self._call("ct_hip_synchronize", self.handle)

    def timings(self) -> np.ndarray:
        t = (ctypes.c_double * 4)()
        self._call("ct_hip_last_timings", self.handle, t)
        return np.array(t[:], np.float64)

    def kernel_timings(self) -> dict:
        """Per-kernel milliseconds of the last call (summed over its slabs); when
        the two Radon walks run as one launch (rows and columns of at most
        516 px), walk_image holds both and walk_transpose is zero."""
        t = (ctypes.c_double * 6)()
        self._call("ct_hip_last_kernel_timings", self.handle, t, 6)
        keys = ["minmax", "normalise", "walk_image", "walk_transpose", "row_filter",
                "backprojection"]
        return dict(zip(keys, [float(x) for x in t]))

    def radon_timings(self):
        """(ms of the row walk over the image, ms of the row walk over the
        transposed image, launches per walk) of the last Radon stage; both
        walks' time in the first and zero in the second when they run as one
        launch."""
        t = (ctypes.c_double * 2)()
        n = ctypes.c_int(0)
        self._call("ct_hip_last_radon_timings", self.handle, t, ctypes.byref(n))
        return float(t[0]), float(t[1]), int(n.value)

    def design_filter(self, n: int, ftype: int) -> np.ndarray:
        out = np.empty(n // 2 + 1, np.float32)
        self._call("ct_hip_design_filter", self.handle, n, int(ftype), _p(out))
        return out

    def bp_configs(self):
        """[(tile, angles_per_stage, slices_per_thread)] of the instantiated
        backprojection configurations."""
        out = []
        for k in range(self.lib.ct_hip_bp_config_count()):
            t, a, z = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            self.lib.ct_hip_bp_config(k, ctypes.byref(t), ctypes.byref(a), ctypes.byref(z))
            out.append((t.value, a.value, z.value))
        return out

    def set_bp_config(self, cfg: int):
        self._call("ct_hip_set_bp_config", self.handle, int(cfg))

    def autotune(self, N: int, na: int):
        """(best config index, per-config milliseconds)."""
        n = self.lib.ct_hip_bp_config_count()
        best = ctypes.c_int()
        ms = (ctypes.c_double * n)()
        self._call("ct_hip_autotune", self.handle, N, na, ctypes.byref(best), ms, n)
        return best.value, list(ms)

    def tuned_config(self, N: int, na: int) -> int:
        """The autotuner's tile configuration for N x N slices and na angles
        (tuning on first use); reconstructions use it unless set_bp_config
        fixed one."""
        best = ctypes.c_int()
        self._call("ct_hip_autotune", self.handle, N, na, ctypes.byref(best), None, 0)
        return best.value

    def fbp_batch(self, jobs, flags: int = CT_BATCH_GRAPH | CT_BATCH_AUTOTUNE):
        """jobs: list of (d_sinos DeviceBuffer, d_out DeviceBuffer, angles, nz, na, bins, filter)."""
        keep = []
        arr = (FbpJob * len(jobs))()
        for k, (ds, do, ang, nz, na, bins, f) in enumerate(jobs):
            a = _f32(ang)
            keep.append(a)
            arr[k] = FbpJob(ds.ptr, do.ptr, _p(a), nz, na, bins, int(f))
        self._call("ct_hip_fbp_batch_dev", self.handle, arr, len(jobs), int(flags))
        return self.timings()

    def set_filter_backend(self, backend: int):
        self._call("ct_hip_set_filter_backend", self.handle, int(backend))

    def filter_sinograms(self, sinos, ftype: int) -> np.ndarray:
        """The row-filter step alone: (nz,na,bins) -> (nz,na,sq)."""
        s = _f32(sinos)
        nz, na, bins = s.shape
        sq = int(np.ceil(np.sqrt(2.0) * bins))
        out = np.empty((nz, na, sq), np.float32)
        self._call("ct_hip_filter_sinograms", self.handle, _p(s), nz, na, bins, int(ftype), _p(out))
        return out

    def reconstruct_volume_dev(self, d_vol: DeviceBuffer, w: int, h: int, d: int, na: int,
                               ftype: int, d_out: DeviceBuffer,
                               d_sino: Optional[DeviceBuffer] = None):
        """reconstructVolume over device memory (ct_hip_reconstruct_volume_dev);
        d_sino (optional, d*na*w floats) receives the sinograms."""
        self._call("ct_hip_reconstruct_volume_dev", self.handle, d_vol.ptr, w, h, d, na, int(ftype),
                   d_out.ptr, d_sino.ptr if d_sino is not None else None)

    def reconstruct_from_sinograms_dev(self, d_sinos: DeviceBuffer, nz: int, na: int, bins: int,
                                       angles, ftype: int, d_out: DeviceBuffer):
        """reconstructVolumeFromSinograms over device memory: (nz,na,bins) ->
        (nz,bins,bins), no denormalisation (ct_hip_reconstruct_from_sinograms_dev)."""
        a = _f32(angles)
        self._call("ct_hip_reconstruct_from_sinograms_dev", self.handle, d_sinos.ptr, nz, na, bins,
                   _p(a), int(ftype), d_out.ptr)

    def radon_dev(self, d_vol: DeviceBuffer, w: int, h: int, d: int, na: int, bins: int,
                  d_sino: DeviceBuffer) -> np.ndarray:
        mm = np.empty(2 * d, np.float32)
        self._call("ct_hip_radon_dev", self.handle, d_vol.ptr, w, h, d, na, bins, d_sino.ptr, _p(mm))
        return mm.reshape(d, 2)

    def noisy_log_dev(self, d_sino: DeviceBuffer, n: int, I0: float, kappa: float, seed: int):
        self._call("ct_hip_noisy_log_dev", self.handle, d_sino.ptr, n, float(I0), float(kappa),
                   ctypes.c_uint64(seed))

    def phantom(self, seed: int, n: int, nx: int, ny: int, nz: int) -> np.ndarray:
        e = phantom_params(seed, n)
        out = np.empty((nz, ny, nx), np.float32)
        self._call("ct_hip_phantom", self.handle, e, n, nx, ny, nz, _p(out))
        return out

    def phantom_dev(self, seed: int, n: int, nx: int, ny: int, nz: int,
                    buf: Optional[DeviceBuffer] = None) -> DeviceBuffer:
        e = phantom_params(seed, n)
        buf = buf or self.alloc(4 * nx * ny * nz)
        self._call("ct_hip_phantom_dev", self.handle, e, n, nx, ny, nz, buf.ptr)
        return buf

    def volume_metrics_dev(self, d_orig: DeviceBuffer, d_recon: DeviceBuffer, n: int) -> dict:
        """Reconstruction metrics of n voxels held in device memory (ct_metrics as a dict)."""
        m = Metrics()
        self._call("ct_hip_volume_metrics_dev", self.handle, d_orig.ptr, d_recon.ptr, int(n),
                   ctypes.byref(m))
        return m.as_dict()


class Server:
    """Serving form (ct_hip_server_*): continuous batching over device-resident sinograms."""

    def __init__(self, pool_bytes: int, max_batch: int = 16, slots: int = 2, device: int = 0):
        self.lib = load_library()
        h = _vp()
        rc = self.lib.ct_hip_server_create(device, int(pool_bytes), max_batch, slots, ctypes.byref(h))
        if rc != CT_OK:
            raise BackendError(rc, "ct_hip_server_create failed")
        self.handle = h.value
        self._angles = {}

    def _err(self, rc, what):
        msg = self.lib.ct_hip_server_last_error(self.handle) or b""
        raise BackendError(rc, f"{what}: {msg.decode(errors='replace')}")

    def set_flags(self, flags: int):
        self.lib.ct_hip_server_set_flags(self.handle, int(flags))

    def set_filter_backend(self, backend: int):
        """Row filter backend (FilterBackend) of the batches formed from now on;
        a Toeplitz-GEMM backend filters each job with one MFMA launch."""
        rc = self.lib.ct_hip_server_set_filter_backend(self.handle, int(backend))
        if rc != CT_OK:
            raise BackendError(rc, f"set_filter_backend: unknown backend {int(backend)}")

    def set_policy(self, policy: int = 1, batch_cap_ms: float = 5.0, age_weight: float = 0.25):
        """Admission order: SERVE_FIFO, or SERVE_SIZE_AGED (default) = ascending
        estimated device time minus age_weight x time waited; batch_cap_ms caps
        a batch's estimated device time (0 = no cap)."""
        rc = self.lib.ct_hip_server_set_policy(self.handle, int(policy), float(batch_cap_ms),
                                               float(age_weight))
        if rc != CT_OK:
            self._err(rc, "set_policy")

    def submit(self, d_sinos: "DeviceBuffer", angles, nz: int, na: int, bins: int, ftype: int) -> int:
        a = _f32(angles)
        job = FbpJob(d_sinos.ptr, None, _p(a), nz, na, bins, int(ftype))
        t = ctypes.c_uint64()
        rc = self.lib.ct_hip_server_submit(self.handle, ctypes.byref(job), ctypes.byref(t))
        if rc != CT_OK:
            self._err(rc, "submit")
        return t.value

    def wait(self, timeout_ms: float = 1000.0):
        """(ticket, latency_ms) of the next finished job, or None on timeout."""
        t = ctypes.c_uint64()
        p = _vp()
        lat = ctypes.c_double()
        rc = self.lib.ct_hip_server_wait(self.handle, timeout_ms, ctypes.byref(t), ctypes.byref(p),
                                         ctypes.byref(lat))
        if rc == 1:
            return None
        if rc != CT_OK:
            self._err(rc, f"job {t.value}")
        return t.value, lat.value

    def download(self, ticket: int, shape) -> np.ndarray:
        return self.download_into(ticket, np.empty(shape, np.float32))

    def download_into(self, ticket: int, out: np.ndarray) -> np.ndarray:
        """A finished job's output into a caller's float32 C-contiguous array."""
        if out.dtype != np.float32 or not out.flags.c_contiguous:
            raise ValueError("download_into: float32 C-contiguous array required")
        rc = self.lib.ct_hip_server_download(self.handle, ticket, _p(out), out.nbytes)
        if rc != CT_OK:
            self._err(rc, "download")
        return out

    def release(self, ticket: int):
        rc = self.lib.ct_hip_server_release(self.handle, ticket)
        if rc != CT_OK:
            self._err(rc, "release")

    def stats(self) -> dict:
        v = (ctypes.c_double * 7)()
        self.lib.ct_hip_server_stats(self.handle, v, 7)
        keys = ["batches", "completed", "max_inflight_jobs", "overlapped_launches",
                "pool_used", "pool_size", "span_ms"]
        return dict(zip(keys, list(v)))

    def close(self):
        if getattr(self, "handle", None):
            self.lib.ct_hip_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipBackend:
    """IReconstructionBackend (IReconstructionBackend.h:10-55) on the HIP library."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self.ctx = Context(device, hbm_budget_bytes)
        self._sino_ms = 0.0
        self._recon_ms = 0.0

    def name(self) -> str:
        return "HIP (MI355X)"

    def isAvailable(self) -> bool:
        return load_library().ct_hip_device_count() > 0

    def computeSinogram(self, slice2d, num_angles: int, detector_bins: int,
                        use_parallel: bool = True) -> Sinogram:
        s = _f32(slice2d)
        if s.size == 0 or num_angles == 0 or detector_bins == 0:
            return Sinogram(np.zeros((0, 0), np.float32), np.zeros(0, np.float32))
        h, w = s.shape
        out = np.empty((num_angles, detector_bins), np.float32)
        mm = np.empty(2, np.float32)
        self.ctx._call("ct_hip_compute_sinogram", self.ctx.handle, _p(s), w, h, num_angles,
                       detector_bins, _p(out), _p(mm))
        step = np.float32(180.0) / np.float32(num_angles)
        ang = (step * np.arange(num_angles, dtype=np.float32)).astype(np.float32)
        return Sinogram(out, ang, 1.0, float(mm[0]), float(mm[1]))

    def reconstructSlice(self, sino: Sinogram, output_size: int,
                         params: ReconstructionParams) -> np.ndarray:
        d = _f32(sino.data)
        if d.size == 0 or output_size == 0:
            return np.zeros((0, 0), np.float32)
        na, bins = d.shape
        ang = _f32(sino.angles_deg)
        out = np.empty((output_size, output_size), np.float32)
        self.ctx._call("ct_hip_reconstruct_slice", self.ctx.handle, _p(d), na, bins, _p(ang),
                       output_size, int(params.filter), _p(out))
        return out

    def reconstructVolume(self, vol, params: ReconstructionParams,
                          onSliceDone: Optional[Callable[[int, np.ndarray], None]] = None,
                          want_sinograms: bool = False):
        v = _f32(vol)
        if v.size == 0:
            return np.zeros((0, 0, 0), np.float32)
        d, h, w = v.shape
        out = np.empty_like(v)
        sino = np.empty((d, params.num_angles, w), np.float32) if want_sinograms else None
        self.ctx._call("ct_hip_reconstruct_volume", self.ctx.handle, _p(v), w, h, d,
                       params.num_angles, int(params.filter), _p(out),
                       _p(sino) if sino is not None else None)
        t = self.ctx.timings()
        self._sino_ms, self._recon_ms = float(t[0]), float(t[1] + t[2])
        if onSliceDone:
            for z in range(d):
                onSliceDone(z, out[z])
        return (out, sino) if want_sinograms else out

    def reconstructVolumeFromSinograms(self, sinos, angles_deg, params: ReconstructionParams,
                                       onSliceDone=None) -> np.ndarray:
        s = _f32(sinos)
        if s.size == 0:
            return np.zeros((0, 0, 0), np.float32)
        nz, na, bins = s.shape
        ang = _f32(angles_deg)
        out = np.empty((nz, bins, bins), np.float32)
        self.ctx._call("ct_hip_reconstruct_from_sinograms", self.ctx.handle, _p(s), nz, na, bins,
                       _p(ang), int(params.filter), _p(out))
        t = self.ctx.timings()
        self._sino_ms, self._recon_ms = 0.0, float(t[1] + t[2])
        if onSliceDone:
            for z in range(nz):
                onSliceDone(z, out[z])
        return out

    def extractPointCloud(self, vol, threshold: float, x_coords=None, y_coords=None,
                          z_coords=None) -> np.ndarray:
        """Points (x, y, z, hu) of voxels with hu > threshold, in z,y,x order."""
        v = _f32(vol)
        if v.size == 0:
            return np.zeros((0, 4), np.float32)
        d, h, w = v.shape
        cs = [None if c is None else _f32(c) for c in (x_coords, y_coords, z_coords)]
        cp = [None if c is None else _p(c) for c in cs]
        n = ctypes.c_size_t(0)
        self.ctx._call("ct_hip_extract_point_cloud", self.ctx.handle, _p(v), w, h, d, *cp,
                       float(threshold), None, 0, ctypes.byref(n))
        out = np.empty((n.value, 4), np.float32)
        if n.value:
            self.ctx._call("ct_hip_extract_point_cloud", self.ctx.handle, _p(v), w, h, d, *cp,
                           float(threshold), _p(out), n.value, ctypes.byref(n))
        return out

    def volumeMetrics(self, original, reconstruction) -> dict:
        """The controller's metrics of a reconstruction against its original
        (CtReconstructionController.cpp:1199-1277): mse, mae, psnr, max_abs,
        vmin, vmax, count. Both volumes must hold the same number of voxels."""
        o, r = _f32(original), _f32(reconstruction)
        if o.size != r.size:
            raise ValueError("volumeMetrics: volumes differ in size")
        m = Metrics()
        self.ctx._call("ct_hip_volume_metrics", self.ctx.handle, _p(o), _p(r), o.size,
                       ctypes.byref(m))
        return m.as_dict()

    def lastSinogramTimeMs(self) -> float:
        return self._sino_ms

    def lastReconstructionTimeMs(self) -> float:
        return self._recon_ms


def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


class FileIO:
    """FileIO (FileIO.h:9-21): NPY volumes and point clouds, in the
    reference's byte format. Loaders return None where the reference returns
    false; savers return the reference's bool. Slice PNG/BMP export belongs to
    the UI and is not provided."""

    @staticmethod
    def saveVolumeNPY(volume, path) -> bool:
        v = _f32(volume)
        if v.ndim != 3 or v.size == 0:
            return False
        d, h, w = v.shape
        return load_library().ct_hip_save_volume_npy(_p(v), w, h, d, _path(path)) == CT_OK

    @staticmethod
    def loadVolumeNPY(path, input_is_xyz_layout: bool = False) -> Optional[np.ndarray]:
        """Volume (depth, height, width) as the reference builds it (FileIO.cpp:84-145).
        Its axis coordinates are linspace(-100, 100, n) per axis (volume_coords)."""
        L = load_library()
        dims = (ctypes.c_size_t * 3)()
        if L.ct_hip_load_volume_npy(_path(path), int(input_is_xyz_layout), None, 0, dims) != CT_OK:
            return None
        w, h, d = dims[0], dims[1], dims[2]
        out = np.empty((d, h, w), np.float32)
        if L.ct_hip_load_volume_npy(_path(path), int(input_is_xyz_layout), _p(out), out.size,
                                    dims) != CT_OK:
            return None
        return out

    @staticmethod
    def savePointCloudNPY(cloud, path) -> bool:
        c = _f32(cloud).reshape(-1, 4) if np.size(cloud) else np.zeros((0, 4), np.float32)
        if c.shape[0] == 0:
            return False
        return load_library().ct_hip_save_point_cloud_npy(_p(c), c.shape[0], _path(path)) == CT_OK

    @staticmethod
    def loadPointCloudNPY(path) -> Optional[np.ndarray]:
        L = load_library()
        n = ctypes.c_size_t(0)
        if L.ct_hip_load_point_cloud_npy(_path(path), None, 0, ctypes.byref(n)) != CT_OK:
            return None
        out = np.empty((n.value, 4), np.float32)
        if L.ct_hip_load_point_cloud_npy(_path(path), _p(out), n.value, ctypes.byref(n)) != CT_OK:
            return None
        return out


def volume_coords(n: int) -> np.ndarray:
    """Axis coordinates a loaded volume carries: linspace(-100, 100, n) in
    float (FileIO.cpp:141-143; Utils linspace: start + i*step, step = (b-a)/(n-1))."""
    if n <= 0:
        return np.zeros(0, np.float32)
    if n == 1:
        return np.array([-100.0], np.float32)
    step = np.float32(200.0) / np.float32(n - 1)
    return (np.float32(-100.0) + np.arange(n, dtype=np.float32) * step).astype(np.float32)


@dataclass
class LoadedSinograms:
    """What CtReconstructionController::loadSinograms keeps
    (CtReconstructionController.cpp:1600-1607)."""
    sinograms: np.ndarray          # (nz, na, bins)
    angles_deg: np.ndarray         # na floats
    detector_spacing_mm: float
    min_hu: float
    max_hu: float
    meta_found: bool
    meta_minmax: bool
    meta_angles: bool


def exportAllSinograms(backend: "HipBackend", volume, num_angles: int, npy_path) -> bool:
    """CtReconstructionController::exportAllSinograms
    (CtReconstructionController.cpp:1353-1482) for a volume in memory: the
    sinograms of every slice (bins = width) computed on the device, written as
    NPY (depth, num_angles, width) plus the .meta.json sidecar. Raises
    BackendError on a device failure; returns False when the files cannot be
    written, as the reference does."""
    v = _f32(volume)
    if v.ndim != 3 or v.size == 0 or num_angles <= 0:
        return False
    d, h, w = v.shape
    rc = backend.ctx.lib.ct_hip_export_sinograms(backend.ctx.handle, _p(v), w, h, d,
                                                 int(num_angles), _path(npy_path))
    if rc == CT_ERR_IO:
        return False
    if rc != CT_OK:
        msg = backend.ctx.lib.ct_hip_last_error(backend.ctx.handle) or b""
        raise BackendError(rc, f"ct_hip_export_sinograms: {msg.decode(errors='replace')}")
    return True


def loadSinograms(npy_path) -> Optional[LoadedSinograms]:
    """CtReconstructionController::loadSinograms
    (CtReconstructionController.cpp:1494-1616) without the file dialog: the
    sinogram stack and its sidecar fields with the reader's fallbacks. None
    where the reference returns false. Host only; needs no device."""
    L = load_library()
    info = SinoInfo()
    if L.ct_hip_load_sinograms(_path(npy_path), None, 0, None, ctypes.byref(info)) != CT_OK:
        return None
    s = np.empty((info.nz, info.na, info.bins), np.float32)
    ang = np.empty(info.na, np.float32)
    if L.ct_hip_load_sinograms(_path(npy_path), _p(s), s.size, _p(ang), ctypes.byref(info)) != CT_OK:
        return None
    return LoadedSinograms(s, ang, float(info.detector_spacing_mm), float(info.min_hu),
                           float(info.max_hu), bool(info.meta_found), bool(info.meta_minmax),
                           bool(info.meta_angles))
