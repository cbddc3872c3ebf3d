"""ctypes wrapper around the CPU oracle (oracle/_build/libct_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker. The product package never
imports this module.
"""
from __future__ import annotations

import ctypes
import os
import platform
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
REF_BIN = os.path.join(HERE, "_ref", "ref_filter")

RAMP, SHEPP_LOGAN, HAMMING, COSINE, HANN, BARTLETT = range(6)

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int)
_sz = ctypes.c_size_t


class Ellipsoid(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("cx", "cy", "cz", "ax", "ay", "az", "phi", "theta", "psi", "rho")]


def build() -> None:
    """Compile the oracle libraries (and oracle/_ref when the reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/QtCTReconstruction"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(BUILD, "libct_oracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.ct_o_deg2rad.restype = ctypes.c_float
        L.ct_o_deg2rad.argtypes = [ctypes.c_float]
        L.ct_o_next_pow2.restype = _sz
        L.ct_o_next_pow2.argtypes = [_sz]
        L.ct_o_linspace.argtypes = [ctypes.c_float, ctypes.c_float, _sz, _f32p]
        L.ct_o_forward_real.argtypes = [_f32p, _sz, _f32p]
        L.ct_o_inverse_real.argtypes = [_f32p, _sz, _f32p]
        L.ct_o_create_filter.argtypes = [_sz, ctypes.c_int, _f32p]
        L.ct_o_compute_sinogram.argtypes = [_f32p, _sz, _sz, _sz, _sz, _f32p, _f32p]
        L.ct_o_geometry.argtypes = [_sz, ctypes.POINTER(_sz), ctypes.POINTER(_sz), ctypes.POINTER(_sz)]
        L.ct_o_reconstruct_slice.argtypes = [_f32p, _sz, _sz, _f32p, _sz, ctypes.c_int, _f32p, _f32p]
        L.ct_o_reconstruct_volume.argtypes = [_f32p, _sz, _sz, _sz, _sz, ctypes.c_int, _i32p, _sz,
                                              _f32p, _f32p, ctypes.c_int]
        L.ct_o64_filter.argtypes = [_sz, ctypes.c_int, _f64p]
        L.ct_o64_kernel.argtypes = [_sz, ctypes.c_int, _f64p]
        L.ct_o64_compute_sinogram.argtypes = [_f32p, _sz, _sz, _sz, _sz, _f64p]
        L.ct_o64_reconstruct_slice.argtypes = [_f64p, _sz, _sz, _f64p, _sz, ctypes.c_int, _f64p, _f64p]
        L.ct_o64_reconstruct_volume.argtypes = [_f32p, _sz, _sz, _sz, _sz, ctypes.c_int, _i32p, _sz,
                                                _f64p, ctypes.c_int]
        L.ct_o_metrics.argtypes = [_f32p, _f32p, _sz, _f64p, _f32p]
        L.ct_ph_params.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(Ellipsoid)]
        L.ct_ph_rasterize.argtypes = [ctypes.POINTER(Ellipsoid), ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_int]
        _lib = L
    return _lib


def _p32(a):
    return a.ctypes.data_as(_f32p)


def _p64(a):
    return a.ctypes.data_as(_f64p)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed (rc={rc})")


# ---------------------------------------------------------------- Utils / FFT
def deg2rad(d: float) -> np.float32:
    return np.float32(lib().ct_o_deg2rad(float(d)))


def next_pow2(v: int) -> int:
    return int(lib().ct_o_next_pow2(v))


def linspace(a: float, b: float, n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    lib().ct_o_linspace(a, b, n, _p32(out))
    return out


def forward_real(x) -> np.ndarray:
    x = _c32(x)
    out = np.empty(2 * (x.size // 2 + 1), np.float32)
    _check(lib().ct_o_forward_real(_p32(x), x.size, _p32(out)), "forward_real")
    return out.view(np.complex64)


def inverse_real(spec) -> np.ndarray:
    spec = np.ascontiguousarray(spec, dtype=np.complex64)
    n2 = spec.size - 1
    out = np.empty(2 * n2, np.float32)
    _check(lib().ct_o_inverse_real(spec.view(np.float32).ctypes.data_as(_f32p), n2, _p32(out)),
           "inverse_real")
    return out


def create_filter(n: int, ftype: int) -> np.ndarray:
    out = np.empty(n // 2 + 1, np.float32)
    _check(lib().ct_o_create_filter(n, ftype, _p32(out)), "create_filter")
    return out


def geometry(bins: int):
    sq, pad, nfft = _sz(), _sz(), _sz()
    lib().ct_o_geometry(bins, ctypes.byref(sq), ctypes.byref(pad), ctypes.byref(nfft))
    return sq.value, pad.value, nfft.value


# ---------------------------------------------------------------- Serial path
def compute_sinogram(slice2d, na: int, bins: int):
    s = _c32(slice2d)
    h, w = s.shape
    sino = np.empty((na, bins), np.float32)
    mm = np.empty(2, np.float32)
    _check(lib().ct_o_compute_sinogram(_p32(s), w, h, na, bins, _p32(sino), _p32(mm)),
           "compute_sinogram")
    return sino, mm


def serial_angles(na: int) -> np.ndarray:
    step = np.float32(180.0) / np.float32(na)
    return (step * np.arange(na, dtype=np.float32)).astype(np.float32)


def reconstruct_slice(sino, angles_deg, output_size: int, ftype: int, want_filtered=False):
    sino = _c32(sino)
    na, bins = sino.shape
    ang = _c32(angles_deg)
    sq, _, _ = geometry(bins)
    filt = np.empty((na, sq), np.float32)
    rec = np.empty((output_size, output_size), np.float32)
    _check(lib().ct_o_reconstruct_slice(_p32(sino), na, bins, _p32(ang), output_size, ftype,
                                        _p32(filt), _p32(rec)), "reconstruct_slice")
    return (rec, filt) if want_filtered else rec


def reconstruct_volume(vol, na: int, ftype: int, z_list=None, threads: int = 0, want_sino=False):
    """Serial reconstructVolume for the listed slices; returns (out, sino|None).

    Slices not in z_list are left as NaN so they cannot be mistaken for checked ones.
    """
    vol = _c32(vol)
    d, h, w = vol.shape
    zl = np.arange(d, dtype=np.int32) if z_list is None else np.asarray(z_list, np.int32)
    out = np.full(vol.shape, np.nan, np.float32)
    sino = np.empty((zl.size, na, w), np.float32) if want_sino else None
    _check(lib().ct_o_reconstruct_volume(_p32(vol), w, h, d, na, ftype,
                                         zl.ctypes.data_as(_i32p), zl.size, _p32(out),
                                         _p32(sino) if want_sino else None, threads),
           "reconstruct_volume")
    return out, sino


# ---------------------------------------------------------------- FP64 twin
def filter64(n: int, ftype: int) -> np.ndarray:
    out = np.empty(n // 2 + 1, np.float64)
    _check(lib().ct_o64_filter(n, ftype, _p64(out)), "filter64")
    return out


def kernel64(n: int, ftype: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    _check(lib().ct_o64_kernel(n, ftype, _p64(out)), "kernel64")
    return out


def compute_sinogram64(slice2d, na: int, bins: int) -> np.ndarray:
    s = _c32(slice2d)
    h, w = s.shape
    out = np.empty((na, bins), np.float64)
    _check(lib().ct_o64_compute_sinogram(_p32(s), w, h, na, bins, _p64(out)), "sinogram64")
    return out


def reconstruct_slice64(sino, angles_deg, output_size: int, ftype: int, want_filtered=False):
    sino = np.ascontiguousarray(sino, np.float64)
    na, bins = sino.shape
    ang = np.ascontiguousarray(angles_deg, np.float64)
    sq, _, _ = geometry(bins)
    filt = np.empty((na, sq), np.float64)
    rec = np.empty((output_size, output_size), np.float64)
    _check(lib().ct_o64_reconstruct_slice(_p64(sino), na, bins, _p64(ang), output_size, ftype,
                                          _p64(filt), _p64(rec)), "reconstruct_slice64")
    return (rec, filt) if want_filtered else rec


def reconstruct_volume64(vol, na: int, ftype: int, z_list=None, threads: int = 0):
    vol = _c32(vol)
    d, h, w = vol.shape
    zl = np.arange(d, dtype=np.int32) if z_list is None else np.asarray(z_list, np.int32)
    out = np.full(vol.shape, np.nan, np.float64)
    _check(lib().ct_o64_reconstruct_volume(_p32(vol), w, h, d, na, ftype,
                                           zl.ctypes.data_as(_i32p), zl.size, _p64(out), threads),
           "reconstruct_volume64")
    return out


# ---------------------------------------------------------------- phantom
def phantom_params(seed: int, n: int):
    arr = (Ellipsoid * max(n, 1))()
    _check(lib().ct_ph_params(seed, n, arr), "phantom_params")
    return arr


def phantom(seed: int, n: int, nx: int, ny: int, nz: int, threads: int = 0) -> np.ndarray:
    e = phantom_params(seed, n)
    vol = np.empty((nz, ny, nx), np.float32)
    _check(lib().ct_ph_rasterize(e, n, nx, ny, nz, _p32(vol), threads), "phantom")
    return vol


def phantom_from(e, n: int, nx: int, ny: int, nz: int, threads: int = 0) -> np.ndarray:
    vol = np.empty((nz, ny, nx), np.float32)
    _check(lib().ct_ph_rasterize(e, n, nx, ny, nz, _p32(vol), threads), "phantom")
    return vol


# ---------------------------------------------------------------- OpenMP baseline
def _host_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " avx512f " in line + " "
    except OSError:
        pass
    return False


_omp = None


def ompbase():
    """The OpenMPBackend::reconstructVolume restatement (timed baseline only)."""
    global _omp
    if _omp is None:
        tag = "v4" if _host_has_avx512() else "v3"
        path = os.path.join(BUILD, f"libct_ompbase_{tag}.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.ct_omp_reconstruct_volume.argtypes = [_f32p, _sz, _sz, _sz, _sz, ctypes.c_int, _f32p,
                                                ctypes.c_int, _f64p]
        L.ct_omp_radon_forward.argtypes = [_f32p, _sz, _sz, _sz, _sz, _f32p]
        L.isa = tag
        _omp = L
    return _omp


def omp_reconstruct_volume(vol, na: int, ftype: int, threads: int):
    vol = _c32(vol)
    d, h, w = vol.shape
    out = np.empty_like(vol)
    ph = np.zeros(3, np.float64)
    _check(ompbase().ct_omp_reconstruct_volume(_p32(vol), w, h, d, na, ftype, _p32(out), threads,
                                               _p64(ph)), "omp_reconstruct_volume")
    return out, ph


def omp_radon_forward(slice2d, na: int, bins: int) -> np.ndarray:
    """The OpenMP baseline's Radon stage of one slice (RadonTransform.cpp:25-113):
    -sin convention, (int) truncation; [na][bins]."""
    s = _c32(slice2d)
    h, w = s.shape
    out = np.empty((na, bins), np.float32)
    _check(ompbase().ct_omp_radon_forward(_p32(s), w, h, na, bins, _p32(out)), "omp_radon_forward")
    return out


def metrics(orig, recon) -> dict:
    """CtReconstructionController.cpp:1199-1277 metrics, sequential (ct_o_metrics)."""
    o, r = _c32(orig).ravel(), _c32(recon).ravel()
    assert o.size == r.size
    d = np.empty(3, np.float64)
    f = np.empty(3, np.float32)
    _check(lib().ct_o_metrics(_p32(o), _p32(r), o.size, _p64(d), _p32(f)), "metrics")
    return {"mse": float(d[0]), "mae": float(d[1]), "psnr": float(d[2]),
            "max_abs": np.float32(f[0]), "vmin": np.float32(f[1]), "vmax": np.float32(f[2]),
            "count": o.size}


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def host_info() -> dict:
    return {"machine": platform.machine(), "cpus": os.cpu_count()}
