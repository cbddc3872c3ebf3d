"""Reconstruction metrics and the on-disk formats (NPY volumes, point clouds,
sinogram stacks with their JSON sidecar).

CPU part: the metrics oracle (CtReconstructionController.cpp:1199-1277) is
pinned by numpy FP64, closed forms and the special cases; the file formats
(FileIO.cpp:56-206, CtReconstructionController.cpp:1353-1616) are host code
of the library and are checked against numpy's own NPY reader/writer, a
literal header and the sidecar fallback rules.

GPU part: the device metrics pass against the oracle, and the sinogram export
against the device Radon and the consumer path.
"""
import json
import math
import os

import numpy as np
import pytest

import qtctreconstruction_amd as ct

FLT_MAX = float(np.finfo(np.float32).max)


# ------------------------------------------------------------ metrics oracle
def _np_metrics(o, r):
    """Independent FP64 statement of the same figures with numpy's pairwise sums."""
    o = np.asarray(o, np.float32).ravel()
    r = np.asarray(r, np.float32).ravel()
    d = (r - o).astype(np.float32)  # float subtraction, as the controller
    n = d.size
    d64 = d.astype(np.float64)
    mse = float(np.sum(d64 * d64) / n)
    mae = float(np.sum(np.abs(d64)) / n)
    mx = np.float32(np.max(np.abs(d)))
    lo, hi = np.float32(o.min()), np.float32(o.max())
    rng = float(np.float32(hi - lo))
    psnr = math.inf if (mse <= 0 or rng <= 0) else 10 * math.log10(rng * rng / mse)
    return dict(mse=mse, mae=mae, psnr=psnr, max_abs=np.float32(1.0) if mx < 1e-6 else mx,
                vmin=lo, vmax=hi)


def test_oracle_metrics_vs_numpy(O):
    rng = np.random.default_rng(7)
    o = rng.normal(0, 300, 200_003).astype(np.float32)
    r = (o + rng.normal(0, 5, o.size)).astype(np.float32)
    a, b = O.metrics(o, r), _np_metrics(o, r)
    for k in ("mse", "mae", "psnr"):
        assert a[k] == pytest.approx(b[k], rel=1e-12), k
    for k in ("max_abs", "vmin", "vmax"):
        assert a[k] == b[k], k


def test_oracle_metrics_closed_forms(O):
    o = np.arange(1000, dtype=np.float32)
    m = O.metrics(o, o + np.float32(0.25))  # every diff exactly 0.25
    assert m["mse"] == 0.0625 and m["mae"] == 0.25 and m["max_abs"] == np.float32(0.25)
    assert m["psnr"] == pytest.approx(10 * math.log10(999.0 ** 2 / 0.0625), rel=1e-15)
    # identical volumes: PSNR infinite, max difference reported as 1
    m = O.metrics(o, o)
    assert m["mse"] == 0 and m["mae"] == 0 and math.isinf(m["psnr"]) and m["max_abs"] == 1.0
    # constant original: range 0 -> PSNR infinite however large the error
    m = O.metrics(np.full(10, 3.0, np.float32), np.zeros(10, np.float32))
    assert m["mse"] == 9.0 and math.isinf(m["psnr"])
    # differences below 1e-6 are reported as max difference 1
    m = O.metrics(np.ones(4, np.float32), np.ones(4, np.float32) + np.float32(2**-23))
    assert m["max_abs"] == 1.0 and m["mse"] > 0


def test_oracle_metrics_float_subtraction_and_range(O):
    # the difference is rounded to float before it is squared: 1 - 1e8 -> -1e8
    m = O.metrics(np.array([1e8], np.float32), np.array([1.0], np.float32))
    assert m["mse"] == 1e16 and m["mae"] == 1e8
    # the range is a float difference: 1e8 - (-1) rounds to 1e8
    o = np.array([-1.0, 1e8], np.float32)
    m = O.metrics(o, o + np.float32(4.0))
    assert m["psnr"] == pytest.approx(10 * math.log10(1e16 / m["mse"]), rel=1e-13)
    assert m["psnr"] != pytest.approx(10 * math.log10(100000001.0 ** 2 / m["mse"]), rel=1e-13)


def test_oracle_metrics_nan_and_empty(O):
    o = np.array([0.0, 1.0, 2.0], np.float32)
    r = np.array([0.5, np.nan, 2.0], np.float32)
    m = O.metrics(o, r)
    assert math.isnan(m["mse"]) and math.isnan(m["mae"])
    assert m["max_abs"] == np.float32(0.5)  # std::max keeps the running value over NaN
    m = O.metrics(np.zeros(0, np.float32), np.zeros(0, np.float32))
    assert m["mse"] == 0 and m["mae"] == 0 and math.isinf(m["psnr"]) and m["max_abs"] == 1.0
    assert m["vmin"] == np.float32(FLT_MAX) and m["vmax"] == np.float32(-FLT_MAX)


# --------------------------------------------------------------- NPY volumes
def test_save_volume_npy_bytes(tmp_path):
    v = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    p = tmp_path / "v.npy"
    assert ct.FileIO.saveVolumeNPY(v, p)
    raw = p.read_bytes()
    head = b"{'descr': '<f4', 'fortran_order': False, 'shape': (2, 3, 4), }"
    # 62 characters + 7 spaces + '\n' = 70: the smallest length >= 63 with
    # 10 + length a multiple of 16
    assert raw[:8] == b"\x93NUMPY\x01\x00"
    assert raw[8:10] == (70).to_bytes(2, "little")
    assert raw[10:80] == head + b" " * 7 + b"\n"
    assert raw[80:] == v.tobytes()
    np.testing.assert_array_equal(np.load(p), v)  # numpy's reader agrees
    assert not ct.FileIO.saveVolumeNPY(np.zeros((0, 3, 4), np.float32), tmp_path / "e.npy")
    assert not ct.FileIO.saveVolumeNPY(v, tmp_path / "no_such_dir" / "x.npy")


def test_load_volume_npy_numpy_files(tmp_path):
    rng = np.random.default_rng(3)
    v = rng.normal(size=(5, 7, 9)).astype(np.float32)
    p = tmp_path / "a.npy"
    np.save(p, v)  # numpy pads its header to 64 bytes
    np.testing.assert_array_equal(ct.FileIO.loadVolumeNPY(p), v)
    # version 2.0 header (uint32 length)
    p2 = tmp_path / "b.npy"
    with open(p2, "wb") as f:
        np.lib.format.write_array(f, v, version=(2, 0))
    np.testing.assert_array_equal(ct.FileIO.loadVolumeNPY(p2), v)
    # (x, y, z) files are transposed into (z, y, x)
    got = ct.FileIO.loadVolumeNPY(p, input_is_xyz_layout=True)
    assert got.shape == (9, 7, 5)
    np.testing.assert_array_equal(got, v.transpose(2, 1, 0))
    # our own writer round-trips
    p3 = tmp_path / "c.npy"
    assert ct.FileIO.saveVolumeNPY(v, p3)
    np.testing.assert_array_equal(ct.FileIO.loadVolumeNPY(p3), v)


def test_load_volume_npy_rejects(tmp_path):
    v = np.ones((2, 2, 2), np.float32)
    np.save(tmp_path / "f8.npy", v.astype(np.float64))
    assert ct.FileIO.loadVolumeNPY(tmp_path / "f8.npy") is None
    np.save(tmp_path / "2d.npy", np.ones((3, 4), np.float32))
    assert ct.FileIO.loadVolumeNPY(tmp_path / "2d.npy") is None
    np.save(tmp_path / "zero.npy", np.ones((0, 2, 2), np.float32))
    assert ct.FileIO.loadVolumeNPY(tmp_path / "zero.npy") is None
    np.save(tmp_path / "ok.npy", v)
    raw = (tmp_path / "ok.npy").read_bytes()
    (tmp_path / "short.npy").write_bytes(raw[:-4])
    assert ct.FileIO.loadVolumeNPY(tmp_path / "short.npy") is None
    (tmp_path / "magic.npy").write_bytes(b"\x93NUMPZ" + raw[6:])
    assert ct.FileIO.loadVolumeNPY(tmp_path / "magic.npy") is None
    assert ct.FileIO.loadVolumeNPY(tmp_path / "missing.npy") is None


def test_point_cloud_npy(tmp_path):
    rng = np.random.default_rng(5)
    pc = rng.normal(size=(11, 4)).astype(np.float32)
    p = tmp_path / "pc.npy"
    assert ct.FileIO.savePointCloudNPY(pc, p)
    raw = p.read_bytes()
    # 60 characters + 9 spaces + '\n' = 70 (10 + 70 = 5 * 16)
    assert raw[8:10] == (70).to_bytes(2, "little")
    assert raw[10:80] == b"{'descr': '<f4', 'fortran_order': False, 'shape': (11, 4), }" + b" " * 9 + b"\n"
    assert raw[80:] == pc.tobytes()
    np.testing.assert_array_equal(np.load(p), pc)
    np.testing.assert_array_equal(ct.FileIO.loadPointCloudNPY(p), pc)
    assert not ct.FileIO.savePointCloudNPY(np.zeros((0, 4), np.float32), tmp_path / "e.npy")
    np.save(tmp_path / "three.npy", np.ones((5, 3), np.float32))
    assert ct.FileIO.loadPointCloudNPY(tmp_path / "three.npy") is None


def test_volume_coords_match_reference_linspace(O):
    for n in (1, 2, 7, 128, 513):
        np.testing.assert_array_equal(ct.volume_coords(n), O.linspace(-100.0, 100.0, n))


# ----------------------------------------------------------- sinogram files
def _write_stack(tmp_path, name="s.npy", shape=(3, 8, 16), seed=11):
    s = np.random.default_rng(seed).normal(size=shape).astype(np.float32)
    p = tmp_path / name
    with open(p, "wb") as f:  # np.save(path) would append ".npy" to "X.NPY"
        np.save(f, s)
    return p, s


def _uniform(na):
    step = np.float32(180.0) / np.float32(na)
    return (step * np.arange(na, dtype=np.float32)).astype(np.float32)


def _sidecar(p):
    q = str(p)
    return q[:-4] + ".meta.json" if q.lower().endswith(".npy") else q + ".meta.json"


def test_load_sinograms_without_sidecar(tmp_path):
    p, s = _write_stack(tmp_path)
    L = ct.loadSinograms(p)
    np.testing.assert_array_equal(L.sinograms, s)
    assert not L.meta_found and not L.meta_minmax and not L.meta_angles
    np.testing.assert_array_equal(L.angles_deg, _uniform(8))
    assert L.detector_spacing_mm == 1.0
    assert L.min_hu == float(s.min()) and L.max_hu == float(s.max())
    assert ct.loadSinograms(tmp_path / "none.npy") is None


def test_load_sinograms_with_sidecar_and_fallbacks(tmp_path):
    p, s = _write_stack(tmp_path)
    ang = [0.1 * a + 1e-7 for a in range(8)]
    meta = {"version": 1, "depth": 3, "num_angles": 8, "detector_bins": 16,
            "detector_spacing_mm": 0.5, "original_min_hu": -1000.25, "original_max_hu": 2.5e3,
            "angles_deg": ang, "extra": {"nested": [1, {"x": "é\\"}], "e": -1.5e-3}}
    with open(_sidecar(p), "w") as f:
        json.dump(meta, f, indent=2)
    L = ct.loadSinograms(p)
    assert L.meta_found and L.meta_minmax and L.meta_angles
    np.testing.assert_array_equal(L.angles_deg, np.array(ang, np.float64).astype(np.float32))
    assert L.detector_spacing_mm == 0.5 and L.min_hu == -1000.25 and L.max_hu == 2500.0
    # wrong angle count -> uniform angles; one min/max key -> data extrema;
    # a spacing that is not a number -> 1
    meta.update(angles_deg=ang[:-1], detector_spacing_mm="wide")
    del meta["original_max_hu"]
    with open(_sidecar(p), "w") as f:
        json.dump(meta, f)
    L = ct.loadSinograms(p)
    assert L.meta_found and not L.meta_angles and not L.meta_minmax
    np.testing.assert_array_equal(L.angles_deg, _uniform(8))
    assert L.detector_spacing_mm == 1.0
    assert L.min_hu == float(s.min()) and L.max_hu == float(s.max())
    # both keys present but not numbers -> 0 each (QJsonValue::toDouble(0.0))
    meta.update(original_min_hu=None, original_max_hu="x")
    with open(_sidecar(p), "w") as f:
        json.dump(meta, f)
    L = ct.loadSinograms(p)
    assert L.meta_minmax and L.min_hu == 0.0 and L.max_hu == 0.0
    # unparsable sidecar -> every default
    with open(_sidecar(p), "w") as f:
        f.write('{"angles_deg": [1, 2,], }')
    L = ct.loadSinograms(p)
    assert not L.meta_found and not L.meta_angles and L.min_hu == float(s.min())


def test_sidecar_name_for_upper_case_extension(tmp_path):
    p, s = _write_stack(tmp_path, name="U.NPY", shape=(1, 2, 3))
    with open(tmp_path / "U.meta.json", "w") as f:
        json.dump({"detector_spacing_mm": 2.0}, f)
    L = ct.loadSinograms(p)
    assert L.meta_found and L.detector_spacing_mm == 2.0


# ------------------------------------------------------------------- GPU part
@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 3, 4, 5, 1023, 4097, (1 << 20) + 3, 1 << 24])
def test_device_metrics_vs_oracle(O, backend, n):
    rng = np.random.default_rng(n)
    o = rng.normal(0, 400, n).astype(np.float32)
    r = (o + rng.normal(0, 3, n)).astype(np.float32)
    got = backend.volumeMetrics(o, r)
    ref = O.metrics(o, r)
    # double sums of non-negative terms in another order: the sequential
    # oracle's relative error is at most (n-1) u (u = 2^-53), the device's
    # tree sum's far less, so 2 n u bounds their difference
    tol = max(1e-13, 2.0 * n * 2.0 ** -53)
    for k in ("mse", "mae", "psnr"):
        assert got[k] == pytest.approx(ref[k], rel=tol), (k, n)
    for k in ("max_abs", "vmin", "vmax"):
        assert np.float32(got[k]) == ref[k], (k, n)
    assert got["count"] == n
    again = backend.volumeMetrics(o, r)
    assert again == got  # fixed reduction order: same bits on every call


@pytest.mark.gpu
def test_device_metrics_special_cases(O, backend):
    o = np.arange(1000, dtype=np.float32)
    got = backend.volumeMetrics(o, o)
    assert got["mse"] == 0 and math.isinf(got["psnr"]) and got["max_abs"] == 1.0
    got = backend.volumeMetrics(o, o + np.float32(0.25))
    assert got["mse"] == 0.0625 and got["mae"] == 0.25
    r = o.copy()
    r[17] = np.nan
    got = backend.volumeMetrics(o, r)
    assert math.isnan(got["mse"]) and got["max_abs"] == 1.0
    got = backend.volumeMetrics(np.zeros(0, np.float32), np.zeros(0, np.float32))
    ref = O.metrics(np.zeros(0, np.float32), np.zeros(0, np.float32))
    assert got["count"] == 0 and math.isinf(got["psnr"]) and got["max_abs"] == 1.0
    assert np.float32(got["vmin"]) == ref["vmin"] and np.float32(got["vmax"]) == ref["vmax"]


@pytest.mark.gpu
def test_device_metrics_of_a_reconstruction(O, backend):
    """The producer's use: phantom against its reconstruction, device-resident."""
    ctx = backend.ctx
    N, d, na = 128, 9, 180
    vol = ctx.phantom(21, 10, N, N, d)
    rec = backend.reconstructVolume(vol, ct.ReconstructionParams(ct.FilterType.Ramp, na))
    dv = ctx.alloc(vol.nbytes).upload(vol)
    dr = ctx.alloc(rec.nbytes).upload(rec)
    got = ctx.volume_metrics_dev(dv, dr, vol.size)
    ref = O.metrics(vol, rec)
    for k in ("mse", "mae", "psnr"):
        assert got[k] == pytest.approx(ref[k], rel=1e-9), k
    for k in ("max_abs", "vmin", "vmax"):
        assert np.float32(got[k]) == ref[k], k
    dv.free()
    dr.free()


@pytest.mark.gpu
def test_device_metrics_rejects_misaligned(backend):
    ctx = backend.ctx
    b = ctx.alloc(4096)
    with pytest.raises(ct.BackendError):
        ctx._call("ct_hip_volume_metrics_dev", ctx.handle, b.ptr + 4, b.ptr, 16,
                  ct.ctypes.byref(ct.Metrics()))
    b.free()


@pytest.mark.gpu
def test_export_and_load_sinograms(O, backend, tmp_path):
    ctx = backend.ctx
    w, h, d, na = 48, 40, 5, 60
    vol = ctx.phantom(33, 8, w, h, d)
    p = tmp_path / "sinos.npy"
    assert ct.exportAllSinograms(backend, vol, na, p)
    s = np.load(p)
    assert s.shape == (d, na, w)
    lo, hi = np.float32(np.inf), np.float32(-np.inf)
    for z in range(d):
        sz = backend.computeSinogram(vol[z], na, w)
        np.testing.assert_array_equal(s[z], sz.data)  # the same device Radon
        ref, mm = O.compute_sinogram(vol[z], na, w)
        assert O.rel_l2(s[z], ref) <= 1e-5
        lo, hi = min(lo, np.float32(mm[0])), max(hi, np.float32(mm[1]))
    meta = json.load(open(_sidecar(p)))
    assert meta["version"] == 1 and meta["depth"] == d and meta["num_angles"] == na
    assert meta["detector_bins"] == w and meta["detector_spacing_mm"] == 1
    assert np.float32(meta["original_min_hu"]) == lo and np.float32(meta["original_max_hu"]) == hi
    step = float(np.float32(180.0) / np.float32(na))
    assert meta["angles_deg"] == [step * a for a in range(na)]
    assert list(meta) == sorted(meta)
    # consumer: load and reconstruct = reconstruct from the in-memory stack
    L = ct.loadSinograms(p)
    assert L.meta_found and L.meta_minmax and L.meta_angles
    params = ct.ReconstructionParams(ct.FilterType.SheppLogan, na)
    a = backend.reconstructVolumeFromSinograms(L.sinograms, L.angles_deg, params)
    b = backend.reconstructVolumeFromSinograms(s, np.float32(step) * np.arange(na, dtype=np.float32),
                                               params)
    np.testing.assert_array_equal(a, b)


def test_sidecar_number_forms(tmp_path):
    """JSON numbers in exponent form, an underflowing one and a negative zero;
    a malformed number makes the sidecar unparsable (every default)."""
    p, s = _write_stack(tmp_path, shape=(1, 4, 8))
    with open(_sidecar(p), "w") as f:
        f.write('{"angles_deg": [2.5E+1, 1e-400, -0.0, 4.5e-1],'
                ' "detector_spacing_mm": 125e-3, "original_min_hu": -1.5E3, "original_max_hu": 2e3}')
    L = ct.loadSinograms(p)
    assert L.meta_found and L.meta_angles and L.meta_minmax
    np.testing.assert_array_equal(L.angles_deg, np.array([25.0, 0.0, -0.0, 0.45], np.float32))
    assert L.detector_spacing_mm == 0.125 and L.min_hu == -1500.0 and L.max_hu == 2000.0
    for bad in ('{"detector_spacing_mm": 01}', '{"detector_spacing_mm": 1.}',
                '{"detector_spacing_mm": +1}', '{"detector_spacing_mm": .5}'):
        with open(_sidecar(p), "w") as f:
            f.write(bad)
        L = ct.loadSinograms(p)
        assert not L.meta_found and L.detector_spacing_mm == 1.0, bad
