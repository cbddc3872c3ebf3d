"""Pins the CPU oracle (no GPU).

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a swapped sin/cos, a wrong offset or sign, a kept Nyquist bin) fails:
  * bit-exact agreement with the reference's own FFT.cpp / FilterDesign.cpp
    (golden vectors in tests/golden/, made by tests/golden/make_golden.py);
  * textbook transforms (numpy FFT in FP64) and closed forms;
  * special angles whose projections are column / row sums;
  * invariants (mass, linearity, the raw-zero skip, the y-flip).
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_fft_filter.npz")


# ----------------------------------------------------------- golden vectors
def test_filters_bit_exact_vs_reference_build(O):
    g = np.load(GOLD)
    keys = [k for k in g.files if k.startswith("filter_")]
    assert len(keys) == 42
    for k in keys:
        n = int(k.split("_")[1][1:])
        t = int(k.split("_")[2][1:])
        mine = O.create_filter(n, t)
        assert mine.view(np.uint32).tolist() == g[k].view(np.uint32).tolist(), k


@pytest.mark.parametrize("n", [8, 64, 512, 2048])
def test_fft_bit_exact_vs_reference_build(O, n):
    g = np.load(GOLD)
    fwd = O.forward_real(g[f"fwd_in_n{n}"]).view(np.float32)
    assert np.array_equal(fwd.view(np.uint32), g[f"fwd_out_n{n}"].view(np.uint32))
    inv = O.inverse_real(g[f"inv_in_n{n}"].view(np.complex64))
    assert np.array_equal(inv.view(np.uint32), g[f"inv_out_n{n}"].view(np.uint32))


# ----------------------------------------------------------- textbook FFT
@pytest.mark.parametrize("n", [64, 512, 2048])
def test_forward_real_matches_numpy(O, n):
    x = np.random.default_rng(n).standard_normal(n).astype(np.float32)
    ref = np.fft.rfft(x.astype(np.float64))
    got = O.forward_real(x).astype(np.complex128)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 2e-6


@pytest.mark.parametrize("n", [64, 2048])
def test_inverse_real_drops_nyquist(O, n):
    rng = np.random.default_rng(7)
    X = (rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)).astype(np.complex64)
    X[0] = X[0].real
    got = O.inverse_real(X).astype(np.float64)
    # closed form of the quirk: the unpacking index (n2-i)%n2 reads X[0] where
    # X[n/2] belongs, so the transform is irfft with X[n/2] := X[0] (X[0] real)
    Xz = X.astype(np.complex128).copy()
    Xz[-1] = Xz[0]
    ref = np.fft.irfft(Xz, n)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 2e-6
    # in FBP the spectrum's DC is already 0 (H[0] = 0), so Nyquist acts as 0
    X0 = X.copy()
    X0[0] = 0.0
    Xz = X0.astype(np.complex128)
    Xz[-1] = 0.0
    ref0 = np.fft.irfft(Xz, n)
    got0 = O.inverse_real(X0).astype(np.float64)
    assert np.linalg.norm(got0 - ref0) / np.linalg.norm(ref0) < 2e-6
    # and changing the Nyquist bin changes nothing, bit for bit
    X2 = X.copy()
    X2[-1] = 123.0 + 45.0j
    assert np.array_equal(O.inverse_real(X2), O.inverse_real(X))


# ----------------------------------------------------------- filter design
@pytest.mark.parametrize("n", [64, 512, 2048])
def test_ramp_closed_form(O, n):
    # DFT of the band-limited ramp kernel: H[k] = 2k/n up to the truncation
    # of the kernel at n/2 taps (|error| <= 4/(pi^2 n) * n/ (n) ~ 1/n).
    H = O.create_filter(n, O.RAMP).astype(np.float64)
    k = np.arange(n // 2 + 1)
    assert H[0] == 0.0
    assert np.max(np.abs(H[1:] - 2.0 * k[1:] / n)) < 2.0 / n


@pytest.mark.parametrize("t", range(6))
def test_filter_fp32_vs_fp64(O, t):
    n = 2048
    h32 = O.create_filter(n, t).astype(np.float64)
    h64 = O.filter64(n, t)
    assert h64[-1] == 0.0 and h64[0] == 0.0
    # the FP64 design zeroes Nyquist; the FP32 one keeps it but it is never read
    assert np.max(np.abs(h32[:-1] - h64[:-1])) < 1e-5


def test_windows_closed_form(O):
    n = 1024
    k = np.arange(n // 2 + 1)
    freq = k / n
    ramp = O.filter64(n, O.RAMP)
    np.testing.assert_allclose(O.filter64(n, O.HANN)[1:-1],
                               (ramp * (0.5 + 0.5 * np.cos(2 * np.pi * freq)))[1:-1], rtol=1e-12)
    sl = ramp * np.where(k > 0, np.sin(np.pi * freq) / np.where(k > 0, np.pi * freq, 1), 1)
    np.testing.assert_allclose(O.filter64(n, O.SHEPP_LOGAN)[1:-1], sl[1:-1], rtol=1e-12)
    # unknown ids fall back to Shepp-Logan (FilterDesign.cpp:25-27)
    assert np.array_equal(O.create_filter(n, 17), O.create_filter(n, O.SHEPP_LOGAN))


def test_row_filter_is_padded_toeplitz(O):
    """reconstructSlice's FFT filter equals the exact circular convolution with
    the FP64 kernel (DC and Nyquist removed), restricted to [0, sq)."""
    bins = 128
    sq, pad, nfft = O.geometry(bins)
    assert (sq, pad, nfft) == (182, 27, 512)
    rng = np.random.default_rng(3)
    sino = rng.random((6, bins)).astype(np.float32)
    ang = O.serial_angles(6)
    _, filt = O.reconstruct_slice(sino, ang, 8, O.RAMP, want_filtered=True)
    h = O.kernel64(nfft, O.RAMP)
    i = np.arange(sq)[:, None]
    j = np.arange(bins)[None, :] + pad
    T = h[(i - j) % nfft]
    ref = sino.astype(np.float64) @ T.T
    assert O.rel_l2(filt, ref) < 2e-6
    # keeping the Nyquist bin would be a visible mistake at this tolerance
    H = O.filter64(nfft, O.RAMP)
    H[-1] = 2.0 * (nfft // 2) / nfft
    hk = np.fft.irfft(np.concatenate([H]).astype(np.complex128), nfft)
    bad = sino.astype(np.float64) @ hk[(i - j) % nfft].T
    assert O.rel_l2(bad, ref) > 1e-4


def test_geometry_rules(O):
    assert O.geometry(512) == (725, 106, 2048)
    assert O.geometry(128) == (182, 27, 512)
    assert O.geometry(10) == (15, 2, 64)
    assert O.next_pow2(1) == 1 and O.next_pow2(725 * 2) == 2048 and O.next_pow2(4096) == 4096


# ----------------------------------------------------------- Serial Radon
def _random_slice(rng, h, w, zero_frac=0.3):
    s = rng.uniform(-0.5, 1.0, (h, w)).astype(np.float32)
    s[rng.random((h, w)) < zero_frac] = 0.0
    return s


def test_sinogram_angle0_is_column_sums(O):
    rng = np.random.default_rng(11)
    w = h = 32
    s = _random_slice(rng, h, w)
    sino, mm = O.compute_sinogram(s, 4, w)
    span = mm[1] - mm[0]
    v = np.where(s != 0, (s.astype(np.float64) - mm[0]) / span, 0.0)
    col = v.sum(axis=0)  # theta = 0: s = x + 0.5 -> half to bin x, half to x+1
    ref = np.zeros(w)
    ref[:-1] += 0.5 * col[:-1]
    ref[1:] += 0.5 * col[:-1]
    ref[-1] += col[-1]  # floor(s) == bins-1 takes the full value
    np.testing.assert_allclose(sino[0], ref, rtol=1e-5, atol=1e-5)


def test_sinogram_angle90_is_row_sums(O):
    rng = np.random.default_rng(12)
    w = h = 32
    s = _random_slice(rng, h, w)
    sino, mm = O.compute_sinogram(s, 4, w)  # a=2 -> 90 degrees
    v = np.where(s != 0, (s.astype(np.float64) - mm[0]) / (mm[1] - mm[0]), 0.0)
    row = v.sum(axis=1)
    ref = np.zeros(h)
    ref[:-1] += 0.5 * row[:-1]
    ref[1:] += 0.5 * row[:-1]
    ref[-1] += row[-1]
    np.testing.assert_allclose(sino[2], ref, rtol=1e-4, atol=1e-4)


def test_sinogram_mass_conservation(O):
    # object well inside the field: every deposit lands in [0, bins-1)
    rng = np.random.default_rng(13)
    s = np.zeros((48, 48), np.float32)
    s[16:32, 14:30] = rng.uniform(0.1, 1.0, (16, 16)).astype(np.float32)
    sino, mm = O.compute_sinogram(s, 37, 48)
    v = np.where(s != 0, (s.astype(np.float64) - mm[0]) / (mm[1] - mm[0]), 0.0)
    np.testing.assert_allclose(sino.sum(axis=1), v.sum(), rtol=1e-5)


def test_sinogram_skips_raw_zero(O):
    # min < 0: a raw-zero background pixel would carry (0-min)/span > 0 if it
    # were not skipped (SerialBackend.cpp:104)
    s = np.zeros((16, 16), np.float32)
    s[5, 5] = -0.25
    s[9, 9] = 1.0
    a, mm = O.compute_sinogram(s, 5, 16)
    s2 = s.copy()
    s2[2, 3] = np.float32(1e-30)  # nonzero: now included with weight ~ (0 - min)/span
    b, _ = O.compute_sinogram(s2, 5, 16)
    extra = (b - a).sum(axis=1)
    np.testing.assert_allclose(extra, 0.25 / 1.25, rtol=1e-5)


def test_sinogram_fp32_vs_fp64(O):
    rng = np.random.default_rng(14)
    s = _random_slice(rng, 64, 64)
    a, _ = O.compute_sinogram(s, 90, 64)
    b = O.compute_sinogram64(s, 90, 64)
    assert O.rel_l2(a, b) < 1e-5


# ----------------------------------------------------------- Serial FBP
def _disk(n, cx, cy, r, val=1.0):
    y, x = np.mgrid[0:n, 0:n]
    return np.where((x - cx) ** 2 + (y - cy) ** 2 <= r * r, val, 0.0).astype(np.float32)


def test_fbp_reconstructs_disk_flipped_in_y(O):
    n = 64
    s = _disk(n, 20.0, 40.0, 5.0, 1.0)
    sino, mm = O.compute_sinogram(s, 180, n)
    rec = O.reconstruct_slice(sino, O.serial_angles(180), n, O.RAMP)
    # a point (x0, y0) reconstructs at (x0 + 1/2, N - 1/2 - y0): Serial's forward
    # uses +sin, its backprojection -sin (SerialBackend.cpp:110 vs :229)
    w = np.where(rec > 0.5 * rec.max(), rec, 0.0)
    y, x = np.mgrid[0:n, 0:n]
    cx, cy = (w * x).sum() / w.sum(), (w * y).sum() / w.sum()
    assert abs(cx - 20.5) < 0.3 and abs(cy - (n - 0.5 - 40.0)) < 0.3
    # interior amplitude: the disk's normalised value is 1 (min 0, max 1)
    assert abs(rec[n - 1 - 40, 20] - 1.0) < 0.1 or abs(rec[n - 40, 20] - 1.0) < 0.1
    # a y-flip plus half-pixel shift explains the image; the unflipped one does not
    tgt = _disk(n, 20.5, n - 0.5 - 40.0, 5.0, 1.0)
    m = _disk(n, 32, 32, 28) > 0
    assert np.abs(rec - tgt)[m].mean() < 0.3 * np.abs(rec - s)[m].mean()


def test_fbp_linearity(O):
    rng = np.random.default_rng(15)
    na, bins = 24, 40
    s1 = rng.random((na, bins)).astype(np.float32)
    s2 = rng.random((na, bins)).astype(np.float32)
    ang = O.serial_angles(na)
    r1 = O.reconstruct_slice(s1, ang, bins, O.HANN).astype(np.float64)
    r2 = O.reconstruct_slice(s2, ang, bins, O.HANN).astype(np.float64)
    r12 = O.reconstruct_slice((2 * s1 + s2).astype(np.float32), ang, bins, O.HANN)
    assert O.rel_l2(r12, 2 * r1 + r2) < 1e-5


def test_fbp_mask_and_scale(O):
    # constant sinogram rows: outside the inscribed circle the output is 0
    na, bins = 16, 32
    sino = np.ones((na, bins), np.float32)
    rec = O.reconstruct_slice(sino, O.serial_angles(na), bins, O.RAMP)
    R = bins // 2
    y, x = np.mgrid[0:bins, 0:bins]
    outside = (x - R) ** 2 + (y - R) ** 2 > R * R
    assert np.all(rec[outside] == 0.0)
    assert np.all(np.isfinite(rec))


def test_fbp_fp32_vs_fp64(O):
    rng = np.random.default_rng(16)
    s = _random_slice(rng, 48, 48)
    sino32, _ = O.compute_sinogram(s, 60, 48)
    sino64 = O.compute_sinogram64(s, 60, 48)
    r32 = O.reconstruct_slice(sino32, O.serial_angles(60), 48, O.SHEPP_LOGAN)
    r64 = O.reconstruct_slice64(sino64, np.arange(60) * 3.0, 48, O.SHEPP_LOGAN)
    assert O.rel_l2(r32, r64) < 1e-4


def test_volume_denormalisation_and_setslice_quirk(O):
    vol = O.phantom(5, 6, 24, 24, 5)
    out, sino = O.reconstruct_volume(vol, 30, O.RAMP, want_sino=True)
    for z in range(5):
        s, mm = O.compute_sinogram(vol[z], 30, 24)
        assert np.array_equal(sino[z], s)
        r = O.reconstruct_slice(s, O.serial_angles(30), 24, O.RAMP)
        span = mm[1] - mm[0]
        eff = span if span > 0 else np.float32(1.0)
        assert np.array_equal(out[z], (r * eff + mm[0]).astype(np.float32))
    # h != w: Buffer3D::setSlice silently keeps the zero-initialised slices
    vol2 = np.ones((2, 20, 24), np.float32)
    out2, _ = O.reconstruct_volume(vol2, 8, O.RAMP)
    assert np.all(np.isnan(out2)) or np.all(out2[np.isfinite(out2)] == 0)


# ----------------------------------------------------------- phantom
def test_phantom_param_ranges_and_determinism(O):
    e = O.phantom_params(42, 200)
    arr = np.array([[getattr(q, f) for f, _ in O.Ellipsoid._fields_] for q in e])
    assert np.all(np.abs(arr[:, 0:3]) <= 0.6)
    assert np.all((arr[:, 3:6] >= 0.05) & (arr[:, 3:6] <= 0.4))
    assert np.all((arr[:, 6:9] >= 0) & (arr[:, 6:9] < np.pi))
    assert np.all((arr[:, 9] >= -0.3) & (arr[:, 9] <= 1.0))
    e2 = O.phantom_params(42, 200)
    assert all(bytes(a) == bytes(b) for a, b in zip(e, e2))
    e3 = O.phantom_params(43, 200)
    assert bytes(e[0]) != bytes(e3[0])


def test_phantom_single_ellipsoid_membership_fp64(O):
    e = O.phantom_params(9, 1)
    q = e[0]
    q.rho = 1.0
    n = 48
    vol = O.phantom_from(e, 1, n, n, n)
    # FP64 membership with the same ZYZ rotation, independent of the float path
    a, b, g = float(q.phi), float(q.theta), float(q.psi)
    Rz = lambda t: np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1]])
    Ry = lambda t: np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])
    R = Rz(a) @ Ry(b) @ Rz(g)
    c = np.linspace(-1, 1, n)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    d = np.stack([X - q.cx, Y - q.cy, Z - q.cz], -1) @ R  # = (R^T d) per voxel
    f = (d[..., 0] / q.ax) ** 2 + (d[..., 1] / q.ay) ** 2 + (d[..., 2] / q.az) ** 2
    ref = (f <= 1.0).astype(np.float32)
    boundary = np.abs(f - 1.0) < 1e-4
    assert np.array_equal(vol[~boundary], ref[~boundary])
    # voxel count ~ volume of the ellipsoid / voxel volume
    dv = (2.0 / (n - 1)) ** 3
    expect = 4.0 / 3.0 * np.pi * q.ax * q.ay * q.az / dv
    assert abs(vol.sum() - expect) / expect < 0.05


def test_phantom_additive(O):
    e = O.phantom_params(21, 4)
    vol_all = O.phantom_from(e, 4, 20, 20, 20).astype(np.float64)
    acc = np.zeros_like(vol_all)
    for k in range(4):
        one = (O.Ellipsoid * 1)(e[k])
        acc += O.phantom_from(one, 1, 20, 20, 20)
    np.testing.assert_allclose(vol_all, acc, atol=1e-6)


def test_sinogram_minmax_skips_nan(O):
    # Serial's scan compares with < and > from FLT_MAX / lowest
    # (SerialBackend.cpp:69-76): a NaN pixel never becomes an extreme
    rng = np.random.default_rng(17)
    s = _random_slice(rng, 24, 24)
    s[3, 4] = np.nan
    _, mm = O.compute_sinogram(s, 6, 24)
    assert mm[0] == np.nanmin(s) and mm[1] == np.nanmax(s)
    _, mm = O.compute_sinogram(np.full((4, 4), np.nan, np.float32), 3, 4)
    assert mm[0] == np.finfo(np.float32).max and mm[1] == -np.finfo(np.float32).max


# ----------------------------------------------------------- OpenMP baseline
# The timed CPU baseline restates OpenMPBackend::reconstructVolume
# (OpenMPBackend.cpp:61-190, RadonTransform.cpp:25-113,
# OptimizedBackprojectionCPU.cpp:22-174). Its conventions differ from
# Serial's, and each check below fails if the restatement took Serial's
# instead: -sin in both directions (no y-flip), (int) truncation of the bin,
# the (sq-1)/2 detector centre of the backprojection.

@pytest.mark.parametrize("w", [32, 33])
def test_omp_radon_angle0_is_column_sums(O, w):
    rng = np.random.default_rng(21)
    s = np.abs(_random_slice(rng, w, w))
    sino = O.omp_radon_forward(s, 4, w)
    v = s.astype(np.float64) / s.max()  # min 0 (zeros present), max -> 1
    col = v.sum(axis=0)
    ref = np.zeros(w)
    if w % 2 == 0:  # u = x + 1/2: half to bin x, half to x + 1; the last bin takes all
        ref[:-1] += 0.5 * col[:-1]
        ref[1:] += 0.5 * col[:-1]
        ref[-1] += col[-1]
    else:  # u = x exactly: the whole pixel to bin x
        ref[:] = col
    np.testing.assert_allclose(sino[0], ref, rtol=1e-5, atol=1e-5)


def test_omp_radon_angle90_reverses_rows(O):
    # -sin: at 90 degrees row y lands on bin ~ (h - 1 - y) + 1/2, the reverse
    # of Serial's +sin order
    rng = np.random.default_rng(22)
    n = 32
    s = np.abs(_random_slice(rng, n, n))
    sino = O.omp_radon_forward(s, 4, n)  # a = 2 -> 90 degrees
    v = s.astype(np.float64) / s.max()
    row = v.sum(axis=1)
    ref = np.zeros(n)
    ref[n - 1] += row[0]  # u = n - 1/2: the last bin takes all
    for y in range(1, n):
        ref[n - 1 - y] += 0.5 * row[y]
        ref[n - y] += 0.5 * row[y]
    np.testing.assert_allclose(sino[2], ref, rtol=1e-4, atol=1e-4)
    serial, _ = O.compute_sinogram(s, 4, n)
    assert O.rel_l2(sino[2], serial[2]) > 0.1  # the two conventions really differ


def test_omp_radon_truncates_negative_bins(O):
    # pixel (0, 15) of a 16 x 16 slice at 10 degrees: u = 8 - 7.5 (cos + sin)
    # = -0.69. (int) truncation sends it to bins 0 and 1 with weights 1 - u > 1
    # and u < 0 (mass kept); floor would drop it
    s = np.zeros((16, 16), np.float32)
    s[15, 0] = 1.0
    sino = O.omp_radon_forward(s, 18, 16)  # a = 1 -> 10 degrees
    row = sino[1]
    assert row[0] > 1.0 and row[1] < 0.0 and np.all(row[2:] == 0.0)
    assert abs(row.sum() - 1.0) < 1e-6


def _omp_disk_volume(n, d, cx, cy, r):
    y, x = np.mgrid[0:n, 0:n]
    sl = np.where((x - cx) ** 2 + (y - cy) ** 2 <= r * r, 1.0, 0.0).astype(np.float32)
    return np.repeat(sl[None], d, axis=0)


def test_omp_volume_disk_unflipped(O):
    # no y-flip: a disk at (x0, y0) reconstructs at (x0, y0) with amplitude ~1
    # (min 0, max 1, so denormalisation is the identity); sq = 91 is odd, so
    # the (sq - 1)/2 centre meets the forward's bins/2 exactly
    n, d = 64, 4
    vol = _omp_disk_volume(n, d, 20.0, 40.0, 5.0)
    out, ph = O.omp_reconstruct_volume(vol, 180, O.RAMP, 2)
    assert ph[0] > 0 and ph[1] > 0 and ph[2] >= ph[0] + ph[1]
    y, x = np.mgrid[0:n, 0:n]
    for z in range(d):
        rec = out[z]
        wgt = np.where(rec > 0.5 * rec.max(), rec, 0.0)
        cx, cy = (wgt * x).sum() / wgt.sum(), (wgt * y).sum() / wgt.sum()
        assert abs(cx - 20.0) < 0.3 and abs(cy - 40.0) < 0.3, (z, cx, cy)
        assert abs(rec[40, 20] - 1.0) < 0.1, rec[40, 20]


def test_omp_volume_linearity(O):
    # with min 0 the per-slice normalisation and denormalisation cancel, so the
    # baseline is linear in the volume (to float rounding)
    rng = np.random.default_rng(23)
    n, d = 40, 3
    v1 = np.abs(rng.standard_normal((d, n, n))).astype(np.float32)
    v2 = np.abs(rng.standard_normal((d, n, n))).astype(np.float32)
    v1[:, 0, 0] = 0.0
    v2[:, 0, 0] = 0.0
    a, _ = O.omp_reconstruct_volume(v1, 36, O.HANN, 2)
    b, _ = O.omp_reconstruct_volume(v2, 36, O.HANN, 2)
    c, _ = O.omp_reconstruct_volume(v1 + v2, 36, O.HANN, 2)
    assert O.rel_l2(c, a.astype(np.float64) + b) < 1e-5
