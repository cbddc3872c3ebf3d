"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars: phantom voxels bit-exact; every pixel's detector bin and fraction
bit-exact (Serial's float expression for s; a bin with one deposit is
Serial's rounded product bit for bit); sinograms within relative L2 <= 1e-5
of Serial (each lane sums its interval's two deposit kinds apart with fused
multiply-adds, so sums of several deposits differ from Serial's raster-order
sums by rounding only, ~5e-7); filtered reconstructions within relative L2
<= 1e-5 of the Serial restatement (FP32 FFT vs Serial's FP32 FFT, FMA
interpolation); every reconstruction also reports its error vs the FP64
oracle.
"""
import threading

import numpy as np
import pytest

import qtctreconstruction_amd as ct

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative L2 vs the FP32 Serial restatement


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                          np.asarray(b, np.float32).view(np.uint32))


SINO_TOL = 2e-6  # observed ~5e-7: rounding of the fused, split sums only


def _sino_parity(O, got, ref):
    """Every angle within SINO_TOL (below the 1e-5 bar) of Serial."""
    got = np.asarray(got, np.float32)
    assert np.isfinite(got).all()
    err = O.rel_l2(got, ref)
    assert err <= SINO_TOL, err
    return err


# ----------------------------------------------------------------- phantom
@pytest.mark.parametrize("shape,n,seed", [((64, 64, 64), 10, 1), ((96, 80, 40), 24, 7),
                                          ((128, 128, 16), 200, 42), ((33, 17, 5), 3, 9),
                                          ((16, 16, 4), 0, 3)])
def test_phantom_bit_exact(backend, O, shape, n, seed):
    nx, ny, nz = shape
    got = backend.ctx.phantom(seed, n, nx, ny, nz)
    ref = O.phantom(seed, n, nx, ny, nz)
    assert _bits_equal(got, ref)


# ----------------------------------------------------------------- Radon
def _slice(rng, h, w, zero_frac=0.3):
    s = rng.uniform(-0.5, 1.0, (h, w)).astype(np.float32)
    s[rng.random((h, w)) < zero_frac] = 0.0
    return s


@pytest.mark.parametrize("h,w,na,bins", [(64, 64, 180, 64), (97, 97, 33, 97), (50, 64, 45, 70),
                                         (64, 50, 7, 40), (32, 32, 1, 32), (128, 128, 360, 128),
                                         (20, 20, 12, 3), (9, 8, 16, 1), (12, 10, 30, 2),
                                         (24, 600, 20, 600), (600, 24, 20, 300),
                                         (40, 772, 30, 772), (12, 770, 16, 1600),
                                         (40, 1028, 24, 1028), (10, 1100, 10, 1100),
                                         (1100, 10, 10, 700)])
def test_sinogram_parity(backend, O, h, w, na, bins):
    """Row lengths at the edges of each walk form (whole rows staged by 512,
    768 or 1024 lanes up to 516, 772 and 1028 pixels; the windowed 256-lane
    form beyond), with detector widths that need several bin blocks."""
    rng = np.random.default_rng(h * 1000 + na)
    s = _slice(rng, h, w)
    got = backend.computeSinogram(s, na, bins)
    ref, mm = O.compute_sinogram(s, na, bins)
    _sino_parity(O, got.data, ref)
    assert got.original_min_hu == mm[0] and got.original_max_hu == mm[1]
    assert _bits_equal(got.angles_deg, O.serial_angles(na))


def test_sinogram_phantom_slice_and_constant(backend, O):
    vol = O.phantom(5, 24, 128, 128, 3)
    for z in range(3):
        got = backend.computeSinogram(vol[z], 180, 128)
        ref, _ = O.compute_sinogram(vol[z], 180, 128)
        _sino_parity(O, got.data, ref)
    const = np.full((16, 16), 2.5, np.float32)  # span < 1e-6 -> 1
    got = backend.computeSinogram(const, 8, 16)
    ref, _ = O.compute_sinogram(const, 8, 16)
    _sino_parity(O, got.data, ref)


@pytest.mark.parametrize("h,w,na,bins", [(97, 97, 180, 97), (64, 50, 90, 70), (50, 64, 45, 40),
                                         (20, 20, 12, 3), (33, 33, 360, 33), (16, 700, 24, 700),
                                         (700, 16, 24, 100), (8, 1100, 12, 1100)])
def test_sinogram_indexing_bit_exact(backend, O, h, w, na, bins):
    """Every pixel's bins and fraction are Serial's at every angle: with one
    nonzero pixel in the slice each bin receives at most one deposit, so the
    sinogram is bit-identical whatever the order of the sums. Corners, edges
    (the zero pads of both walks) and random interior pixels."""
    rng = np.random.default_rng(h * 7 + na)
    spots = [(0, 0), (0, w - 1), (h - 1, 0), (h - 1, w - 1), (h // 2, 0), (0, w // 2)]
    spots += [(int(rng.integers(h)), int(rng.integers(w))) for _ in range(6)]
    for y, x in spots:
        s = np.zeros((h, w), np.float32)
        s[y, x] = rng.uniform(0.5, 2.0)
        got = backend.computeSinogram(s, na, bins)
        ref, _ = O.compute_sinogram(s, na, bins)
        assert _bits_equal(got.data, ref), (y, x)


@pytest.mark.parametrize("h,w,na,bins", [(9, 9, 180, 1), (7, 9, 180, 2), (9, 8, 90, 3),
                                         (24, 24, 180, 24), (29, 31, 90, 31)])
def test_sinogram_first_bin_edge_bit_exact(backend, O, h, w, na, bins):
    """Pixels whose s lies just either side of 0 at some angle (i0 = -1 drops
    both deposits, i0 = 0 keeps them): with a detector of 1 to 3 bins, dc is
    0 or 1 and s carries bits below 2^-23, so a fraction formed as s + 1
    would round. One pixel per slice, bit for bit against Serial."""
    ang = O.serial_angles(na)
    cs = np.array([np.cos(O.deg2rad(a)) for a in ang], np.float32)
    sn = np.array([np.sin(O.deg2rad(a)) for a in ang], np.float32)
    cx, cy = np.float32(w - 1) * np.float32(0.5), np.float32(h - 1) * np.float32(0.5)
    dc = np.float32(bins // 2)
    xx = (np.arange(w, dtype=np.float32) - cx)[None, :]
    yy = (np.arange(h, dtype=np.float32) - cy)[:, None]
    picks = set()
    for a in range(na):
        s = ((xx * cs[a]).astype(np.float32) + (yy * sn[a]).astype(np.float32)).astype(np.float32)
        s = (s + dc).astype(np.float32)
        m = np.abs(s) + (s == 0) * 9  # nearest nonzero s
        yx = np.unravel_index(np.argmin(m), m.shape)
        if m[yx] < 0.05:
            picks.add((int(yx[0]), int(yx[1])))
    assert picks
    for y, x in sorted(picks)[:10]:
        sl = np.zeros((h, w), np.float32)
        sl[y, x] = np.float32(1.7)
        got = backend.computeSinogram(sl, na, bins)
        ref, _ = O.compute_sinogram(sl, na, bins)
        assert _bits_equal(got.data, ref), (y, x)


@pytest.mark.parametrize("n,na", [(512, 90), (600, 24), (1100, 24)])
def test_sinogram_zero_row_skip_bit_exact(backend, O, n, na):
    """The walks skip the rows (over the transposed image, the columns) that
    hold no nonzero normalised pixel: the 16-slice walk (512 px) per pair of
    z-blocks, with each row's nonzero extent; the 8-slice walk per z-block,
    whole rows (600 px) or windows (1100 px). Slices with one or two nonzero
    pixels: rows marked in only one z-block of a pair, the last pair, whose
    second z-block stands in for a missing one (24 slices: 3 z-blocks), a
    pixel that normalises to exactly 0 (the slice's minimum, for which
    Serial adds +0), and an empty slice. Every bin takes at most one nonzero
    deposit: bit for bit against Serial."""
    d = 24
    vol = np.zeros((d, n, n), np.float32)
    vol[2, 5, 100] = 1.7       # z-block 0 of pair 0 only
    vol[13, 300, 7] = 0.9      # z-block 1 of pair 0 only
    vol[8, 50, 60] = -1.0      # the slice minimum: normalises to 0, its row stays unmarked
    vol[8, 400, 300] = 2.0
    vol[17, 511, 511] = 3.0    # z-block 2, pair 1 (its second z-block stands in)
    vol[23, 0, 0] = 0.25
    # slice 20 stays empty
    _, sino = backend.reconstructVolume(vol, ct.ReconstructionParams(ct.FilterType.Ramp, na),
                                        want_sinograms=True)
    for z in range(d):
        ref, _ = O.compute_sinogram(vol[z], na, n)
        assert _bits_equal(sino[z], ref), z
    assert np.all(sino[20] == 0.0)


def test_empty_inputs(backend):
    s = backend.computeSinogram(np.zeros((0, 0), np.float32), 10, 10)
    assert s.data.size == 0
    r = backend.reconstructSlice(ct.Sinogram(np.zeros((0, 0), np.float32), np.zeros(0)), 8,
                                 ct.ReconstructionParams())
    assert r.size == 0


# ----------------------------------------------------------------- FBP
@pytest.mark.parametrize("filt", list(ct.FilterType))
def test_reconstruct_slice_all_filters(backend, O, filt):
    rng = np.random.default_rng(int(filt) + 1)
    s = _slice(rng, 96, 96)
    sino, _ = O.compute_sinogram(s, 120, 96)
    ang = O.serial_angles(120)
    sg = ct.Sinogram(sino, ang)
    got = backend.reconstructSlice(sg, 96, ct.ReconstructionParams(filt, 120))
    ref = O.reconstruct_slice(sino, ang, 96, int(filt))
    assert O.rel_l2(got, ref) <= TOL
    ref64 = O.reconstruct_slice64(O.compute_sinogram64(s, 120, 96), np.arange(120) * 1.5, 96,
                                  int(filt))
    print(f"filter {filt.name}: relL2 vs Serial {O.rel_l2(got, ref):.2e}, "
          f"vs FP64 {O.rel_l2(got, ref64):.2e}")


@pytest.mark.parametrize("bins", [16, 40, 80, 150, 300, 600, 1000])
def test_fft_filter_every_length(backend, O, bins):
    """Detector widths whose padded rows need every transform length from 64
    to 4096 (each remainder of log2(nfft) mod 3 of the fused passes)."""
    nfft = O.geometry(bins)[2]
    assert nfft == {16: 64, 40: 128, 80: 256, 150: 512, 300: 1024, 600: 2048, 1000: 4096}[bins]
    rng = np.random.default_rng(bins)
    na = 6
    sino = rng.standard_normal((na, bins)).astype(np.float32)
    ang = O.serial_angles(na)
    for filt in (ct.FilterType.Ramp, ct.FilterType.Hann):
        got = backend.reconstructSlice(ct.Sinogram(sino, ang), 48, ct.ReconstructionParams(filt, na))
        ref = O.reconstruct_slice(sino, ang, 48, int(filt))
        assert O.rel_l2(got, ref) <= TOL, (bins, filt)


def _filtered_rows_parity(backend, O, bins):
    """The filtered rows of a stack whose last z-block holds one row match
    Serial's filtered rows at 1e-5, row by row, and the FP64 padded
    convolution, for three filters."""
    rng = np.random.default_rng(bins)
    nz, na = 5, 7
    sinos = rng.standard_normal((nz, na, bins)).astype(np.float32)
    ang = O.serial_angles(na)
    sq, pad, nfft = O.geometry(bins)
    for filt in (ct.FilterType.Ramp, ct.FilterType.SheppLogan, ct.FilterType.Hann):
        got = backend.ctx.filter_sinograms(sinos, filt)
        assert got.shape == (nz, na, sq)
        h = O.kernel64(nfft, int(filt))
        T = h[(np.arange(sq)[:, None] - np.arange(bins)[None, :] - pad) % nfft]
        for z in range(nz):
            _, ref = O.reconstruct_slice(sinos[z], ang, bins, int(filt), want_filtered=True)
            for a in range(na):
                assert O.rel_l2(got[z, a], ref[a]) <= TOL, (bins, filt, z, a)
            ref64 = sinos[z].astype(np.float64) @ T.T
            assert O.rel_l2(got[z], ref64) <= TOL, (bins, filt, z)


@pytest.mark.parametrize("bins", [91, 150, 212, 213, 300, 424, 425, 512])
def test_wave_filter_every_split(backend, O, bins):
    """Rows of 91 .. 512 bins take the short-transform filter: R = 1, 2, 3
    sub-transforms of 512 points (the first and last width of each R); the
    stack's last z-block holds one row, so the second wavefront's rows and
    the first's odd row read as zero."""
    _filtered_rows_parity(backend, O, bins)


@pytest.mark.parametrize("bins", [23, 26, 50, 724, 725, 727, 768, 848, 849])
def test_lds_filter_short_length(backend, O, bins):
    """Widths whose kernel offsets (2D + 1) fit a power of two below Serial's
    nfft (23-26, 46-53, 725-848: 64, 128 and 2048 points instead of 128, 256
    and 4096) take the LDS FFT at that length with the FP64-designed short
    spectrum; 724 and 849 are the neighbours that keep Serial's length."""
    _filtered_rows_parity(backend, O, bins)


def test_filter_design_bit_exact(backend, O):
    """designFilter returns FilterDesign::createFilter's taps bit for bit
    (unknown types fall back to Shepp-Logan), within 2.5e-7 of the FP64
    design."""
    for n in (4, 64, 512, 2048, 4096):
        for t in (0, 1, 2, 3, 4, 5, 9):
            got = backend.ctx.design_filter(n, t)
            assert _bits_equal(got, O.create_filter(n, t)), (n, t)
            if n >= 64 and t <= 5:
                ref = O.filter64(n, t)[: n // 2]  # the FP64 design zeroes bin n/2
                assert np.max(np.abs(got[: n // 2] - ref)) <= 3e-7, (n, t)


@pytest.mark.parametrize("out_size,angles", [(64, "serial"), (100, "serial"), (40, "random")])
def test_reconstruct_slice_geometry_variants(backend, O, out_size, angles):
    rng = np.random.default_rng(out_size)
    na, bins = 50, 64
    sino = rng.random((na, bins)).astype(np.float32)
    ang = O.serial_angles(na) if angles == "serial" else \
        np.sort(rng.uniform(0, 360, na)).astype(np.float32)
    got = backend.reconstructSlice(ct.Sinogram(sino, ang), out_size,
                                   ct.ReconstructionParams(ct.FilterType.SheppLogan, na))
    ref = O.reconstruct_slice(sino, ang, out_size, O.SHEPP_LOGAN)
    assert O.rel_l2(got, ref) <= TOL


def test_volume_c1_config(backend, O):
    """128^3, 10 ellipsoids, 180 angles, Ram-Lak: the full Serial volume."""
    vol = backend.ctx.phantom(2026, 10, 128, 128, 128)
    params = ct.ReconstructionParams(ct.FilterType.Ramp, 180)
    out, sino = backend.reconstructVolume(vol, params, want_sinograms=True)
    ref, ref_sino = O.reconstruct_volume(vol, 180, O.RAMP, want_sino=True)
    _sino_parity(O, sino, ref_sino)
    err = O.rel_l2(out, ref)
    zl = [0, 37, 64, 127]
    ref64 = O.reconstruct_volume64(vol, 180, O.RAMP, z_list=zl)
    err64 = O.rel_l2(out[zl], ref64[zl])
    print(f"C1: relL2 vs Serial {err:.2e}, vs FP64 (4 slices) {err64:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("d", [1, 3, 5, 6, 13, 20, 41])
def test_volume_ragged_depth(backend, O, d):
    vol = O.phantom(d, 8, 40, 40, d)
    params = ct.ReconstructionParams(ct.FilterType.Hann, 36)
    out = backend.reconstructVolume(vol, params)
    ref, _ = O.reconstruct_volume(vol, 36, O.HANN)
    assert O.rel_l2(out, ref) <= TOL


@pytest.mark.parametrize("n,d,na", [(48, 20, 30), (64, 41, 24), (600, 17, 6)])
def test_volume_sinograms_every_slice(backend, O, n, d, na):
    """The volume Radon stage (8-slice z-blocks, an odd count among them,
    the last one partial) against Serial slice by slice; 600 px rows take
    the windowed walk."""
    vol = O.phantom(n + d, 10, n, n, d)
    _, sino = backend.reconstructVolume(vol, ct.ReconstructionParams(ct.FilterType.Ramp, na),
                                        want_sinograms=True)
    for z in range(d):
        ref, _ = O.compute_sinogram(vol[z], na, n)
        _sino_parity(O, sino[z], ref)


def test_volume_non_square_slices_stay_zero(backend, O):
    vol = O.phantom(3, 8, 40, 24, 4)  # h != w: Buffer3D::setSlice drops every slice
    out, sino = backend.reconstructVolume(vol, ct.ReconstructionParams(ct.FilterType.Ramp, 20),
                                          want_sinograms=True)
    assert np.all(out == 0.0)
    # the sinograms are still Serial's (bins = w), and the timings are this
    # call's: the Radon stage, and 0 for the filter and the backprojection
    for z in range(4):
        ref, _ = O.compute_sinogram(vol[z], 20, 40)
        _sino_parity(O, sino[z], ref)
    t = backend.ctx.timings()
    assert t[0] > 0.0 and t[1] == 0.0 and t[2] == 0.0 and t[3] == t[0]
    assert backend.lastReconstructionTimeMs() == 0.0


def test_nan_pixel_is_not_an_extreme(backend, O):
    """Serial's min/max scan skips NaN (SerialBackend.cpp:69-76): one NaN pixel
    leaves original_min/max_hu at the finite extremes, on the slice entry and
    the volume path."""
    rng = np.random.default_rng(5)
    s = _slice(rng, 40, 40)
    s[7, 9] = np.nan
    got = backend.computeSinogram(s, 12, 40)
    _, mm = O.compute_sinogram(s, 12, 40)
    assert _bits_equal([got.original_min_hu, got.original_max_hu], mm)
    vol = np.stack([_slice(rng, 40, 40) for _ in range(9)])
    vol[4, 20, 30] = np.nan
    vol[8, :, :] = np.nan  # all NaN: Serial keeps FLT_MAX / lowest
    d_vol = backend.ctx.alloc(vol.nbytes).upload(vol)
    d_sino = backend.ctx.alloc(4 * 9 * 12 * 40)
    mmv = backend.ctx.radon_dev(d_vol, 40, 40, 9, 12, 40, d_sino)
    for z in range(9):
        _, ref = O.compute_sinogram(vol[z], 12, 40)
        assert _bits_equal(mmv[z], ref), z
    d_vol.free()
    d_sino.free()


def test_trig_cache_is_bounded(backend, O):
    """Every distinct angle list gets a device table; beyond 32 lists the least
    recently used are freed, so device memory stays flat over many lists."""
    rng = np.random.default_rng(8)
    sino = rng.random((8, 16)).astype(np.float32)
    params = ct.ReconstructionParams(ct.FilterType.Ramp, 8)

    def call(k):
        ang = np.sort(rng.uniform(0.0, 180.0, 8)).astype(np.float32)
        return backend.reconstructSlice(ct.Sinogram(sino, ang), 16, params), ang

    for k in range(40):
        call(k)
    used = backend.ctx.bytes_in_use()
    for k in range(200):
        out, ang = call(k)
    assert backend.ctx.bytes_in_use() == used
    # an evicted list is rebuilt correctly when it comes back
    assert O.rel_l2(out, O.reconstruct_slice(sino, ang, 16, O.RAMP)) <= TOL


def test_radon_shape_limit_is_an_error(backend):
    """Edges beyond the proven 2048 px are refused before any device work."""
    with pytest.raises(ct.BackendError) as ei:
        backend.computeSinogram(np.ones((1, 2049), np.float32), 4, 16)
    assert ei.value.code == ct.CT_ERR_UNSUPPORTED
    with pytest.raises(ct.BackendError) as ei:
        backend.reconstructVolume(np.ones((1, 2, 2049), np.float32),
                                  ct.ReconstructionParams(ct.FilterType.Ramp, 4))
    assert ei.value.code == ct.CT_ERR_UNSUPPORTED


def test_from_sinograms(backend, O):
    rng = np.random.default_rng(77)
    nz, na, bins = 5, 60, 72
    sinos = rng.random((nz, na, bins)).astype(np.float32)
    ang = O.serial_angles(na)
    got = backend.reconstructVolumeFromSinograms(sinos, ang,
                                                 ct.ReconstructionParams(ct.FilterType.Cosine, na))
    for z in range(nz):
        ref = O.reconstruct_slice(sinos[z], ang, bins, O.COSINE)
        assert O.rel_l2(got[z], ref) <= TOL


def test_concurrent_compute_sinogram(backend, O):
    """The controller calls computeSinogram from an OpenMP loop on one backend
    (CtReconstructionController.cpp:1232-1241): the context serialises."""
    vol = O.phantom(11, 12, 48, 48, 8)
    res = [None] * 8

    def work(z):
        res[z] = backend.computeSinogram(vol[z], 30, 48).data

    th = [threading.Thread(target=work, args=(z,)) for z in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for z in range(8):
        ref, _ = O.compute_sinogram(vol[z], 30, 48)
        _sino_parity(O, res[z], ref)


def test_hbm_budget_refusal():
    be = ct.HipBackend(0, hbm_budget_bytes=1 << 20)
    vol = np.ones((16, 128, 128), np.float32)
    with pytest.raises(ct.BackendError) as ei:
        be.reconstructVolume(vol, ct.ReconstructionParams(ct.FilterType.Ramp, 90))
    assert ei.value.code == ct.CT_ERR_BUDGET


def test_budget_forces_ragged_slabs_bit_identical(O):
    """Under an HBM budget too small for the whole volume's workspace, the
    device path and the host-buffer path both run z-slabs (multiples of 8
    slices, the last one ragged) and give the unbudgeted result bit for bit."""
    n, d, na = 256, 45, 180
    vol = O.phantom(77, 12, n, n, d)
    params = ct.ReconstructionParams(ct.FilterType.Hann, na)
    ref, ref_sino = ct.HipBackend(0).reconstructVolume(vol, params, want_sinograms=True)
    nbytes = vol.nbytes
    # host path: ~2 MB of device memory per slice -> slabs of 16
    hb = ct.HipBackend(0, hbm_budget_bytes=40 << 20)
    out, sino = hb.reconstructVolume(vol, params, want_sinograms=True)
    assert _bits_equal(out, ref) and _bits_equal(sino, ref_sino)
    assert hb.ctx.bytes_in_use() <= 40 << 20
    # device path: input and output resident, ~1 MB of workspace per slice
    ctx = ct.Context(0, hbm_budget_bytes=2 * nbytes + (20 << 20))
    d_vol = ctx.alloc(nbytes).upload(vol)
    d_out = ctx.alloc(nbytes)
    ctx.reconstruct_volume_dev(d_vol, n, n, d, na, int(params.filter), d_out)
    assert _bits_equal(d_out.download(vol.shape), ref)
    assert ctx.bytes_in_use() <= 2 * nbytes + (20 << 20)
    ctx.close()
    hb.ctx.close()


def test_host_slab_ramp_bit_identical(O):
    """The host-buffer paths at 512 x 512 x 300 run slabs of 32, 64, 64, 64,
    40 and 36 slices (a ramp, a rest of 40 slices in the middle, the ragged
    last slab), their sinogram counterpart slabs of the same counts; both
    give the device-resident calls' results bit for bit."""
    n, d, na = 512, 300, 90
    vol = O.phantom(91, 10, n, n, d)
    params = ct.ReconstructionParams(ct.FilterType.SheppLogan, na)
    ctx = ct.Context(0)
    d_vol = ctx.alloc(vol.nbytes).upload(vol)
    d_out = ctx.alloc(vol.nbytes)
    d_sino = ctx.alloc(4 * d * na * n)
    ctx.reconstruct_volume_dev(d_vol, n, n, d, na, int(params.filter), d_out, d_sino)
    ref = d_out.download(vol.shape)
    ref_sino = d_sino.download((d, na, n))
    be = ct.HipBackend(0)
    out, sino = be.reconstructVolume(vol, params, want_sinograms=True)
    assert _bits_equal(out, ref) and _bits_equal(sino, ref_sino)
    ang = O.serial_angles(na)
    got = be.reconstructVolumeFromSinograms(ref_sino, ang, params)
    ctx.reconstruct_from_sinograms_dev(d_sino, d, na, n, ang, int(params.filter), d_out)
    assert _bits_equal(got, d_out.download(vol.shape))
    be.ctx.close()
    ctx.close()


def test_unsupported_width_is_an_error(backend):
    sino = np.ones((4, 3000), np.float32)
    with pytest.raises(ct.BackendError) as ei:
        backend.reconstructSlice(ct.Sinogram(sino, np.zeros(4, np.float32)), 16,
                                 ct.ReconstructionParams())
    assert ei.value.code == ct.CT_ERR_UNSUPPORTED


# ----------------------------------------------------------------- point cloud
def _serial_point_cloud(vol, thr, xc=None, yc=None, zc=None):
    """SerialBackend::extractPointCloud (SerialBackend.cpp:325-346) in numpy."""
    d, h, w = vol.shape
    z, y, x = np.nonzero(vol > thr)  # C order = z, y, x raster order
    X = (np.asarray(xc, np.float32)[x] if xc is not None else x.astype(np.float32))
    Y = (np.asarray(yc, np.float32)[y] if yc is not None else y.astype(np.float32))
    Z = (np.asarray(zc, np.float32)[z] if zc is not None else z.astype(np.float32))
    return np.stack([X, Y, Z, vol[z, y, x]], 1).astype(np.float32)


@pytest.mark.parametrize("shape,thr", [((33, 40, 29), 0.3), ((64, 64, 64), 0.5), ((5, 7, 3), 10.0)])
def test_point_cloud_order_and_values(backend, O, shape, thr):
    d, h, w = shape
    vol = O.phantom(31, 12, w, h, d)
    got = backend.extractPointCloud(vol, thr)
    ref = _serial_point_cloud(vol, thr)
    assert got.shape == ref.shape and _bits_equal(got, ref)
    xc = np.linspace(-1, 1, w).astype(np.float32)
    zc = np.arange(d, dtype=np.float32) * 2.5
    got = backend.extractPointCloud(vol, thr, x_coords=xc, z_coords=zc)
    assert _bits_equal(got, _serial_point_cloud(vol, thr, xc=xc, zc=zc))


# ----------------------------------------------------------------- C++ adapter
def test_cpp_hip_backend_through_the_interface(O, tmp_path):
    """HipBackend (C++) obtained from BackendFactory::createBestAvailable, called
    through IReconstructionBackend as the controller calls it."""
    import os
    import subprocess
    drv = os.path.join(os.path.dirname(ct.__file__), "hip_backend_driver")
    w = h = 64
    d, na, filt, thr = 6, 90, int(ct.FilterType.Hann), 0.4
    vol = O.phantom(404, 10, w, h, d)
    vol.tofile(tmp_path / "vol.f32")
    r = subprocess.run([drv, str(tmp_path / "vol.f32"), str(w), str(h), str(d), str(na), str(filt),
                        str(thr), str(tmp_path), "404", "10"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    meta = (tmp_path / "meta.txt").read_text().split("\n")
    # the phantom through IVolumeGenerator::generateBrainHU (HipVolumeGenerator):
    # the same voxels as the oracle's phantom of that seed and count, bit for
    # bit, with the reference's coordinate axes (linspace(-50, 50, n) mm)
    assert meta[16] == "GPU (HIP, MI355X)" and meta[17] == "1"
    assert meta[18].split() == [str(w), str(h), str(d)]
    x0, x1, nzc = meta[19].split()
    assert float(x0) == -50.0 and float(x1) == 50.0 and int(nzc) == d
    assert _bits_equal(np.fromfile(tmp_path / "phantom.f32", np.float32).reshape(d, h, w), vol)
    assert meta[0] == "HIP (MI355X)" and int(meta[1]) == d and meta[3] == "1"
    sino0 = np.fromfile(tmp_path / "sino0.f32", np.float32).reshape(na, w)
    ref_s, _ = O.compute_sinogram(vol[0], na, w)
    _sino_parity(O, sino0, ref_s)
    sl = np.fromfile(tmp_path / "slice0.f32", np.float32).reshape(w, w)
    assert O.rel_l2(sl, O.reconstruct_slice(ref_s, O.serial_angles(na), w, filt)) <= TOL
    v = np.fromfile(tmp_path / "volume.f32", np.float32).reshape(d, h, w)
    ref_v, _ = O.reconstruct_volume(vol, na, filt)
    assert O.rel_l2(v, ref_v) <= TOL
    fs = np.fromfile(tmp_path / "fromsino.f32", np.float32).reshape(d, w, w)
    for z in range(d):
        s, _ = O.compute_sinogram(vol[z], na, w)
        assert O.rel_l2(fs[z], O.reconstruct_slice(s, O.serial_angles(na), w, filt)) <= TOL
    pts = np.fromfile(tmp_path / "points.f32", np.float32).reshape(-1, 4)
    xc = 0.5 * np.arange(w, dtype=np.float32)
    assert _bits_equal(pts, _serial_point_cloud(vol, thr, xc=xc))
    # files: exported + reloaded stack reconstructs exactly as the in-memory one
    # (na = 90: the sidecar's angles equal the float product 2.0f * a)
    ff = np.fromfile(tmp_path / "fromfile.f32", np.float32).reshape(d, w, w)
    assert _bits_equal(ff, fs)
    assert meta[6] == "1" and meta[7] == "1"  # volume and point-cloud NPY round trips
    np.testing.assert_array_equal(np.load(tmp_path / "volume.npy"), v)
    # the interface's default reconstructVolumeFromSinograms body: one
    # reconstructSlice per z (the FFT then pairs each row with zeros instead of
    # the next slice's row, so the bits may differ by rounding)
    assert int(meta[8]) == d
    fd = np.fromfile(tmp_path / "fromsino_default.f32", np.float32).reshape(d, w, w)
    assert O.rel_l2(fd, fs) <= 1e-6
    for z in range(d):
        s, _ = O.compute_sinogram(vol[z], na, w)
        assert O.rel_l2(fd[z], O.reconstruct_slice(s, O.serial_angles(na), w, filt)) <= TOL
    m = O.metrics(vol, v)
    got = [float(x) for x in meta[9:16]]
    for k, g in zip(("mse", "mae", "psnr"), got[:3]):
        assert g == pytest.approx(m[k], rel=1e-9), k
    assert np.float32(got[3]) == m["max_abs"] and np.float32(got[4]) == m["vmin"]
    assert np.float32(got[5]) == m["vmax"] and int(got[6]) == vol.size
