"""Batched form: uneven volumes in one persistent filter launch + one
persistent backprojection launch, captured in a hipGraph and replayed.

Each volume's angle sum runs in the same order whatever the tile
configuration or batch composition, so a volume reconstructed in a batch is
bit-identical to the same volume reconstructed alone.
"""
import numpy as np
import pytest

import qtctreconstruction_amd as ct

pytestmark = pytest.mark.gpu

JOBS = [(40, 5, 36, ct.FilterType.Ramp), (64, 3, 90, ct.FilterType.Hann),
        (97, 4, 45, ct.FilterType.SheppLogan), (32, 9, 180, ct.FilterType.Cosine),
        (128, 2, 60, ct.FilterType.Bartlett)]


def _make_sinos(O, seed, N, nz, na):
    vol = O.phantom(seed, 8, N, N, nz)
    return np.stack([O.compute_sinogram(vol[z], na, N)[0] for z in range(nz)])


def _alone(be, sinos, ang, f):
    return be.reconstructVolumeFromSinograms(sinos, ang, ct.ReconstructionParams(f, len(ang)))


@pytest.mark.parametrize("flags", [ct.CT_BATCH_GRAPH | ct.CT_BATCH_AUTOTUNE, ct.CT_BATCH_GRAPH, 0])
def test_batch_equals_alone(O, flags):
    be = ct.HipBackend(0)
    ctx = be.ctx
    rng = np.random.default_rng(flags)
    bufs, refs = [], []
    for k, (N, nz, na, f) in enumerate(JOBS):
        sinos = _make_sinos(O, 100 + k, N, nz, na)
        ang = O.serial_angles(na) if k % 2 == 0 else np.sort(rng.uniform(0, 180, na)).astype(np.float32)
        ds = ctx.alloc(sinos.nbytes).upload(sinos)
        do = ctx.alloc(4 * nz * N * N)
        bufs.append((ds, do, ang, nz, na, N, f))
        refs.append((_alone(be, sinos, ang, f), sinos, ang))
    for rep in range(3):  # replays of the captured graph, eager work between them
        ctx.fbp_batch(bufs, flags)
        for (ds, do, ang, nz, na, N, f), (ref, sinos, _) in zip(bufs, refs):
            got = do.download((nz, N, N))
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (rep, N, nz, na)
        _ = _alone(be, refs[0][1], refs[0][2], JOBS[0][3])
    # and against the oracle
    for (ref, sinos, ang), (N, nz, na, f) in zip(refs, JOBS):
        for z in range(nz):
            r = O.reconstruct_slice(sinos[z], ang, N, int(f))
            assert O.rel_l2(ref[z], r) <= 1e-5


def test_batch_changes_between_replays(O):
    """A second batch with other inputs (same shapes) reuses the graph."""
    be = ct.HipBackend(0)
    ctx = be.ctx
    N, nz, na, f = 48, 4, 30, ct.FilterType.Ramp
    ang = O.serial_angles(na)
    ds = ctx.alloc(4 * nz * na * N)
    do = ctx.alloc(4 * nz * N * N)
    for seed in (1, 2, 3):
        sinos = _make_sinos(O, seed, N, nz, na)
        ds.upload(sinos)
        ctx.fbp_batch([(ds, do, ang, nz, na, N, f)], ct.CT_BATCH_GRAPH)
        got = do.download((nz, N, N))
        ref = _alone(be, sinos, ang, f)
        assert np.array_equal(got, ref)


def test_tile_configs_bit_identical(O):
    be = ct.HipBackend(0)
    vol = O.phantom(8, 10, 80, 80, 6)
    params = ct.ReconstructionParams(ct.FilterType.SheppLogan, 72)
    outs = []
    for cfg in range(len(be.ctx.bp_configs())):
        be.ctx.set_bp_config(cfg)
        outs.append(be.reconstructVolume(vol, params))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


@pytest.mark.parametrize("bins,nz", [(256, 9), (300, 6)])
def test_tiles_outside_circle(O, bins, nz):
    """Slices of 256 and 300 voxels: 4 and 12 of the 32-voxel tiles (32 and
    50 of the 16-voxel ones) lie wholly outside the reconstruction circle and
    skip the angle sums. Random sinograms give every voxel inside the circle
    a value of the slice's scale, so a skipped voxel inside it would fail the
    per-voxel bound; voxels outside are exactly 0; every tile configuration
    writes the same volume."""
    be = ct.HipBackend(0)
    rng = np.random.default_rng(bins)
    na = 24
    sinos = rng.standard_normal((nz, na, bins)).astype(np.float32)
    ang = O.serial_angles(na)
    params = ct.ReconstructionParams(ct.FilterType.Hann, na)
    outs = []
    for cfg in range(len(be.ctx.bp_configs())):
        be.ctx.set_bp_config(cfg)
        outs.append(be.reconstructVolumeFromSinograms(sinos, ang, params))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    got = outs[0]
    R = bins // 2
    yy, xx = np.mgrid[0:bins, 0:bins].astype(np.float32) - np.float32(R)
    outside = xx * xx + yy * yy > np.float32(R * R)
    assert np.all(got[:, outside] == 0.0)
    for z in range(nz):
        ref = O.reconstruct_slice(sinos[z], ang, bins, O.HANN)
        assert O.rel_l2(got[z], ref) <= 1e-5
        assert np.max(np.abs(got[z] - ref)) <= 1e-4 * np.max(np.abs(ref)), z


def test_autotune_returns_a_timed_config():
    ctx = ct.Context(0)
    best, ms = ctx.autotune(256, 180)
    assert 0 <= best < len(ctx.bp_configs())
    assert all(t > 0 for t in ms) and ms[best] == min(ms)


def test_graph_replay_with_changing_class_composition(O):
    """Batches with the same FFT-length classes but different numbers of jobs
    per class replay one captured graph; each volume must still equal its
    standalone reconstruction (the per-class work-list offsets live in device
    memory, not in the captured kernel arguments)."""
    be = ct.HipBackend(0)
    ctx = be.ctx
    small = [(40, 3, 24), (48, 2, 30), (44, 4, 18)]   # LDS FFT of 128 points
    large = [(96, 2, 36), (100, 3, 20)]               # wave form, one 512-point transform
    pool = {}
    for k, (N, nz, na) in enumerate(small + large):
        s = _make_sinos(O, 300 + k, N, nz, na)
        pool[(N, nz, na)] = (s, ctx.alloc(s.nbytes).upload(s), ctx.alloc(4 * nz * N * N))
    compositions = [small[:1] + large, small + large[:1], small[:2] + large, small + large]
    for comp in compositions:
        jobs = []
        for (N, nz, na) in comp:
            s, ds, do = pool[(N, nz, na)]
            jobs.append((ds, do, O.serial_angles(na), nz, na, N, ct.FilterType.Ramp))
        ctx.fbp_batch(jobs, ct.CT_BATCH_GRAPH)
        for (N, nz, na) in comp:
            s, ds, do = pool[(N, nz, na)]
            ref = _alone(be, s, O.serial_angles(na), ct.FilterType.Ramp)
            assert np.array_equal(do.download((nz, N, N)), ref), (comp, N)


def test_batch_every_filter_class(O):
    """One batch holding every row filter class: the LDS FFT below 91 bins
    and past 512, and the wave form with 1, 2 and 3 sub-transforms of 512
    points (one persistent launch per class). Each volume equals its
    standalone reconstruction bit for bit, and its first slice matches the
    Serial oracle."""
    be = ct.HipBackend(0)
    ctx = be.ctx
    shapes = [(40, 3, 12), (96, 2, 10), (300, 2, 8), (512, 1, 6), (600, 2, 5)]
    jobs, keep = [], []
    for k, (N, nz, na) in enumerate(shapes):
        s = _make_sinos(O, 500 + k, N, nz, na)
        ds, do = ctx.alloc(s.nbytes).upload(s), ctx.alloc(4 * nz * N * N)
        jobs.append((ds, do, O.serial_angles(na), nz, na, N, ct.FilterType.Hann))
        keep.append((s, do, N, nz, na))
    for flags in (0, ct.CT_BATCH_GRAPH):
        ctx.fbp_batch(jobs, flags)
        for (s, do, N, nz, na) in keep:
            got = do.download((nz, N, N))
            ref = _alone(be, s, O.serial_angles(na), ct.FilterType.Hann)
            assert np.array_equal(got, ref), (flags, N)
            ser = O.reconstruct_slice(s[0], O.serial_angles(na), N, O.HANN)
            assert O.rel_l2(got[0], ser) <= 1e-5, (flags, N)


@pytest.mark.parametrize("fb", [ct.FilterBackend.GEMM_SPLIT_BF16, ct.FilterBackend.GEMM_BF16,
                                ct.FilterBackend.GEMM_FP8])
def test_batch_with_gemm_filter_equals_alone(O, fb):
    """With a Toeplitz-GEMM filter backend set on the context, the batch filters
    each volume with one MFMA launch ahead of the persistent backprojection:
    every volume equals its standalone reconstruction on the same context bit
    for bit, with and without the graph, and the split form stays within 1e-4
    of the FP64 oracle."""
    be = ct.HipBackend(0)
    ctx = be.ctx
    ctx.set_filter_backend(fb)
    shapes = [(40, 3, 24, ct.FilterType.Ramp), (97, 5, 30, ct.FilterType.Hann),
              (300, 2, 12, ct.FilterType.SheppLogan), (64, 6, 90, ct.FilterType.Hann)]
    jobs, keep = [], []
    for k, (N, nz, na, f) in enumerate(shapes):
        s = _make_sinos(O, 700 + k, N, nz, na)
        ds, do = ctx.alloc(s.nbytes).upload(s), ctx.alloc(4 * nz * N * N)
        jobs.append((ds, do, O.serial_angles(na), nz, na, N, f))
        keep.append((s, do, N, nz, na, f))
    for flags in (0, ct.CT_BATCH_GRAPH, ct.CT_BATCH_GRAPH):
        ctx.fbp_batch(jobs, flags)
        for (s, do, N, nz, na, f) in keep:
            got = do.download((nz, N, N))
            ref = _alone(be, s, O.serial_angles(na), f)
            assert np.array_equal(got, ref), (fb, flags, N)
            if fb == ct.FilterBackend.GEMM_SPLIT_BF16:
                ref64 = O.reconstruct_slice64(s[0], O.serial_angles(na), N, int(f))
                err = O.rel_l2(got[0], ref64)
                assert err <= 1e-4, (N, err)


def test_tune_cache_file_is_read_by_every_context(tmp_path, monkeypatch):
    """CT_HIP_TUNE_CACHE entries apply to each new context, not only the
    first one of the process."""
    import ctypes
    cache = tmp_path / "tune.txt"
    cache.write_text("48 30 6\n")
    monkeypatch.setenv("CT_HIP_TUNE_CACHE", str(cache))
    for _ in range(2):
        ctx = ct.Context(0)
        best = ctypes.c_int(-1)
        ctx._call("ct_hip_autotune", ctx.handle, 48, 30, ctypes.byref(best), None, 0)
        assert best.value == 6
        ctx.close()
