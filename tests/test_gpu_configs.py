"""The four GPU workloads at their real sizes, checked against the oracle.

* 512^3 plain FP32 volume: a 24-ellipsoid phantom, 720 angles, 512 bins
  (sq 725, nfft 2048), Shepp-Logan, through reconstructVolume on the device.
  Phantom bit-exact over the whole volume; detector indexing bit-exact at
  every angle (slices with one nonzero pixel); sinograms and sampled slices
  within 1e-5 of Serial, both denormalised and before denormalisation (the
  consumer path on the same sinograms), and against the FP64 oracle.
* Precision: the full 368,640-row sinogram stack of that volume filtered by
  every row-filter backend (FFT, Toeplitz GEMM in BF16, split BF16, FP8), for
  Ram-Lak and Hann; sampled rows against the exact FP64 padded Toeplitz, and
  sampled reconstructed slices against the FP64 oracle, per variant.
* Batched: 16 volumes drawn from 128^3, 256^3, 384^2x192, 512^2x256, 768^2x128
  with 180-1440 angles and detector bins ceil(sqrt2 N)+2, reconstructed by one
  persistent launch pair (graph + autotuner): every volume bit-identical to
  its standalone consumer-path run, and a slice of each against Serial.
* Served: a job of every (edge 128-1024, angles 360/720/1440, filter
  Ram-Lak/Shepp-Logan/Hann) class, from noisy log sinograms (Poisson counts at
  I0 = 1e4), submitted at Poisson times to the continuous-batching server:
  each result bit-identical to its standalone run, a slice of every (edge,
  angles) class against Serial and against FP64.

Oracle calls on sampled slices run in a thread pool (the C oracle releases
the GIL) so the suite stays within a couple of minutes.
"""
import concurrent.futures as cf
import math
import time

import numpy as np
import pytest

import qtctreconstruction_amd as ct

pytestmark = pytest.mark.gpu

TOL = 1e-5        # relative L2 vs the FP32 Serial restatement
SINO_TOL = 2e-6   # sinograms: rounding of the split, fused sums only
FP64_TOL = 1e-4   # relative L2 vs the FP64 oracle
POOL = cf.ThreadPoolExecutor(max_workers=8)


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                          np.asarray(b, np.float32).view(np.uint32))


def _pmap(fn, items):
    return list(POOL.map(fn, items))


# =============================================================== 512^3 plain
N2, NA2, ELL2, SEED2 = 512, 720, 24, 20261203
SL = int(ct.FilterType.SheppLogan)
ZL = [0, 170, 256, 511]


@pytest.fixture(scope="module")
def c2():
    ctx = ct.Context(0)
    n = N2
    print(f"[configs] 512^3 volume: phantom -> reconstructVolume on the device", flush=True)
    d_vol = ctx.phantom_dev(SEED2, ELL2, n, n, n)
    d_out = ctx.alloc(4 * n ** 3)
    d_sino = ctx.alloc(4 * n * NA2 * n)
    ctx.reconstruct_volume_dev(d_vol, n, n, n, NA2, SL, d_out, d_sino)
    vol = d_vol.download((n, n, n))
    out = d_out.download((n, n, n))
    yield dict(ctx=ctx, vol=vol, out=out, d_vol=d_vol, d_out=d_out, d_sino=d_sino)
    ctx.close()


def test_c2_phantom_bit_exact(c2, O):
    ref = O.phantom(SEED2, ELL2, N2, N2, N2, threads=16)
    assert _bits_equal(c2["vol"], ref)


def test_c2_sinogram_indexing_bit_exact(c2, O):
    """One nonzero pixel per slice: every bin receives at most one deposit, so
    the sinogram is Serial's bit for bit exactly when every pixel's bin and
    fraction are, at all 720 angles (corners, edges, centre and random
    pixels; 8 slices = one z-block of the walk)."""
    ctx = c2["ctx"]
    n = N2
    rng = np.random.default_rng(9)
    pix = [(0, 0), (n - 1, n - 1), (0, n - 1), (n - 1, 0), (255, 256), (256, 255)]
    pix += [tuple(int(v) for v in rng.integers(0, n, 2)) for _ in range(2)]
    vol = np.zeros((8, n, n), np.float32)
    for z, (y, x) in enumerate(pix):
        vol[z, y, x] = np.float32(rng.uniform(0.1, 2.0))
    d_v = ctx.alloc(vol.nbytes).upload(vol)
    d_s = ctx.alloc(4 * 8 * NA2 * n)
    ctx.radon_dev(d_v, n, n, 8, NA2, n, d_s)
    got = d_s.download((8, NA2, n))
    for z in range(8):
        ref, _ = O.compute_sinogram(vol[z], NA2, n)
        assert _bits_equal(got[z], ref), pix[z]
    d_v.free()
    d_s.free()


def test_c2_volume_vs_serial_and_fp64(c2, O):
    """Sampled slices of the denormalised volume and, before denormalisation,
    the consumer path's reconstruction of the same device sinograms (the
    backprojection's accumulator before the fused epilogue)."""
    ctx, vol, out = c2["ctx"], c2["vol"], c2["out"]
    n = N2
    ref, ref_sino = O.reconstruct_volume(vol, NA2, SL, z_list=ZL, threads=16, want_sino=True)
    sino = c2["d_sino"].download((n, NA2, n))
    for k, z in enumerate(ZL):
        assert O.rel_l2(sino[z], ref_sino[k]) <= SINO_TOL, z
    err = O.rel_l2(out[ZL], ref[ZL])
    ref64 = O.reconstruct_volume64(vol, NA2, SL, z_list=ZL, threads=16)
    err64 = O.rel_l2(out[ZL], ref64[ZL])
    # before denormalisation: Serial's reconstructSlice output (normalised)
    d_pre = ctx.alloc(4 * n ** 3)
    ang = O.serial_angles(NA2)
    ctx.reconstruct_from_sinograms_dev(c2["d_sino"], n, NA2, n, ang, SL, d_pre)
    pre = d_pre.download((n, n, n))[ZL]
    d_pre.free()
    pre_ref = np.stack(_pmap(lambda k: O.reconstruct_slice(ref_sino[k], ang, n, SL),
                             range(len(ZL))))
    err_pre = O.rel_l2(pre, pre_ref)
    pre64 = np.stack(_pmap(lambda z: O.reconstruct_slice64(O.compute_sinogram64(vol[z], NA2, n),
                                                           ang.astype(np.float64), n, SL), ZL))
    err_pre64 = O.rel_l2(pre, pre64)
    print(f"512^3 x 720: relL2 vs Serial {err:.2e} (before denormalisation {err_pre:.2e}); "
          f"vs FP64 {err64:.2e} (before denormalisation {err_pre64:.2e})")
    assert err <= TOL and err_pre <= TOL
    assert err64 <= FP64_TOL and err_pre64 <= FP64_TOL


def test_c2_host_buffers_pipelined_bit_identical(c2, O):
    """reconstructVolume and reconstructVolumeFromSinograms on host buffers run
    the volume in z-slabs (32, 64, 128, 128, 64, 64, 32 slices here) through
    pinned staging on three streams; the result is the device-resident
    call's bit for bit."""
    be = ct.HipBackend(0)
    params = ct.ReconstructionParams(ct.FilterType.SheppLogan, NA2)
    t0 = time.perf_counter()
    out, sino = be.reconstructVolume(c2["vol"], params, want_sinograms=True)
    dt = time.perf_counter() - t0
    assert _bits_equal(out, c2["out"])
    ref_sino = c2["d_sino"].download((N2, NA2, N2))
    assert _bits_equal(sino, ref_sino)
    t = be.ctx.timings()
    print(f"host-buffer reconstructVolume 512^3 x 720: {dt * 1e3:.1f} ms wall "
          f"(stages {t[0]:.1f} + {t[1]:.1f} + {t[2]:.1f} ms on the device)")
    assert t[0] > 0 and t[1] > 0 and t[2] > 0
    ang = O.serial_angles(NA2)
    got = be.reconstructVolumeFromSinograms(ref_sino, ang, params)
    ctx = c2["ctx"]
    d_ref = ctx.alloc(4 * N2 ** 3)
    ctx.reconstruct_from_sinograms_dev(c2["d_sino"], N2, NA2, N2, ang, SL, d_ref)
    assert _bits_equal(got, d_ref.download((N2, N2, N2)))
    d_ref.free()
    be.ctx.close()


# =============================================================== precision
# filtered rows vs the exact FP64 padded Toeplitz (measured: FFT 8e-6,
# split BF16 4e-6, BF16 2.4e-3, FP8 2.9e-2) and reconstructed slices vs FP64
ROW_BAR = {ct.FilterBackend.FFT_F32: 2e-5, ct.FilterBackend.GEMM_SPLIT_BF16: 1e-5,
           ct.FilterBackend.GEMM_BF16: 6e-3, ct.FilterBackend.GEMM_FP8: 6e-2}
REC_BAR = {ct.FilterBackend.FFT_F32: 5e-5, ct.FilterBackend.GEMM_SPLIT_BF16: 2e-5,
           ct.FilterBackend.GEMM_BF16: 1e-2, ct.FilterBackend.GEMM_FP8: 1e-1}


@pytest.mark.parametrize("ftype", [int(ct.FilterType.Ramp), int(ct.FilterType.Hann)])
def test_c3_precision_full_stack(c2, O, ftype):
    ctx = c2["ctx"]
    n = N2
    sq, pad, nfft = O.geometry(n)
    sino = c2["d_sino"].download((n, NA2, n))
    rows = np.random.default_rng(ftype).choice(n * NA2, 4096, replace=False)
    P = sino.reshape(-1, n)[rows].astype(np.float64)
    h = O.kernel64(nfft, ftype)
    T = h[(np.arange(sq)[:, None] - (np.arange(n)[None, :] + pad)) % nfft]
    ref_rows = P @ T.T
    ang = O.serial_angles(NA2)
    zs = [100, 400]
    rec64 = np.stack(_pmap(lambda z: O.reconstruct_slice64(sino[z].astype(np.float64),
                                                           ang.astype(np.float64), n, ftype), zs))
    d_rec = ctx.alloc(4 * n ** 3)
    try:
        for fb in ct.FilterBackend:
            ctx.set_filter_backend(fb)
            got = ctx.filter_sinograms(sino, ftype)  # every one of the 368,640 rows
            assert got.shape == (n, NA2, sq)
            g = got.reshape(-1, sq)[rows]
            e_rows = O.rel_l2(g, ref_rows)
            m_rows = float(np.abs(g - ref_rows).max())
            del got
            ctx.reconstruct_from_sinograms_dev(c2["d_sino"], n, NA2, n, ang, ftype, d_rec)
            rec = d_rec.download((n, n, n))[zs]
            e_rec = O.rel_l2(rec, rec64)
            m_rec = float(np.abs(rec - rec64).max())
            print(f"filter {ftype} {fb.name}: rows relL2 {e_rows:.2e} max {m_rows:.2e}; "
                  f"slices relL2 {e_rec:.2e} max {m_rec:.2e} (vs FP64)")
            assert e_rows <= ROW_BAR[fb], (fb.name, e_rows)
            assert e_rec <= REC_BAR[fb], (fb.name, e_rec)
    finally:
        ctx.set_filter_backend(ct.FilterBackend.FFT_F32)
        d_rec.free()


# =============================================================== batched
def _batch_specs(seed=4):
    """16 volumes as the batched benchmark draws them: (N, nz, na, filter,
    ellipsoids); detector bins B = ceil(sqrt2 N) + 2."""
    rng = np.random.default_rng(seed)
    shapes = [(128, 128), (256, 256), (384, 192), (512, 256), (768, 128)]
    out = []
    for _ in range(16):
        N, nz = shapes[rng.integers(len(shapes))]
        na = int(rng.choice([180, 360, 720, 1440]))
        f = int(rng.choice([0, 1, 4]))
        out.append((N, nz, na, f, int(rng.integers(10, 25))))
    return out


def batch_bins(N: int) -> int:
    return int(math.ceil(math.sqrt(2.0) * N)) + 2


def test_c4_batched_16_volumes(O):
    ctx = ct.Context(0)
    try:
        jobs, meta = [], []
        for k, (N, nz, na, f, ne) in enumerate(_batch_specs()):
            B = batch_bins(N)
            d_v = ctx.phantom_dev(1000 + k, ne, N, N, nz)
            d_s = ctx.alloc(4 * nz * na * B)
            ctx.radon_dev(d_v, N, N, nz, na, B, d_s)
            d_v.free()
            d_o = ctx.alloc(4 * B * B * nz)
            ang = O.serial_angles(na)
            jobs.append((d_s, d_o, ang, nz, na, B, f))
            meta.append((N, nz, na, f, B))
        print(f"[configs] batched: {[(m[0], m[1], m[2], m[4]) for m in meta]}", flush=True)
        ctx.fbp_batch(jobs, ct.CT_BATCH_GRAPH | ct.CT_BATCH_AUTOTUNE)
        ctx.fbp_batch(jobs, ct.CT_BATCH_GRAPH | ct.CT_BATCH_AUTOTUNE)  # graph replay
        checks = []
        for k, (d_s, d_o, ang, nz, na, B, f) in enumerate(jobs):
            got = d_o.download((nz, B, B))
            d_a = ctx.alloc(4 * B * B * nz)
            ctx.reconstruct_from_sinograms_dev(d_s, nz, na, B, ang, f, d_a)
            alone = d_a.download((nz, B, B))
            d_a.free()
            assert _bits_equal(got, alone), meta[k]
            z = nz // 2
            s = d_s.download((nz, na, B))[z]
            checks.append((k, s, got[z].copy()))
            del got, alone
        errs = _pmap(lambda c: O.rel_l2(c[2], O.reconstruct_slice(c[1], jobs[c[0]][2],
                                                                  jobs[c[0]][5], jobs[c[0]][6])),
                     checks)
        print("batched: relL2 vs Serial per volume " + " ".join(f"{e:.1e}" for e in errs))
        assert max(errs) <= TOL
    finally:
        ctx.close()


# =============================================================== served
EDGES = [128, 256, 512, 768, 1024]
ANGLES = [360, 720, 1440]
FILTERS = [int(ct.FilterType.Ramp), int(ct.FilterType.SheppLogan), int(ct.FilterType.Hann)]


def test_c5_served_every_class(O):
    ctx = ct.Context(0)
    srv = None
    try:
        rng = np.random.default_rng(55)
        depth = 40
        pool = {}
        for N in EDGES:
            for na in ANGLES:
                d_v = ctx.phantom_dev(7000 + N + na, 24, N, N, depth)
                d_s = ctx.alloc(4 * depth * na * N)
                ctx.radon_dev(d_v, N, N, depth, na, N, d_s)
                d_v.free()
                ctx.noisy_log_dev(d_s, depth * na * N, 1e4, 2.0 / N, seed=N * 10007 + na)
                pool[(N, na)] = d_s
        jobs = [(N, int(rng.integers(8, depth + 1)), na, f)
                for N in EDGES for na in ANGLES for f in FILTERS]
        order = rng.permutation(len(jobs))
        print(f"[configs] served: {len(jobs)} jobs, depths {sorted(j[1] for j in jobs)}", flush=True)
        srv = ct.Server(pool_bytes=24 << 30, max_batch=16, slots=2)
        arrivals = np.cumsum(rng.exponential(1.0 / 400.0, len(jobs)))  # ~400 jobs/s
        tickets = {}
        t0 = time.perf_counter()
        results = {}
        k = 0
        while len(results) < len(jobs):
            now = time.perf_counter() - t0
            while k < len(jobs) and arrivals[k] <= now:
                N, d, na, f = jobs[order[k]]
                t = srv.submit(pool[(N, na)], O.serial_angles(na), d, na, N, f)
                tickets[t] = order[k]
                k += 1
            r = srv.wait(0.5 if k < len(jobs) else 60000)
            if r is None:
                assert k < len(jobs), "server stalled"
                continue
            j = tickets[r[0]]
            N, d, na, f = jobs[j]
            results[j] = srv.download(r[0], (d, N, N))
            srv.release(r[0])
        st = srv.stats()
        print(f"served stats: {st}")
        assert st["completed"] == len(jobs)
        # each result equals its standalone consumer-path run bit for bit
        for j, (N, d, na, f) in enumerate(jobs):
            d_a = ctx.alloc(4 * d * N * N)
            ctx.reconstruct_from_sinograms_dev(pool[(N, na)], d, na, N, O.serial_angles(na), f, d_a)
            alone = d_a.download((d, N, N))
            d_a.free()
            assert _bits_equal(results[j], alone), jobs[j]
        # a slice of every (edge, angles) class against Serial and FP64
        picks = {}
        for j, (N, d, na, f) in enumerate(jobs):
            picks.setdefault((N, na), j)
        checks = []
        for (N, na), j in picks.items():
            _, d, _, f = jobs[j]
            z = d // 2
            s = pool[(N, na)].download((depth, na, N))[z]
            checks.append((j, s, results[j][z]))

        def errs(c):
            j, s, got = c
            N, d, na, f = jobs[j]
            ang = O.serial_angles(na)
            e = O.rel_l2(got, O.reconstruct_slice(s, ang, N, f))
            e64 = O.rel_l2(got, O.reconstruct_slice64(s.astype(np.float64),
                                                      ang.astype(np.float64), N, f))
            return e, e64

        ee = _pmap(errs, checks)
        for c, (e, e64) in zip(checks, ee):
            print(f"served {jobs[c[0]]}: relL2 vs Serial {e:.1e}, vs FP64 {e64:.1e}")
        assert max(e for e, _ in ee) <= TOL
        assert max(e64 for _, e64 in ee) <= FP64_TOL
        for d_s in pool.values():
            d_s.free()
    finally:
        if srv is not None:
            srv.close()
        ctx.close()
