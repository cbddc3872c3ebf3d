"""Benchmark of the FBP hot path on one MI355X.

A step is one reconstructVolume over device-resident input: per-slice
min/max + normalisation, ray-driven Radon (720 angles x 512 bins; the
LDS-staged row walk over the image and over its transpose, each timed with
HIP events), the
Shepp-Logan FFT row filter (sq 725, nfft 2048) and voxel-driven
backprojection with denormalisation, for a 512^3 phantom of 24 seeded
ellipsoids (the plain-FP32 single-GPU configuration). value = voxel-updates/s
= 512^3 * 720 / step time.

Prints one JSON line. Timed region: K steps between two device
synchronisations; stage times come from HIP events on the context's stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FBP voxel-updates/s + p50/p99 latency, served vs plain HIP, relL2≤1e-4 vs FP64"
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 (vector and matrix alike)
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
FLOP_PER_UPDATE = 9        # backprojection: t (3), u (1), frac (1), lerp (3), accumulate (1)
# Radon, per (pixel, angle) deposit of one slice: the same count as an update
# (s (3+1), frac (1), the two weights and products (3), the accumulations (1));
# SerialBackend.cpp:110-120 spends it on every pixel of every slice.
FLOP_PER_SAMPLE = 9
# of which the deposit arithmetic alone (two products, two sums): the rest is
# geometry, which the walk computes once per pixel and angle for 8 slices
FLOP_PER_SAMPLE_DEPOSIT = 4
# HBM bytes per launch of each kernel at the 512^3 x 720 workload, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over one step, each scaled
# by the counter's reading for a kernel of known traffic (minmax_kernel reads
# 4 n^3 bytes, phantom_kernel with no ellipsoids writes 4 n^3 bytes); the
# profile summaries list the raw counts and the command.
TRAFFIC_BYTES = {
    "radon_rows_kernel": 9.644e+09,
    "bp_kernel": 6.459e+09,
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(vol_host: np.ndarray, na: int, ftype: int, target_s: float):
    """OpenMPBackend::reconstructVolume restatement on a z-subset of the same volume."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    n = vol_host.shape[1]
    nz = 4
    t0 = time.perf_counter()
    O.omp_reconstruct_volume(vol_host[:nz], na, ftype, threads)
    dt = time.perf_counter() - t0
    # grow the sample towards target_s of CPU work (bounded by the volume)
    nz2 = int(min(vol_host.shape[0], max(nz, nz * target_s / max(dt, 1e-3))))
    nz2 -= nz2 % 2
    if nz2 > nz:
        t0 = time.perf_counter()
        _, ph = O.omp_reconstruct_volume(vol_host[:nz2], na, ftype, threads)
        dt = time.perf_counter() - t0
        nz = nz2
    upd = float(n) * n * nz * na
    return {"value": upd / dt, "unit": "voxel-updates/s", "cores": threads, "kind": "port",
            "sample": f"OpenMPBackend::reconstructVolume restatement ({O.ompbase().isa}), "
                      f"{n}x{n}x{nz} slices of the same phantom, {na} angles, "
                      f"{dt:.2f} s wall"}


def dominant_roofline(n, na, stage, radon_k, launches):
    """Roofline entry of the kernel with the largest time per step.

    Candidates: the Radon row walk (radon_rows_kernel: one launch over the
    image for the angles with |sin| <= |cos|, one over the transposed image
    for the others), the backprojection (bp_kernel) and the FFT row filter.
    achieved = algorithmic FLOP per step / the kernel's time per step (HIP
    events on the context stream), over its launches per step.
    """
    samples = float(n) * n * n  # pixels x slices per angle
    cands = {
        "radon_rows_kernel": (radon_k[0] + radon_k[1], 2 * max(launches, 1),
                              FLOP_PER_SAMPLE * samples * na,
                              "Radon LDS row walk (image and transposed image)"),
        "bp_kernel": (stage[2], 1, FLOP_PER_UPDATE * samples * na,
                      "voxel-driven backprojection"),
    }
    name = max(cands, key=lambda k: cands[k][0])
    t_ms, nl, flop, what = cands[name]
    per_launch_ms = t_ms / nl
    achieved = flop / (t_ms / 1e3) / 1e12
    tb = TRAFFIC_BYTES.get(name)
    out = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
           "frac": achieved / PEAK_FP32_TFLOPS, "traffic": tb, "kernel": name,
           "launches_per_step": nl, "avg_launch_ms": per_launch_ms,
           "note": f"{what}; FP32 VALU class (no MFMA work on this path); "
                   f"{'9 FLOP per pixel-angle-slice deposit' if name.startswith('radon') else '9 FLOP per voxel-update'}"
                   "; 157.3 TF is MI355X's FP32 peak (packed FP32 counted), the kernels are "
                   "built without packed FP32 ops"}
    if name.startswith("radon"):
        dep = FLOP_PER_SAMPLE_DEPOSIT * samples * na / (t_ms / 1e3) / 1e12
        out["deposit_arithmetic_only"] = {"flop_per_sample": FLOP_PER_SAMPLE_DEPOSIT,
                                          "achieved": dep, "frac": dep / PEAK_FP32_TFLOPS}
    return out


def precision_form(ctx, d_sino, n, na, sample_z, host_sinos_sample):
    """Row filter variants on the full C2 sinogram stack: filter-stage time and
    error vs the exact FP64 padded-Toeplitz filter on sampled slices."""
    import qtctreconstruction_amd as ct
    from oracle import oracle as O
    sq, pad, nfft = O.geometry(n)
    res = {}
    d_out = ctx.alloc(4 * n * n * n)
    ang = O.serial_angles(na)
    for fname, ftype in (("Ramp", 0), ("Hann", 4)):
        h = O.kernel64(nfft, ftype)
        i = np.arange(sq)[:, None]
        j = np.arange(n)[None, :] + pad
        T = h[(i - j) % nfft]
        ref = (host_sinos_sample.astype(np.float64).reshape(-1, n) @ T.T)
        for fb in ct.FilterBackend:
            ctx.set_filter_backend(fb)
            # whole-stack timing through the consumer path (filter + backprojection)
            for _ in range(2):
                ctx._call("ct_hip_reconstruct_from_sinograms_dev", ctx.handle, d_sino.ptr, n, na, n,
                          ang.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_float)),
                          ftype, d_out.ptr)
            t = ctx.timings()
            got = ctx.filter_sinograms(host_sinos_sample, ftype).reshape(ref.shape)
            res[f"{fname}/{fb.name}"] = {"filter_ms": float(t[1]), "backprojection_ms": float(t[2]),
                                         "rel_l2_vs_fp64": O.rel_l2(got, ref),
                                         "max_abs_vs_fp64": float(np.abs(got - ref).max())}
            log(f"[bench] precision {fname}/{fb.name}: {res[f'{fname}/{fb.name}']}")
    ctx.set_filter_backend(ct.FilterBackend.FFT_F32)
    d_out.free()
    flops = 2.0 * n * na * n * sq
    res["gemm_flop"] = flops
    res["rows"] = n * na
    return res


def metrics_form(ctx, d_vol, d_out, nvox, vol_host, reps=10):
    """The controller's reconstruction metrics (MSE/MAE/max/PSNR) of the
    reconstruction against its phantom, both resident on the device: one
    HBM-bound pass over 8 bytes per voxel. Wall time per call includes the
    32-byte result download and the stream synchronisation."""
    for _ in range(2):
        m = ctx.volume_metrics_dev(d_vol, d_out, nvox)
    t = time.perf_counter()
    for _ in range(reps):
        m = ctx.volume_metrics_dev(d_vol, d_out, nvox)
    ms = (time.perf_counter() - t) * 1e3 / reps
    res = {"ms_per_call": ms, "gb_per_s": 8.0 * nvox / (ms / 1e3) / 1e9,
           "mse": m["mse"], "mae": m["mae"], "psnr": m["psnr"], "max_abs": m["max_abs"]}
    if vol_host is not None:
        from oracle import oracle as O
        n = round(nvox ** (1 / 3))
        ref = O.metrics(vol_host, d_out.download((n, n, n)))
        res["rel_err_vs_oracle"] = max(abs(m[k] - ref[k]) / abs(ref[k]) for k in ("mse", "mae", "psnr"))
        res["extrema_bit_exact"] = bool(all(np.float32(m[k]) == ref[k]
                                            for k in ("max_abs", "vmin", "vmax")))
    log(f"[bench] metrics {res}")
    return res


def make_volume_sinos(ctx, seed, N, nz, na, n_ell=16):
    """Device sinograms (nz, na, N) of a seeded phantom, Serial semantics."""
    d_vol = ctx.alloc(4 * N * N * nz)
    ctx.phantom_dev(seed, n_ell, N, N, nz, d_vol)
    d_s = ctx.alloc(4 * nz * na * N)
    ctx.radon_dev(d_vol, N, N, nz, na, N, d_s)
    d_vol.free()
    return d_s


def batched_form(ctx, seed=4):
    """16 uneven volumes, one persistent filter + one persistent backprojection
    launch (hipGraph, autotuned tiles) vs the same volumes one by one."""
    import qtctreconstruction_amd as ct
    rng = np.random.default_rng(seed)
    shapes = [(128, 128), (256, 256), (384, 192), (512, 256), (768, 128)]
    jobs = []
    for k in range(16):
        N, nz = shapes[rng.integers(len(shapes))]
        na = int(rng.choice([180, 360, 720, 1440]))
        f = int(rng.choice([0, 1, 4]))
        d_s = make_volume_sinos(ctx, 1000 + k, N, nz, na, int(rng.integers(10, 25)))
        d_o = ctx.alloc(4 * N * N * nz)
        step = np.float32(180.0) / np.float32(na)
        ang = (step * np.arange(na, dtype=np.float32)).astype(np.float32)
        jobs.append((d_s, d_o, ang, nz, na, N, f))
    updates = float(sum(N * N * nz * na for (_, _, _, nz, na, N, _) in jobs))
    flags = ct.CT_BATCH_GRAPH | ct.CT_BATCH_AUTOTUNE
    ctx.fbp_batch(jobs, flags)  # tunes, captures
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.fbp_batch(jobs, flags)
    t_batch = (time.perf_counter() - t0) / reps
    import ctypes

    def one_by_one():
        for (d_s, d_o, ang, nz, na, N, f) in jobs:
            ctx._call("ct_hip_reconstruct_from_sinograms_dev", ctx.handle, d_s.ptr, nz, na, N,
                      ang.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), f, d_o.ptr)
    one_by_one()  # first use tunes each (N, na) class; not timed
    t0 = time.perf_counter()
    one_by_one()
    t_seq = time.perf_counter() - t0
    for (d_s, d_o, *_r) in jobs:
        d_s.free()
        d_o.free()
    return {"volumes": [[N, N, nz, na] for (_, _, _, nz, na, N, _) in jobs],
            "voxel_updates": updates, "batched_ms": t_batch * 1e3,
            "batched_voxel_updates_per_s": updates / t_batch,
            "one_by_one_ms": t_seq * 1e3, "one_by_one_voxel_updates_per_s": updates / t_seq}


def served_form(ctx, n_jobs, seed=5, pool_gb=96.0, rate_factor=0.7):
    """Poisson stream of jobs over a device-resident input pool; p50/p99
    latency and throughput, beside plain per-job HIP FBP on the same mix."""
    import ctypes

    import qtctreconstruction_amd as ct
    rng = np.random.default_rng(seed)
    edges = [128, 256, 512, 768, 1024]
    angle_sets = [360, 720, 1440]
    I0 = 1e4
    # one noisy sinogram stack of depth 512 per (edge, angles) class; a job of
    # depth d uses its first d slices
    pool = {}
    log("[bench] serving: generating the input pool")
    for N in edges:
        for na in angle_sets:
            d_s = make_volume_sinos(ctx, 7000 + N + na, N, 512, na, 24)
            ctx.noisy_log_dev(d_s, 512 * na * N, I0, 2.0 / N, seed=N * 10007 + na)
            pool[(N, na)] = d_s
    angles = {na: ((np.float32(180.0) / np.float32(na)) * np.arange(na, dtype=np.float32)).astype(np.float32)
              for na in angle_sets}
    jobs = [(int(rng.choice(edges)), int(rng.integers(64, 513)), int(rng.choice(angle_sets)),
             int(rng.choice([0, 1, 4]))) for _ in range(n_jobs)]
    upd = np.array([N * N * d * na for (N, d, na, f) in jobs], np.float64)
    # plain per-job HIP FBP on a sample of the mix -> jobs/s capacity, rate
    d_out = ctx.alloc(4 * 1024 * 1024 * 512)
    sample = jobs[: min(40, n_jobs)]
    for N in edges:  # tune every class once before timing
        for na in angle_sets:
            ctx._call("ct_hip_reconstruct_from_sinograms_dev", ctx.handle, pool[(N, na)].ptr, 8, na,
                      N, angles[na].ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 0, d_out.ptr)
    t0 = time.perf_counter()
    for (N, d, na, f) in sample:
        ctx._call("ct_hip_reconstruct_from_sinograms_dev", ctx.handle, pool[(N, na)].ptr, d, na, N,
                  angles[na].ctypes.data_as(ctypes.POINTER(ctypes.c_float)), f, d_out.ptr)
    t_plain = time.perf_counter() - t0
    plain_ups = float(upd[: len(sample)].sum()) / t_plain
    d_out.free()
    # The first 12 of those jobs through the host-buffer entry, as a caller of
    # reconstructVolumeFromSinograms hands them over (pageable host sinograms
    # in, host volume out): the plain per-job rate with its PCIe transfers.
    f32p = ctypes.POINTER(ctypes.c_float)
    hsample = sample[:12]
    t_host = 0.0
    for (N, d, na, f) in hsample:
        hs = pool[(N, na)].download((d, na, N))
        ho = np.empty((d, N, N), np.float32)
        t0 = time.perf_counter()
        ctx._call("ct_hip_reconstruct_from_sinograms", ctx.handle, hs.ctypes.data_as(f32p), d, na,
                  N, angles[na].ctypes.data_as(f32p), f, ho.ctypes.data_as(f32p))
        t_host += time.perf_counter() - t0
        del hs, ho
    host_ups = float(sum(N * N * d * na for (N, d, na, f) in hsample)) / t_host
    log(f"[bench] serving: plain through host buffers {host_ups:.3e} upd/s "
        f"({len(hsample)} jobs, PCIe included)")
    mean_job_s = t_plain / len(sample)
    lam = rate_factor / mean_job_s
    log(f"[bench] serving: plain {plain_ups:.3e} upd/s, {1/mean_job_s:.1f} jobs/s -> rate {lam:.1f}/s")

    poisson_arrivals = np.cumsum(rng.exponential(1.0 / lam, n_jobs))

    def run(open_loop, policy=ct.SERVE_SIZE_AGED, cap_ms=5.0):
        srv = ct.Server(pool_bytes=int(pool_gb * (1 << 30)), max_batch=32, slots=2)
        srv.set_policy(policy, batch_cap_ms=cap_ms)
        arrivals = poisson_arrivals if open_loop else np.zeros(n_jobs)
        lat = []
        done = 0
        start = time.perf_counter()
        k = 0
        while done < n_jobs:
            now = time.perf_counter() - start
            while k < n_jobs and arrivals[k] <= now:
                N, d, na, f = jobs[k]
                srv.submit(pool[(N, na)], angles[na], d, na, N, f)
                k += 1
            wait_ms = 0.2 if k < n_jobs else 1000.0
            r = srv.wait(wait_ms)
            while r is not None:
                lat.append(r[1])
                srv.release(r[0])
                done += 1
                r = srv.wait(0.0) if done < n_jobs else None
        span = time.perf_counter() - start
        st = srv.stats()
        srv.close()
        lat = np.array(lat)
        return {"jobs": n_jobs, "span_s": span, "voxel_updates_per_s": float(upd.sum()) / span,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "batches": st["batches"], "max_jobs_in_flight": st["max_inflight_jobs"],
                "overlapped_launches": st["overlapped_launches"]}

    log("[bench] serving: closed loop (all jobs queued at t=0)")
    saturated = run(False)
    log(f"[bench] serving saturated: {saturated}")
    log("[bench] serving: Poisson arrivals")
    poisson = run(True)
    log(f"[bench] serving poisson: {poisson}")
    log("[bench] serving: the same Poisson arrivals, FIFO admission, no batch cap")
    poisson_fifo = run(True, ct.SERVE_FIFO, 0.0)
    log(f"[bench] serving poisson fifo: {poisson_fifo}")
    for d_s in pool.values():
        d_s.free()
    return {"jobs": n_jobs, "rate_per_s": lam, "rate_factor_of_plain": rate_factor,
            "plain_per_job_host_buffers_voxel_updates_per_s": host_ups,
            "poisson_vs_plain_device": poisson["voxel_updates_per_s"] / plain_ups,
            "poisson_vs_plain_host_buffers": poisson["voxel_updates_per_s"] / host_ups,
            "plain_host_buffers_note": f"first {len(hsample)} jobs of the mix through "
                                       "ct_hip_reconstruct_from_sinograms (pageable host "
                                       "sinograms in, host volume out, PCIe included)",
            "plain_per_job_voxel_updates_per_s": plain_ups, "saturated": saturated,
            "poisson": poisson, "poisson_fifo": poisson_fifo,
            "policy": "size-ordered with waiting credit 0.25, batch cap max(5 ms, queue/4); "
                      "poisson_fifo: FIFO without a cap on the same arrivals",
            "latency": "submit -> completion in device memory",
            "noise": f"counts ~ Poisson(1e4 exp(-kappa p)) (normal approx.), kappa = 2/N, log"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--angles", type=int, default=720)
    ap.add_argument("--ellipsoids", type=int, default=24)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-target-s", type=float, default=15.0)
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--forms", default="metrics,precision,batched,served",
                    help="extra forms to measure: metrics,precision,batched,served (or none)")
    ap.add_argument("--serve-jobs", type=int, default=2000)
    ap.add_argument("--bp-config", type=int, default=-1,
                    help="fix the backprojection tile configuration (index into the "
                         "instantiated set) instead of autotuning it on first use")
    ap.add_argument("--calibrate", action="store_true",
                    help="also launch the known-traffic kernels used to scale HBM counters: "
                         "phantom_kernel with 0 ellipsoids writes exactly 4*n^3 bytes, "
                         "minmax_kernel reads exactly 4*n^3 bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards by "
              "z-slab with no exchange step, and no multi-GPU figure is produced here.",
              file=sys.stderr)
        sys.exit(2)

    import qtctreconstruction_amd as ct
    n, na = args.n, args.angles
    ftype = int(ct.FilterType.SheppLogan)
    ctx = ct.Context(0)
    nvox = n * n * n
    d_vol = ctx.alloc(4 * nvox)
    d_out = ctx.alloc(4 * nvox)
    if args.calibrate:
        ctx.phantom_dev(args.seed, 0, n, n, n, d_vol)
    ctx.phantom_dev(args.seed, args.ellipsoids, n, n, n, d_vol)
    lib = ctx.lib
    if args.bp_config >= 0:
        ctx.set_bp_config(args.bp_config)

    def step():
        ctx._call("ct_hip_reconstruct_volume_dev", ctx.handle, d_vol.ptr, n, n, n, na, ftype,
                  d_out.ptr, None)

    log(f"[bench] warmup {args.warmup} steps of {n}^3 x {na}")
    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    stage = np.zeros(4)
    radon_k = np.zeros(2)
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        stage += ctx.timings()
        r_img, r_tr, walk_launches = ctx.radon_timings()
        radon_k += (r_img, r_tr)
    ctx.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    stage /= args.steps
    radon_k /= args.steps
    updates = float(nvox) * na
    value = updates / (ms / 1e3)
    log(f"[bench] {ms:.2f} ms/step; stages radon {stage[0]:.2f} (walk over the image "
        f"{radon_k[0]:.2f}, over the transpose {radon_k[1]:.2f}) filter {stage[1]:.2f} "
        f"bp {stage[2]:.2f} ms")
    roofline = dominant_roofline(n, na, stage, radon_k, walk_launches)
    log(f"[bench] roofline {roofline}")

    extra = {}
    vol_host = None
    if not args.no_parity or not args.no_cpu_baseline:
        vol_host = d_vol.download((n, n, n))
    if not args.no_parity:
        from oracle import oracle as O
        zl = [0, n // 3, n // 2, n - 1]
        out = d_out.download((n, n, n))
        t = time.perf_counter()
        ref, _ = O.reconstruct_volume(vol_host, na, ftype, z_list=zl)
        ref64 = O.reconstruct_volume64(vol_host, na, ftype, z_list=zl)
        extra["parity"] = {"slices": zl, "rel_l2_vs_serial": O.rel_l2(out[zl], ref[zl]),
                           "rel_l2_vs_fp64": O.rel_l2(out[zl], ref64[zl]),
                           "oracle_s": round(time.perf_counter() - t, 1)}
        log(f"[bench] parity {extra['parity']}")
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline")
        cpu = cpu_baseline(vol_host, na, ftype, args.cpu_target_s)
        log(f"[bench] cpu {cpu}")

    forms = [f for f in args.forms.split(",") if f and f != "none"]
    if "metrics" in forms:
        extra["metrics_form"] = metrics_form(ctx, d_vol, d_out, nvox, vol_host)
    if "precision" in forms:
        from oracle import oracle as O
        zl = [n // 2, n // 2 + 1]
        d_sino = ctx.alloc(4 * n * na * n)
        ctx.radon_dev(d_vol, n, n, n, na, n, d_sino)
        host_sample = np.stack([d_sino.download((n, na, n))[z] for z in zl]) if vol_host is not None \
            else d_sino.download((n, na, n))[zl]
        extra["precision_form"] = precision_form(ctx, d_sino, n, na, zl, host_sample)
        d_sino.free()
    if "batched" in forms:
        log("[bench] batched form")
        extra["batched_form"] = batched_form(ctx)
        log(f"[bench] batched {extra['batched_form']['batched_voxel_updates_per_s']:.3e} upd/s")
    if "served" in forms:
        d_vol.free()
        d_out.free()
        extra["served_form"] = served_form(ctx, args.serve_jobs)

    line = {
        "metric": METRIC, "value": value, "unit": "voxel-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom generated on the device)",
        "config": {"workload": f"plain FP32 reconstructVolume: {n}^3 phantom "
                               f"({args.ellipsoids} ellipsoids) -> Radon {na} angles x {n} bins "
                               f"-> Shepp-Logan FFT filter (sq {int(np.ceil(np.sqrt(2) * n))}) "
                               f"-> backprojection {n}^3, Serial semantics",
                   "n": n, "angles": na, "filter": "SheppLogan", "parallelism": "single GPU"},
        "stages_ms": {"normalise_radon": stage[0], "row_filter": stage[1],
                      "backprojection": stage[2]},
        "sinogram_samples_per_s": float(n) * na * n / (ms / 1e3),
        "stages_ms_radon": {"walk_image": radon_k[0], "walk_transpose": radon_k[1]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
