"""Benchmark of the FBP hot path on one MI355X.

A step is one reconstructVolume over device-resident input: per-slice
min/max + normalisation, ray-driven Radon (720 angles x 512 bins; the
LDS-staged row walk over the image and over its transpose, each timed with
HIP events), the Shepp-Logan row filter (sq 725: the 2048-point padded
convolution as short transforms of 3 x 512 points) and
voxel-driven backprojection with denormalisation, for a 512^3 phantom of 24
seeded ellipsoids (the plain-FP32 single-GPU configuration). value =
voxel-updates/s = 512^3 * 720 / step time, the nominal count of the workload
(the host baseline uses the same one); the backprojection's roofline counts
only the voxels inside the reconstruction circle, 78.5% of them, the ones
Serial computes (SerialBackend.cpp:219-223). The Radon walks skip the rows
whose 16 normalised slices are all zero (Serial skips zero pixels,
SerialBackend.cpp:104), and in a row walked a wave whose candidates all lie
outside the row's nonzero extent skips its sums; their roofline counts the
deposits of the rows walked (`radon_rows_walked`), and `dense_input` gives
the same step on a volume with no zero pixel.

Prints one JSON line. Timed region: K steps between two device
synchronisations; per-kernel times come from HIP events on the context's
stream. Beside it: the same call on host buffers (pinned, z-slab pipelined:
PCIe included), parity against the Serial restatement and FP64 (before and
after denormalisation), the OpenMPBackend restatement timed on the host
cores, and the precision, batched and served forms.
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
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 MFMA
PEAK_FP8_TFLOPS = 5000.0   # MI355X dense FP8 MFMA
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
    "radon_rows16_pair_kernel": 7.775e+09,
    "bp_kernel": 5.808e+09,
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu_info() -> dict:
    """lscpu's model, sockets and core counts of the host running the baseline."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    sockets = int(info.get("Socket(s)", "1") or 1)
    cores = int(info.get("Core(s) per socket", "0") or 0) * sockets
    logical = int(info.get("CPU(s)", "0") or 0) or (os.cpu_count() or 1)
    return {"model": info.get("Model name", "unknown"), "sockets": sockets,
            "physical_cores": cores or logical, "logical_cpus": logical,
            "threads_per_core": int(info.get("Thread(s) per core", "1") or 1),
            "hypervisor": info.get("Hypervisor vendor")}


def cpu_baseline_child(args) -> dict:
    """The OpenMP baseline in a process of its own: OMP_PROC_BIND / OMP_PLACES
    bind the OpenMP runtime's initial thread, and threads a process starts
    afterwards inherit that binding, so the GPU process must not carry it.
    The child rebuilds the same phantom with the oracle (bit-identical to the
    device's) and prints the baseline as one JSON line."""
    import subprocess
    env = dict(os.environ, OMP_PROC_BIND="spread", OMP_PLACES="cores")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", "--n", str(args.n),
           "--angles", str(args.angles), "--ellipsoids", str(args.ellipsoids),
           "--seed", str(args.seed), "--cpu-target-s", str(args.cpu_target_s)]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(vol_host: np.ndarray, na: int, ftype: int, target_s: float, runs: int = 3):
    """OpenMPBackend::reconstructVolume restatement (OpenMPBackend.cpp:61-190)
    on the host cores: one thread per physical core (OMP_NUM_THREADS when set),
    threads spread over cores; a z-subset of the same phantom sized to about
    target_s per run (the whole volume when it fits), median of `runs` runs,
    with its phase split (Radon + filter, backprojection)."""
    from oracle import oracle as O
    cpu = host_cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or cpu["physical_cores"]
    n = vol_host.shape[1]
    nz = 16
    t0 = time.perf_counter()
    O.omp_reconstruct_volume(vol_host[:nz], na, ftype, threads)
    per_slice = (time.perf_counter() - t0) / nz
    nz = int(min(vol_host.shape[0], max(16, target_s / max(per_slice, 1e-6))))
    nz -= nz % 2
    times, phases = [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        _, ph = O.omp_reconstruct_volume(vol_host[:nz], na, ftype, threads)
        times.append(time.perf_counter() - t0)
        phases.append(ph)
    k = int(np.argsort(times)[len(times) // 2])
    dt = times[k]
    upd = float(n) * n * nz * na
    return {"value": upd / dt, "unit": "voxel-updates/s", "cores": threads, "kind": "port",
            "sample": f"OpenMPBackend::reconstructVolume restatement ({O.ompbase().isa} build), "
                      f"{n}x{n}x{nz} slices of the same phantom, {na} angles, median of "
                      f"{runs} runs {dt:.2f} s wall",
            "runs_s": [round(t, 3) for t in times],
            "phase_ms": {"radon_and_filter": float(phases[k][0]),
                         "backprojection": float(phases[k][1]), "total": float(phases[k][2])},
            "host": cpu, "omp": {"threads": threads,
                                 "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
                                 "OMP_PLACES": os.environ.get("OMP_PLACES")}}


def walks_paired(n):
    """Whether both Radon walks (over the image for the angles with |sin| <=
    |cos|, over the transposed image for the others) take the 16-slice walk,
    rows of up to 516 px, and so run as one launch (radon_rows16_pair_kernel)."""
    return n + 4 <= 520


def circle_voxels(n: int) -> float:
    """Voxels of an n^3 volume inside the reconstruction circle of each
    slice, the ones Serial backprojects: (x - R)^2 + (y - R)^2 <= R^2 with
    R = n/2 (SerialBackend.cpp:210-223)."""
    r = n // 2
    d = np.arange(n, dtype=np.int64) - r
    return float(np.count_nonzero(d[:, None] ** 2 + d[None, :] ** 2 <= r * r)) * n


def walked_deposits(vol, na: int, n_row: int) -> dict:
    """Deposits (pixel, angle, slice) the Radon walks run on this volume.

    The 16-slice walk (rows of up to 516 px) skips, per pair of 8-slice
    z-blocks, the rows of the image (for the n_row angles walked over it)
    and the columns (for the others, walked over the transpose) whose 16
    normalised slices are all zero; a normalised pixel is zero where the raw
    one is 0 (SerialBackend.cpp:104 skips it) or the slice's minimum.
    Wider images take the 8-slice walk, which walks every row."""
    nz, h, w = vol.shape
    if not walks_paired(w):
        return {"deposits": float(w) * h * nz * na, "rows_walked_frac": 1.0,
                "columns_walked_frac": 1.0}
    mn = vol.reshape(nz, -1).min(axis=1)
    live = (vol != 0) & (vol != mn[:, None, None])
    live |= np.isnan(vol)
    npair = (nz + 15) // 16
    pad = npair * 16 - nz
    if pad:  # the last pair's second z-block stands in for a missing one
        live = np.concatenate([live, live[-8:][:pad]], axis=0)
    g = live.reshape(npair, 16, h, w).any(axis=1)
    rows = g.any(axis=2).sum(axis=1)  # per pair
    cols = g.any(axis=1).sum(axis=1)
    dep = 16.0 * (float(rows.sum()) * w * n_row + float(cols.sum()) * h * (na - n_row))
    return {"deposits": dep, "rows_walked_frac": float(rows.sum()) / (npair * h),
            "columns_walked_frac": float(cols.sum()) / (npair * w),
            "nonzero_pixel_frac": float(live[:nz].mean())}


def dominant_roofline(n, na, kt, launches, radon_dep):
    """Roofline entry of the kernel with the largest time per step.

    Candidates: the Radon row walks (radon_rows16_pair_kernel: one launch
    walking the image for the angles with |sin| <= |cos| and the transposed
    image for the others; past 516 px, one launch per walk) and the
    backprojection (bp_kernel). Both are FP32 VALU work (no MFMA, no HBM
    bound: a few % of HBM); achieved = algorithmic FLOP per step (9 per
    pixel-angle-slice deposit the walks run, radon_dep of them, or per
    voxel-update of the voxels inside the circle) / the kernel's time per
    step from HIP events on the context stream.
    """
    walk = "radon_rows16_pair_kernel" if walks_paired(n) else "radon_rows_kernel"
    cands = {
        walk: (kt["walk_image"] + kt["walk_transpose"],
               max(launches, 1) * (1 if walks_paired(n) else 2), FLOP_PER_SAMPLE * radon_dep,
               "Radon LDS row walks (image and transposed image)"),
        "bp_kernel": (kt["backprojection"], 1, FLOP_PER_UPDATE * circle_voxels(n) * na,
                      "voxel-driven backprojection"),
    }
    name = max(cands, key=lambda k: cands[k][0])
    t_ms, nl, flop, what = cands[name]
    achieved = flop / (t_ms / 1e3) / 1e12
    out = {"bound": "fp32", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
           "frac": achieved / PEAK_FP32_TFLOPS, "traffic": TRAFFIC_BYTES.get(name),
           "kernel": name, "launches_per_step": nl, "avg_launch_ms": t_ms / nl,
           "note": f"{what}; FP32 VALU work, no MFMA; "
                   f"{'9 FLOP per pixel-angle-slice deposit of the rows walked (rows whose 16 slices are all zero are skipped, and waves whose candidates lie outside a row nonzero extent skip their sums)' if name.startswith('radon') else '9 FLOP per voxel-update inside the circle'}"
                   "; 157.3 TF is MI355X's FP32 peak, which counts packed FP32; these kernels "
                   "are built without packed FP32 ops (78.6 TF non-packed issue ceiling)"}
    if name.startswith("radon"):
        dep = FLOP_PER_SAMPLE_DEPOSIT * radon_dep / (t_ms / 1e3) / 1e12
        out["deposit_arithmetic_only"] = {"flop_per_sample": FLOP_PER_SAMPLE_DEPOSIT,
                                          "achieved": dep, "frac": dep / PEAK_FP32_TFLOPS}
    return out


def stage_rooflines(n, na, kt, n_row, radon_dep):
    """Every stage of a step against the roofline that binds it: the streaming
    stages (min/max, normalisation, row filter) against HBM with their
    compulsory bytes, the Radon walks and the backprojection against FP32
    (with their compulsory HBM bytes beside). Times: HIP events per kernel."""
    sq = int(np.ceil(np.sqrt(2.0) * n))
    vox = float(n) ** 3
    def hbm(bytes_, ms):
        gbs = bytes_ / (ms / 1e3) / 1e9
        return {"bound": "hbm", "ms": ms, "bytes": bytes_, "achieved_gbs": gbs,
                "peak_gbs": PEAK_HBM_GBS, "frac": gbs / PEAK_HBM_GBS}
    def fp32(flop, bytes_, ms):
        tf = flop / (ms / 1e3) / 1e12
        return {"bound": "fp32", "ms": ms, "flop": flop, "achieved_tflops": tf,
                "peak_tflops": PEAK_FP32_TFLOPS, "frac": tf / PEAK_FP32_TFLOPS,
                "compulsory_bytes": bytes_,
                "hbm_frac_at_compulsory_bytes": bytes_ / (ms / 1e3) / 1e9 / PEAK_HBM_GBS}
    copies = 2.0 * (n + 4) * n * n * 4  # the two zero-padded normalised copies
    out = {"minmax": hbm(4 * vox, kt["minmax"]), "normalise": hbm(4 * vox + copies, kt["normalise"])}
    if walks_paired(n):
        out["radon_walks"] = fp32(FLOP_PER_SAMPLE * radon_dep, copies + 4.0 * n * na * n,
                                  kt["walk_image"] + kt["walk_transpose"])
    else:
        out["radon_walk_image"] = fp32(FLOP_PER_SAMPLE * vox * n_row,
                                       copies / 2 + 4.0 * n * n_row * n, kt["walk_image"])
        out["radon_walk_transpose"] = fp32(FLOP_PER_SAMPLE * vox * (na - n_row),
                                           copies / 2 + 4.0 * n * (na - n_row) * n,
                                           kt["walk_transpose"])
    out["row_filter"] = hbm(4.0 * n * na * n + 4.0 * n * na * sq, kt["row_filter"])
    out["backprojection"] = fp32(FLOP_PER_UPDATE * circle_voxels(n) * na, 4.0 * n * na * sq + 4 * vox,
                                 kt["backprojection"])
    return out


def precision_form(ctx, d_sino, n, na, host_sinos_sample):
    """Row filter variants on the full sinogram stack (n*na rows): filter-stage
    time with its MFMA and HBM fractions, and the filtered rows' error vs the
    exact FP64 padded-Toeplitz filter on sampled slices."""
    import ctypes

    import qtctreconstruction_amd as ct
    from oracle import oracle as O
    sq, pad, nfft = O.geometry(n)
    rows = n * na
    # the GEMM's work counts the K = n nonzero columns of the padded rows
    flop = 2.0 * rows * n * sq
    hbm_bytes = 4.0 * rows * n + 4.0 * rows * sq  # FP32 rows in, FP32 filtered rows out
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
            ms = []
            for _ in range(3):
                ctx.reconstruct_from_sinograms_dev(d_sino, n, na, n, ang, ftype, d_out)
                ms.append(ctx.timings()[1:3])
            t = np.median(np.array(ms), axis=0)
            got = ctx.filter_sinograms(host_sinos_sample, ftype).reshape(ref.shape)
            r = {"filter_ms": float(t[0]), "backprojection_ms": float(t[1]),
                 "rel_l2_vs_fp64": O.rel_l2(got, ref),
                 "max_abs_vs_fp64": float(np.abs(got - ref).max()),
                 "hbm_gbs": hbm_bytes / (t[0] / 1e3) / 1e9}
            r["hbm_frac"] = r["hbm_gbs"] / PEAK_HBM_GBS
            if fb != ct.FilterBackend.FFT_F32:
                f = flop * (3 if fb == ct.FilterBackend.GEMM_SPLIT_BF16 else 1)
                peak = PEAK_FP8_TFLOPS if fb == ct.FilterBackend.GEMM_FP8 else PEAK_BF16_TFLOPS
                r["mfma_tflops"] = f / (t[0] / 1e3) / 1e12
                r["mfma_peak_tflops"] = peak
                r["mfma_frac"] = r["mfma_tflops"] / peak
                r["bound"] = "hbm" if hbm_bytes / PEAK_HBM_GBS / 1e9 > f / peak / 1e12 else "mfma"
            else:
                r["bound"] = "hbm"
            res[f"{fname}/{fb.name}"] = r
            log(f"[bench] precision {fname}/{fb.name}: {r}")
    ctx.set_filter_backend(ct.FilterBackend.FFT_F32)
    d_out.free()
    res["gemm_flop"] = flop
    res["rows"] = rows
    res["errors"] = "filtered rows of 2 slices vs the exact FP64 padded Toeplitz"
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


def make_volume_sinos(ctx, seed, N, nz, na, n_ell=16, bins=None):
    """Device sinograms (nz, na, bins) of a seeded N x N phantom, Serial
    semantics (computeSinogram with `bins` detector bins, default N)."""
    bins = bins or N
    d_vol = ctx.alloc(4 * N * N * nz)
    ctx.phantom_dev(seed, n_ell, N, N, nz, d_vol)
    d_s = ctx.alloc(4 * nz * na * bins)
    ctx.radon_dev(d_vol, N, N, nz, na, bins, d_s)
    d_vol.free()
    return d_s


def batched_form(ctx, seed=4):
    """16 uneven volumes, one persistent filter + one persistent backprojection
    launch (hipGraph, autotuned tiles) vs the same volumes one by one. Detector
    bins B = ceil(sqrt2 N) + 2 (the whole diagonal of the N x N slice, with two
    bins to spare); each volume is reconstructed as the consumer path does,
    onto B x B (reconstructVolumeFromSinograms: output_size = bins)."""
    import qtctreconstruction_amd as ct
    rng = np.random.default_rng(seed)
    shapes = [(128, 128), (256, 256), (384, 192), (512, 256), (768, 128)]
    jobs = []
    for k in range(16):
        N, nz = shapes[rng.integers(len(shapes))]
        na = int(rng.choice([180, 360, 720, 1440]))
        f = int(rng.choice([0, 1, 4]))
        B = int(np.ceil(np.sqrt(2.0) * N)) + 2
        d_s = make_volume_sinos(ctx, 1000 + k, N, nz, na, int(rng.integers(10, 25)), bins=B)
        d_o = ctx.alloc(4 * B * B * nz)
        step = np.float32(180.0) / np.float32(na)
        ang = (step * np.arange(na, dtype=np.float32)).astype(np.float32)
        jobs.append((d_s, d_o, ang, nz, na, B, f, N))
    updates = float(sum(B * B * nz * na for (_, _, _, nz, na, B, _, _) in jobs))
    bjobs = [j[:7] for j in jobs]
    flags = ct.CT_BATCH_GRAPH | ct.CT_BATCH_AUTOTUNE
    ctx.fbp_batch(bjobs, flags)  # tunes, captures
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.fbp_batch(bjobs, flags)
    t_batch = (time.perf_counter() - t0) / reps

    # the same batch with the split-BF16 Toeplitz GEMM as the row filter (one
    # MFMA launch per volume ahead of the captured backprojection)
    ctx.set_filter_backend(ct.FilterBackend.GEMM_SPLIT_BF16)
    ctx.fbp_batch(bjobs, flags)  # designs the Toeplitz tables, captures
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.fbp_batch(bjobs, flags)
    t_split = (time.perf_counter() - t0) / reps
    ctx.set_filter_backend(ct.FilterBackend.FFT_F32)

    def one_by_one():
        for (d_s, d_o, ang, nz, na, B, f, _) in jobs:
            ctx.reconstruct_from_sinograms_dev(d_s, nz, na, B, ang, f, d_o)
    one_by_one()  # first use tunes each (N, na) class; not timed
    t0 = time.perf_counter()
    one_by_one()
    t_seq = time.perf_counter() - t0
    for (d_s, d_o, *_r) in jobs:
        d_s.free()
        d_o.free()
    return {"volumes": [[N, N, nz, na, B] for (_, _, _, nz, na, B, _, N) in jobs],
            "volumes_key": "[N, N, depth, angles, detector bins]",
            "voxel_updates": updates, "batched_ms": t_batch * 1e3,
            "batched_voxel_updates_per_s": updates / t_batch,
            "one_by_one_ms": t_seq * 1e3, "one_by_one_voxel_updates_per_s": updates / t_seq,
            "batched_vs_one_by_one": t_seq / t_batch,
            "batched_split_bf16_ms": t_split * 1e3,
            "batched_split_bf16_voxel_updates_per_s": updates / t_split,
            "bins": "ceil(sqrt2 N) + 2; output B x B (consumer path)"}


def served_form(ctx, n_jobs, seed=5, pool_gb=96.0, rate_factor=0.7):
    """Poisson stream of jobs over a device-resident input pool; p50/p99
    latency and throughput, beside plain per-job HIP FBP on the same mix (with
    the inputs on the device, and through the pinned host-buffer path). In the
    Poisson run every result is also downloaded to host memory on a separate
    thread, as a client of reconstructVolumeFromSinograms receives it: latency
    arrival -> host result."""
    import ctypes
    import queue
    import threading

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
    sample = jobs[: min(200, n_jobs)]
    for N in edges:  # tune every class once before timing
        for na in angle_sets:
            ctx.reconstruct_from_sinograms_dev(pool[(N, na)], 8, na, N, angles[na], 0, d_out)
    t0 = time.perf_counter()
    for (N, d, na, f) in sample:
        ctx.reconstruct_from_sinograms_dev(pool[(N, na)], d, na, N, angles[na], f, d_out)
    t_plain = time.perf_counter() - t0
    plain_ups = float(upd[: len(sample)].sum()) / t_plain
    d_out.free()
    # The first 12 of those jobs through the host-buffer entry, as a caller of
    # reconstructVolumeFromSinograms hands them over (host sinograms in, host
    # volume out; pinned staging, z-slab pipelined): the plain per-job rate
    # with its PCIe transfers.
    f32p = ctypes.POINTER(ctypes.c_float)
    hsample = sample[:12]
    t_host = 0.0
    for (N, d, na, f) in hsample:
        hs = pool[(N, na)].download((d, na, N))
        ho = np.zeros((d, N, N), np.float32)
        ctx._call("ct_hip_reconstruct_from_sinograms", ctx.handle, hs.ctypes.data_as(f32p), d, na,
                  N, angles[na].ctypes.data_as(f32p), f, ho.ctypes.data_as(f32p))
        t0 = time.perf_counter()
        ctx._call("ct_hip_reconstruct_from_sinograms", ctx.handle, hs.ctypes.data_as(f32p), d, na,
                  N, angles[na].ctypes.data_as(f32p), f, ho.ctypes.data_as(f32p))
        t_host += time.perf_counter() - t0
        del hs, ho
    host_ups = float(sum(N * N * d * na for (N, d, na, f) in hsample)) / t_host
    log(f"[bench] serving: plain through host buffers {host_ups:.3e} upd/s "
        f"({len(hsample)} jobs, PCIe included)")
    # plain jobs/s of this mix: the plain rate over the mix's mean job (the
    # sample's jobs are lighter or heavier than the mix's by chance)
    plain_jobs_s = plain_ups / float(upd.mean())
    lam = rate_factor * plain_jobs_s
    log(f"[bench] serving: plain {plain_ups:.3e} upd/s, {plain_jobs_s:.1f} jobs/s of the mix "
        f"-> rate {lam:.1f}/s")

    poisson_arrivals = np.cumsum(rng.exponential(1.0 / lam, n_jobs))
    host_buf = np.zeros(max(N * N * d for (N, d, na, f) in jobs), np.float32)

    def run(open_loop, policy=ct.SERVE_SIZE_AGED, cap_ms=5.0, host_results=False):
        srv = ct.Server(pool_bytes=int(pool_gb * (1 << 30)), max_batch=32, slots=2)
        srv.set_policy(policy, batch_cap_ms=cap_ms)
        arrivals = poisson_arrivals if open_loop else np.zeros(n_jobs)
        lat, lat_host = [], []
        job_of = {}
        dlq: "queue.Queue" = queue.Queue()
        start = time.perf_counter()

        def downloader():
            while True:
                item = dlq.get()
                if item is None:
                    return
                t, k = item
                N, d, na, f = jobs[k]
                srv.download_into(t, host_buf[: d * N * N].reshape(d, N, N))
                lat_host.append((time.perf_counter() - start - arrivals[k]) * 1e3)
                srv.release(t)

        th = threading.Thread(target=downloader) if host_results else None
        if th:
            th.start()
        done = 0
        k = 0
        while done < n_jobs:
            now = time.perf_counter() - start
            while k < n_jobs and arrivals[k] <= now:
                N, d, na, f = jobs[k]
                job_of[srv.submit(pool[(N, na)], angles[na], d, na, N, f)] = k
                k += 1
            wait_ms = 0.2 if k < n_jobs else 1000.0
            r = srv.wait(wait_ms)
            while r is not None:
                lat.append(r[1])
                if th:
                    dlq.put((r[0], job_of[r[0]]))
                else:
                    srv.release(r[0])
                done += 1
                r = srv.wait(0.0) if done < n_jobs else None
        span_dev = time.perf_counter() - start
        if th:
            dlq.put(None)
            th.join()
        span = time.perf_counter() - start
        st = srv.stats()
        srv.close()
        lat = np.array(lat)
        out = {"jobs": n_jobs, "span_s": span_dev,
               "voxel_updates_per_s": float(upd.sum()) / span_dev,
               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
               "batches": st["batches"], "max_jobs_in_flight": st["max_inflight_jobs"],
               "overlapped_launches": st["overlapped_launches"]}
        if th:
            lh = np.array(lat_host)
            out["host_result"] = {"p50_ms": float(np.percentile(lh, 50)),
                                  "p99_ms": float(np.percentile(lh, 99)), "span_s": span,
                                  "voxel_updates_per_s": float(upd.sum()) / span,
                                  "bytes_downloaded": float(sum(4.0 * N * N * d
                                                                for (N, d, na, f) in jobs)),
                                  "latency": "Poisson arrival -> volume in host memory"}
        return out

    log("[bench] serving: closed loop (all jobs queued at t=0)")
    saturated = run(False)
    log(f"[bench] serving saturated: {saturated}")
    log("[bench] serving: Poisson arrivals, results downloaded to host memory")
    poisson = run(True, host_results=True)
    log(f"[bench] serving poisson: {poisson}")
    log("[bench] serving: the same Poisson arrivals, FIFO admission, no batch cap")
    poisson_fifo = run(True, ct.SERVE_FIFO, 0.0)
    log(f"[bench] serving poisson fifo: {poisson_fifo}")
    for d_s in pool.values():
        d_s.free()
    offered = lam * float(upd.mean())
    return {"jobs": n_jobs, "rate_per_s": lam, "rate_factor_of_plain": rate_factor,
            "mean_job_voxel_updates": float(upd.mean()),
            "offered_load_vs_plain": offered / plain_ups,
            "offered_load_vs_saturated_server": offered / saturated["voxel_updates_per_s"],
            "plain_sample_jobs": len(sample),
            "plain_per_job_voxel_updates_per_s": plain_ups,
            "plain_per_job_host_buffers_voxel_updates_per_s": host_ups,
            "poisson_vs_plain_device": poisson["voxel_updates_per_s"] / plain_ups,
            "poisson_host_results_vs_plain_host_buffers":
                poisson["host_result"]["voxel_updates_per_s"] / host_ups,
            "plain_host_buffers_note": f"first {len(hsample)} jobs of the mix through "
                                       "ct_hip_reconstruct_from_sinograms (host sinograms in, "
                                       "host volume out, pinned staging pipelined by z-slabs)",
            "saturated": saturated, "poisson": poisson, "poisson_fifo": poisson_fifo,
            "policy": "size-ordered with waiting credit 0.25, batch cap max(5 ms, queue/4); "
                      "poisson_fifo: FIFO without a cap on the same arrivals",
            "latency": "submit -> completion in device memory; poisson.host_result: "
                       "arrival -> volume in host memory",
            "noise": f"counts ~ Poisson(1e4 exp(-kappa p)) (normal approx.), kappa = 2/N, log"}


def n_row_angles(na: int) -> int:
    """Angles the Radon stage walks over the image (|sin| <= |cos|); the rest
    walk the transpose."""
    th = (np.float32(180.0) / np.float32(na) * np.arange(na, dtype=np.float32)) * \
        np.float32(np.pi / 180.0)
    return int((np.abs(np.sin(th)) <= np.abs(np.cos(th))).sum())


def dense_input(ctx, vol_host, n, na, ftype, updates):
    """The same step on a volume with no zero pixel (the phantom plus
    1 + 0.001 x, so every slice keeps a span): the Radon walks skip no row,
    which is their cost on a dense volume."""
    nvox = n * n * n
    dense = vol_host + (np.float32(1.0) + np.float32(1e-3) * np.arange(n, dtype=np.float32))
    d_dense = ctx.alloc(4 * nvox).upload(dense)
    d_out = ctx.alloc(4 * nvox)
    dense_walk = walked_deposits(dense, na, n_row_angles(na))
    del dense
    ctx.reconstruct_volume_dev(d_dense, n, n, n, na, ftype, d_out)
    ctx.synchronize()
    kd = {}
    t0 = time.perf_counter()
    for _ in range(3):
        ctx.reconstruct_volume_dev(d_dense, n, n, n, na, ftype, d_out)
        for key, v in ctx.kernel_timings().items():
            kd[key] = kd.get(key, 0.0) + v / 3
    ctx.synchronize()
    ms_dense = (time.perf_counter() - t0) * 1e3 / 3
    d_dense.free()
    d_out.free()
    return {"ms_per_step": ms_dense, "voxel_updates_per_s": updates / (ms_dense / 1e3),
            "radon_walks_ms": kd["walk_image"] + kd["walk_transpose"], "kernels_ms": kd,
            "rows_walked_frac": dense_walk["rows_walked_frac"],
            "columns_walked_frac": dense_walk["columns_walked_frac"],
            "note": "the same step on the phantom plus 1 + 0.001 x (no zero pixel): no row of "
                    "the Radon walks is skipped"}


def api_form(vol_host, n, na, ftype, reps=10, warm_s=0.6):
    """reconstructVolume on host buffers (ct_hip_reconstruct_volume): the
    volume in and out through pinned staging, z-slabs pipelined so the PCIe
    transfers overlap the compute. The output array is allocated and touched
    once, as a caller reusing its buffers does. Warm-up: calls for warm_s
    seconds (their times are reported too). Measured at 512^3 x 720: after
    seconds of host-only work, about the first dozen back-to-back calls
    (0.5 s) take 26-40 ms, some kernels and copies in them up to 3.5x slower
    than later (kernel and copy trace), and the calls after that 25.5-25.9
    ms."""
    import ctypes

    import qtctreconstruction_amd as ct
    be = ct.HipBackend(0)
    f32p = ctypes.POINTER(ctypes.c_float)
    out = np.zeros_like(vol_host)

    def call():
        be.ctx._call("ct_hip_reconstruct_volume", be.ctx.handle, vol_host.ctypes.data_as(f32p),
                     n, n, n, na, ftype, out.ctypes.data_as(f32p), None)

    warm_ms = []
    t_warm = time.perf_counter()
    while len(warm_ms) < 2 or time.perf_counter() - t_warm < warm_s:
        t0 = time.perf_counter()
        call()
        warm_ms.append((time.perf_counter() - t0) * 1e3)
    ms = []
    for _ in range(reps):
        t0 = time.perf_counter()
        call()
        ms.append((time.perf_counter() - t0) * 1e3)
    t = be.ctx.timings()
    med = float(np.median(ms))
    res = {"ms_per_step": med, "runs_ms": ms, "voxel_updates_per_s": float(n) ** 3 * na / (med / 1e3),
           "device_stage_ms": float(t[3]), "pcie_bytes": 2.0 * vol_host.nbytes,
           "note": "host volume in, host volume out: pinned staging, z-slabs of 32, 64, 128, "
                   "128, 64, 64, 32 slices on three streams (upload of slab k+1 and download of "
                   "slab k-1 overlap the compute of slab k; pieces of each slab cross PCIe while "
                   "the host copies the next; a second host thread copies outputs out); median "
                   f"of {reps} runs after {len(warm_ms)} warm-up calls ({warm_s} s)",
           "warmup_runs_ms": [round(x, 2) for x in warm_ms]}
    be.ctx.close()
    return out, res


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
    ap.add_argument("--cpu-target-s", type=float, default=9.0,
                    help="CPU work per baseline run (the sample grows to this; 3 runs)")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-api", action="store_true")
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the step on a volume with no zero pixel (dense_input)")
    ap.add_argument("--cpu-baseline-child", action="store_true", help=argparse.SUPPRESS)
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
    if args.cpu_baseline_child:
        from oracle import oracle as O
        vol = O.phantom(args.seed, args.ellipsoids, args.n, args.n, args.n, threads=16)
        print(json.dumps(cpu_baseline(vol, args.angles, 1, args.cpu_target_s)),  # Shepp-Logan
              flush=True)
        return
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
    if args.bp_config >= 0:
        ctx.set_bp_config(args.bp_config)

    def step():
        ctx.reconstruct_volume_dev(d_vol, n, n, n, na, ftype, d_out)

    log(f"[bench] warmup {args.warmup} steps of {n}^3 x {na}")
    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    stage = np.zeros(4)
    kt = {}
    walk_launches = 1
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        stage += ctx.timings()
        for key, v in ctx.kernel_timings().items():
            kt[key] = kt.get(key, 0.0) + v
        walk_launches = ctx.radon_timings()[2]
    ctx.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    stage /= args.steps
    kt = {key: v / args.steps for key, v in kt.items()}
    updates = float(nvox) * na
    value = updates / (ms / 1e3)
    log(f"[bench] {ms:.2f} ms/step; stages radon {stage[0]:.2f} (min/max {kt['minmax']:.3f}, "
        f"normalise {kt['normalise']:.3f}, walks {kt['walk_image'] + kt['walk_transpose']:.2f}) "
        f"filter {stage[1]:.2f} bp {stage[2]:.2f} ms")
    vol_host = d_vol.download((n, n, n))
    walked = walked_deposits(vol_host, na, n_row_angles(na))
    roofline = dominant_roofline(n, na, kt, walk_launches, walked["deposits"])
    stages_roof = stage_rooflines(n, na, kt, n_row_angles(na), walked["deposits"])
    cfg = args.bp_config if args.bp_config >= 0 else ctx.tuned_config(n, na)
    tile = ctx.bp_configs()[cfg]
    bp_tiles = {"config": cfg, "voxel_tile": tile[0], "angles_per_stage": tile[1],
                "slices_per_thread": tile[2], "chosen_by": "fixed" if args.bp_config >= 0 else "autotuner"}
    log(f"[bench] roofline {roofline}")

    extra = {}
    if not args.no_dense:
        extra["dense_input"] = dense_input(ctx, vol_host, n, na, ftype, updates)
        log(f"[bench] dense input: {extra['dense_input']['ms_per_step']:.2f} ms/step, "
            f"walks {extra['dense_input']['radon_walks_ms']:.2f} ms")
    d_sino = None
    if not args.no_parity or "precision" in args.forms:
        d_sino = ctx.alloc(4 * n * na * n)
        ctx.reconstruct_volume_dev(d_vol, n, n, n, na, ftype, d_out, d_sino)
    if not args.no_parity:
        from oracle import oracle as O
        zl = [0, n // 3, n // 2, n - 1]
        out = d_out.download((n, n, n))
        t = time.perf_counter()
        ref, ref_sino = O.reconstruct_volume(vol_host, na, ftype, z_list=zl, threads=16,
                                             want_sino=True)
        ref64 = O.reconstruct_volume64(vol_host, na, ftype, z_list=zl, threads=16)
        # before denormalisation: the consumer path over the same device
        # sinograms (the backprojection's output before the fused epilogue)
        # against Serial's reconstructSlice of Serial's sinograms, and FP64
        ang = O.serial_angles(na)
        d_pre = ctx.alloc(4 * nvox)
        ctx.reconstruct_from_sinograms_dev(d_sino, n, na, n, ang, ftype, d_pre)
        pre = d_pre.download((n, n, n))[zl]
        d_pre.free()
        pre_ref = np.stack([O.reconstruct_slice(ref_sino[k], ang, n, ftype) for k in range(len(zl))])
        pre64 = np.stack([O.reconstruct_slice64(O.compute_sinogram64(vol_host[z], na, n),
                                                ang.astype(np.float64), n, ftype) for z in zl])
        extra["parity"] = {"slices": zl,
                           "rel_l2_vs_serial": O.rel_l2(out[zl], ref[zl]),
                           "rel_l2_vs_fp64": O.rel_l2(out[zl], ref64[zl]),
                           "before_denormalisation": {"rel_l2_vs_serial": O.rel_l2(pre, pre_ref),
                                                      "rel_l2_vs_fp64": O.rel_l2(pre, pre64)},
                           "oracle_s": round(time.perf_counter() - t, 1)}
        log(f"[bench] parity {extra['parity']}")
    api = None
    if not args.no_api:
        log("[bench] reconstructVolume on host buffers")
        api_out, api = api_form(vol_host, n, na, ftype)
        api["bit_identical_to_device_path"] = bool(
            np.array_equal(api_out.view(np.uint32), d_out.download((n, n, n)).view(np.uint32)))
        api["vs_device_resident"] = api["voxel_updates_per_s"] / value
        del api_out
        log(f"[bench] api {api}")
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (child process)")
        cpu = cpu_baseline_child(args)
        log(f"[bench] cpu {cpu}")

    forms = [f for f in args.forms.split(",") if f and f != "none"]
    if "metrics" in forms:
        extra["metrics_form"] = metrics_form(ctx, d_vol, d_out, nvox, vol_host)
    if "precision" in forms:
        zl = [n // 2, n // 2 + 1]
        host_sample = np.stack([d_sino.download((n, na, n))[z] for z in zl])
        extra["precision_form"] = precision_form(ctx, d_sino, n, na, host_sample)
    if d_sino is not None:
        d_sino.free()
    if "batched" in forms:
        log("[bench] batched form")
        extra["batched_form"] = batched_form(ctx)
        log(f"[bench] batched {extra['batched_form']['batched_voxel_updates_per_s']:.3e} upd/s")
    if "served" in forms:
        d_vol.free()
        d_out.free()
        extra["served_form"] = served_form(ctx, args.serve_jobs)
    # The forms against the plain form's throughput (this line's value: the
    # whole reconstructVolume pipeline, Radon included; the consumer-path forms
    # start from sinograms, so they do less work per voxel-update) and, in
    # their own fields, against plain per-job FBP on the same jobs.
    vs_plain = {}
    if "batched_form" in extra:
        bf = extra["batched_form"]
        vs_plain["batched"] = bf["batched_voxel_updates_per_s"] / value
        vs_plain["batched_split_bf16"] = bf["batched_split_bf16_voxel_updates_per_s"] / value
    if "served_form" in extra:
        sf = extra["served_form"]
        vs_plain["served_saturated"] = sf["saturated"]["voxel_updates_per_s"] / value
        vs_plain["served_poisson_at_0.7_load"] = sf["poisson"]["voxel_updates_per_s"] / value
    if vs_plain:
        extra["forms_vs_plain_reconstructVolume"] = vs_plain

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
        "kernels_ms": kt,
        "backprojection_tiles": bp_tiles,
        "sinogram_samples_per_s": float(n) * na * n / (ms / 1e3),
        "radon_rows_walked": walked,
        "stages_ms_radon": ({"walks_one_launch": kt["walk_image"] + kt["walk_transpose"]}
                            if walks_paired(n) else
                            {"walk_image": kt["walk_image"], "walk_transpose": kt["walk_transpose"]}),
        "roofline": roofline,
        "stages_roofline": stages_roof,
        "api_ms_per_step": api["ms_per_step"] if api else None,
        "api_form": api,
        "cpu_baseline": cpu,
        "vs_cpu_baseline": value / cpu["value"] if cpu else None,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
