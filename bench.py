"""Benchmark of the FBP hot path on one MI355X.

A step is one reconstructVolume over device-resident input: per-slice
min/max + normalisation, ray-driven Radon (720 angles x 512 bins), the
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
FLOP_PER_UPDATE = 9        # t (3), u (1), frac (1), lerp (3), accumulate (1)


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
    ctx.phantom_dev(args.seed, args.ellipsoids, n, n, n, d_vol)
    lib = ctx.lib

    def step():
        ctx._call("ct_hip_reconstruct_volume_dev", ctx.handle, d_vol.ptr, n, n, n, na, ftype,
                  d_out.ptr, None)

    log(f"[bench] warmup {args.warmup} steps of {n}^3 x {na}")
    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    stage = np.zeros(4)
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        stage += ctx.timings()
    ctx.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    stage /= args.steps
    updates = float(nvox) * na
    value = updates / (ms / 1e3)
    bp_ms = stage[2]
    bp_tflops = FLOP_PER_UPDATE * updates / (bp_ms / 1e3) / 1e12
    log(f"[bench] {ms:.2f} ms/step; stages radon {stage[0]:.2f} filter {stage[1]:.2f} "
        f"bp {stage[2]:.2f} ms")

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
        "roofline": {"bound": "mfma", "achieved": bp_tflops, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": bp_tflops / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "bp_kernel (backprojection)",
                     "note": "FP32 VALU-bound; 9 algorithmic FLOP per voxel-update; "
                             "157.3 TF is MI355X's FP32 peak (vector and matrix alike)"},
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
