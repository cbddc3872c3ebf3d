"""Write golden vectors from the reference's own FFT.cpp / FilterDesign.cpp.

Runs oracle/_ref/ref_filter (built by `make -C oracle ref` from the unmodified
reference sources) in this container only, and stores its outputs as small
fixtures next to this script. The fixtures are data: inputs and the reference's
outputs. Usage: python tests/golden/make_golden.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_filter")


def run(args, data=None):
    r = subprocess.run([REF] + [str(a) for a in args], input=data, capture_output=True, check=True)
    return np.frombuffer(r.stdout, dtype=np.float32).copy()


def main():
    if not os.path.exists(REF):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = {}
    # FilterDesign::createFilter for every filter type at the FFT sizes the configs use
    for n in (64, 128, 256, 512, 1024, 2048, 4096):
        for t in range(6):
            out[f"filter_n{n}_t{t}"] = run(["filter", n, t])
    # FFT::forwardReal / inverseReal on seeded inputs
    rng = np.random.default_rng(20261203)
    for n in (8, 64, 512, 2048):
        x = rng.standard_normal(n).astype(np.float32)
        out[f"fwd_in_n{n}"] = x
        out[f"fwd_out_n{n}"] = run(["fwd", n], x.tobytes())
        n2 = n // 2
        spec = (rng.standard_normal(n2 + 1) + 1j * rng.standard_normal(n2 + 1)).astype(np.complex64)
        out[f"inv_in_n{n}"] = spec.view(np.float32)
        out[f"inv_out_n{n}"] = run(["inv", n2], spec.tobytes())
    path = os.path.join(HERE, "ref_fft_filter.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    sys.exit(main())
