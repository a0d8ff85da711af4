"""Tile tuner for the complex MFMA GEMM (zgemm.hip).

Times every wave-tile configuration the library is built with (and a few K splits for the
deep-k A^H B products) on the products of H|psi> and block Davidson for a given plane-wave
count, band count and projector count, and prints the fastest per shape class and precision.
The library's built-in table (zgemm.hip, zgemm_t) is what this printed for the 64-atom Si cell
(npw 51,627, 160 bands, 256 projectors); Q3X_GEMM_TILE="TMWxTNW/ksplit" forces one
configuration at run time, which is how this script measures the candidates.

    python -m q3_amd.tune_gemm [--npw 51627] [--bands 160] [--nproj 256] [--reps 20]
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

TILES = ["2x2", "4x2", "2x5", "4x4", "4x5"]
SPLITS = [-1, 0, 64, 128]


def main(argv=None):
    from q3_amd import FP64, MIXED
    from q3_amd.planewave import Context, DeviceArray, zgemm

    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--npw", type=int, default=51627)
    ap.add_argument("--bands", type=int, default=160)
    ap.add_argument("--nproj", type=int, default=256)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--json", default="")
    args = ap.parse_args(argv)
    npw, nb, npj = args.npw, args.bands, args.nproj
    shapes = [("C", npj, nb, npw, "projector overlap"), ("N", npw, nb, npj, "projector update"),
              ("C", 2 * nb, nb, npw, "Gram block"), ("N", npw, nb, 2 * nb, "basis rotation")]
    ctx = Context(0)
    rng = np.random.default_rng(0)
    warm = True  # the first timings of a fresh process run before the clocks settle

    def rnd(shape):
        return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))

    results = []
    for ta, m, n, k, name in shapes:
        A = DeviceArray.from_host(ctx, rnd((m, k) if ta == "N" else (k, m)))
        B = DeviceArray.from_host(ctx, rnd((k, n)))
        C = DeviceArray.from_host(ctx, rnd((m, n)))
        lda = m if ta == "N" else k
        flops = 8.0 * m * n * k
        for prec, label in ((FP64, "fp64"), (MIXED, "fp32")):
            cands = [("", None)] + [(f"{t}/{s}", t) for t in TILES for s in (SPLITS if ta == "C" else [0])]
            best = None
            for cfg, _ in cands:
                if cfg:
                    os.environ["Q3X_GEMM_TILE"] = cfg
                else:
                    os.environ.pop("Q3X_GEMM_TILE", None)
                for _ in range(60 if warm else 3):
                    zgemm(ctx, ta, m, n, k, 1.0, A, lda, B, k, 0.0, C, m, prec)
                warm = False
                ctx.synchronize()
                t0 = time.perf_counter()
                for _ in range(args.reps):
                    zgemm(ctx, ta, m, n, k, 1.0, A, lda, B, k, 0.0, C, m, prec)
                ctx.synchronize()
                dt = (time.perf_counter() - t0) / args.reps
                rec = {"shape": name, "m": m, "n": n, "k": k, "precision": label, "tile": cfg or "built-in",
                       "ms": dt * 1e3, "tflops": flops / dt / 1e12}
                results.append(rec)
                if best is None or rec["ms"] < best["ms"]:
                    best = rec
            builtin = next(r for r in results if r["shape"] == name and r["precision"] == label
                           and r["tile"] == "built-in")
            print(f"{name:18s} {label}: best {best['tile']:8s} {best['tflops']:6.1f} TF/s, "
                  f"built-in {builtin['tflops']:6.1f} TF/s", flush=True)
    os.environ.pop("Q3X_GEMM_TILE", None)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(results, f, indent=1)


if __name__ == "__main__":
    main()
