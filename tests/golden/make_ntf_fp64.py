"""Writes tests/golden/ntf_fp64.json: the FP64 oracle's relative-error
trajectories of NTF-HALS (dimension tree) on the two full-size NTF
configurations, the values the FP32 and 16-bit GPU runs are judged against:
  "256^3": X 256 x 256 x 256, CP rank 32 + 1 % noise, k = 32, 30 iterations;
  "128^4": X 128^4, CP rank 48 + 1 % noise, k = 48, 20 iterations.

The input is built on the host exactly as the GPU tests and bench.py build it
on the device (planc_gen_cp_tensor from host-drawn factors):
  F_i = float32(U(DATA_SEED, 16*stream + i)),  (dims[i] x rank)
  sigma = 0.01 * sum_r prod_i mean(F_i[:, r]),
  X = float32(sum_r prod_i F_i(idx_i, r) + sigma |N(NOISE_SEED, stream)|_idx)
      (products and sums in FP64, mode 0 fastest),
  initial factors float32(U(INIT_SEED, 16*stream + i)) (dims[i] x k).
The device sums the rank terms in its own order, so an element of X may
differ from these by one FP32 ulp; the FP64 sum of X is stored so a GPU run
can check that it holds the same input.

Usage: python tests/golden/make_ntf_fp64.py [--only 256^3|128^4]
(about 5 GB of host memory for 128^4; a few minutes on 8 cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from planc_amd import datagen as D  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ntf_fp64.json")
CASES = {
    "256^3": {"dims": (256, 256, 256), "rank": 32, "k": 32, "stream": 2, "iters": 30},
    "128^4": {"dims": (128, 128, 128, 128), "rank": 48, "k": 48, "stream": 5, "iters": 20},
}


def build_input(dims, rank, k, stream):
    F = [D.uniform(D.DATA_SEED, 16 * stream + i, (d, rank)).astype(np.float32) for i, d in enumerate(dims)]
    F64 = [f.astype(np.float64) for f in F]
    sigma = 0.01 * float(np.prod([f.mean(0) for f in F64], axis=0).sum())
    N = len(dims)
    subs = "abcdefgh"[:N]
    numel = int(np.prod(dims))
    X = np.empty(numel, dtype=np.float32)
    slab = int(np.prod(dims[:-1]))
    spec = ",".join(s + "z" for s in subs) + "->" + subs
    for c in range(dims[-1]):  # one slab of the last mode at a time
        blk = np.einsum(spec, *F64[:-1], F64[-1][c:c + 1], optimize=True).reshape(-1, order="F")
        w = D.words(D.NOISE_SEED, stream, 2 * slab, offset=2 * slab * c)
        u = (w >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        z = np.sqrt(-2.0 * np.log(1.0 - u[0::2])) * np.cos(2.0 * np.pi * u[1::2])
        X[c * slab:(c + 1) * slab] = (blk + sigma * np.abs(z)).astype(np.float32)
    init = [D.uniform(D.INIT_SEED, 16 * stream + i, (d, k)).astype(np.float32) for i, d in enumerate(dims)]
    return X.reshape(dims, order="F"), init, sigma


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default=None, choices=sorted(CASES))
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    O.set_threads(args.threads)
    blas, _ = O.use_blas(args.threads)
    res = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name, c in CASES.items():
        if args.only and name != args.only:
            continue
        t0 = time.time()
        X, init, sigma = build_input(c["dims"], c["rank"], c["k"], c["stream"])
        X64 = np.asfortranarray(X, dtype=np.float64)
        print(f"{name}: input built in {time.time() - t0:.1f} s; blas={blas}", flush=True)
        t0 = time.time()
        _, lam, rel, rc = O.ntf(1, X64, [f.astype(np.float64) for f in init], c["iters"], dimtree=1)
        el = time.time() - t0
        print(f"{name}: rc={rc} final relerr {rel[-1]:.12f} in {el:.1f} s", flush=True)
        if rc != 0:
            raise SystemExit(f"oracle NTF-HALS {name} failed with rc={rc}")
        res[name] = {
            "config": f"X {'x'.join(map(str, c['dims']))} = CP rank {c['rank']} (U[0,1) factors) + 1% noise "
                      f"(sigma = 0.01 * sum_r prod_i mean F_i(:, r)), k={c['k']}, NTF-HALS dimension tree, "
                      f"{c['iters']} iterations, seeded U[0,1) init (stream {c['stream']})",
            **c,
            "sigma": sigma,
            "sum_X_fp64": float(X64.sum()),
            "relerr": [float(x) for x in rel],
            "oracle_gemm": blas or "oracle blocked GEMM",
            "seconds": round(el, 1),
        }
        del X, X64
        json.dump(res, open(OUT, "w"), indent=1)


if __name__ == "__main__":
    main()
