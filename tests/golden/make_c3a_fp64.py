"""Writes tests/golden/nmf_c3a_fp64.json: the FP64 oracle's relative-error
trajectories of NMF-MU, NMF-HALS and NMF-ANLS-BPP on the mixed-precision NMF
configuration (A 65536 x 16384 = W0 H0 with rank-64 U[0,1) factors plus 1 %
noise, k = 64, 50 iterations), the values the BF16 and FP32 GPU runs are
judged against ("equal relative error": both within 1e-3 of these).

The input is built on the host exactly as bench.py / the GPU tests build it
on the device (planc_gen_lowrank, planc_gen_uniform):
  W0 = float32(U(DATA_SEED, 16*3 + 0)), H0 = float32(U(DATA_SEED, 16*3 + 1)),
  sigma = 0.01 * sum_r mean(W0[:, r]) * mean(H0[r, :]),
  A = float32(W0 H0 + sigma |N(NOISE_SEED, 3)|)   (product in FP64),
  W = float32(U(INIT_SEED, 16*3 + 0)), H = float32(U(INIT_SEED, 16*3 + 1)).
The device forms W0 H0 with its own FP64 summation order and its own
log/cos, so a few elements of A may differ from these by one FP32 ulp; the
FP64 sum of A is stored so a GPU run can report the difference (it moves
the final relative error by far less than 1e-6).

The oracle runs in FP64 on the float32 values; its GEMMs go through the host
CBLAS (MKL) when one loads, else its own blocked GEMM.

Usage: python tests/golden/make_c3a_fp64.py [--iters 50] [--algos mu,hals,anlsbpp]
(needs about 25 GB of host memory; minutes on 8 cores).
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

M, N, R, K, STREAM = 65536, 16384, 64, 64, 3
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nmf_c3a_fp64.json")


def build_input(m=M, n=N, r=R, stream=STREAM, chunk=1024):
    W0 = D.uniform(D.DATA_SEED, 16 * stream + 0, (m, r)).astype(np.float32)
    H0 = D.uniform(D.DATA_SEED, 16 * stream + 1, (r, n)).astype(np.float32)
    mean = float((W0.astype(np.float64).mean(0) * H0.astype(np.float64).mean(1)).sum())
    sigma = 0.01 * mean
    A = np.empty((m, n), dtype=np.float32, order="F")
    W0d = W0.astype(np.float64)
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        w = D.words(D.NOISE_SEED, stream, 2 * m * (c1 - c0), offset=2 * m * c0)
        u = (w >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        z = np.sqrt(-2.0 * np.log(1.0 - u[0::2])) * np.cos(2.0 * np.pi * u[1::2])
        blk = W0d @ H0[:, c0:c1].astype(np.float64) + sigma * np.abs(z.reshape((m, c1 - c0), order="F"))
        A[:, c0:c1] = blk.astype(np.float32)
    Wi = D.uniform(D.INIT_SEED, 16 * stream + 0, (m, K)).astype(np.float32)
    Hi = D.uniform(D.INIT_SEED, 16 * stream + 1, (n, K)).astype(np.float32)
    return A, Wi, Hi, sigma


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--algos", default="mu,hals,anlsbpp")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    O.set_threads(args.threads)
    blas, _ = O.use_blas(args.threads)
    t0 = time.time()
    A, Wi, Hi, sigma = build_input()
    print(f"input built in {time.time() - t0:.1f} s; blas={blas}", flush=True)
    res = json.load(open(OUT)) if os.path.exists(OUT) else {}
    res.update({
        "config": f"A {M}x{N} = W0 H0 (rank {R}, U[0,1)) + 1% noise (sigma = 0.01*mean), "
                  f"k={K}, {args.iters} iterations, seeded U[0,1) init (stream {STREAM})",
        "sigma": sigma,
        "sum_A_fp64": float(A.astype(np.float64).sum()),
        "norm_A_fp64": float(np.sqrt((A.astype(np.float64) ** 2).sum())),
        "oracle_gemm": blas or "oracle blocked GEMM",
        "iters": args.iters,
    })
    A64 = np.asfortranarray(A, dtype=np.float64)
    del A
    codes = {"mu": 0, "hals": 1, "anlsbpp": 2}
    for name in args.algos.split(","):
        t0 = time.time()
        _, _, rel, _, rc = O.nmf(codes[name], A64, Wi.astype(np.float64), Hi.astype(np.float64),
                                 args.iters)
        el = time.time() - t0
        print(f"{name}: rc={rc} final relerr {rel[-1]:.12f} in {el:.1f} s", flush=True)
        if rc != 0:
            raise SystemExit(f"oracle {name} failed with rc={rc}")
        res[name] = [float(x) for x in rel]
        res[f"{name}_seconds"] = round(el, 1)
        json.dump(res, open(OUT, "w"), indent=1)


if __name__ == "__main__":
    main()
