#!/usr/bin/env python
"""Benchmark of the MI355X NMF/NTF hot path (one JSON line on stdout).

Workload (the measured `value`): NTF-HALS on a 256 x 256 x 256 FP32 tensor
built from a CP rank-32 model with U[0,1) factors plus |N(0, 0.01*mean)|
noise, rank k = 32, dimension-tree MTTKRP, relative error every iteration.
A "step" is one outer iteration of computeNTF (all three mode updates and
the error).  W warmup iterations run untimed, then exactly K iterations are
timed between two device synchronisations; value = K / elapsed.

Also reported on the same line:
  roofline      -- the dominant kernel (the right partial MTTKRP
                   X_(65536 x 256) * A_2, one launch per outer iteration)
                   timed with HIP events on the stream it runs on; achieved =
                   algorithmic bytes (X read once + projection written) / time.
  cpu_baseline  -- the FP64 oracle restatement of the reference (OpenMP, all
                   allowed host threads) on a bounded sample of the same
                   workload from the same initial factors.
  parity        -- GPU FP32 vs the FP64 oracle relative-error trajectory over
                   that sample.
  legs          -- the mixed-precision comparisons of the headline metric:
                   NTF-HALS 256^3 in BF16/FP16, NMF-MU and NMF-HALS on
                   65536 x 16384 (k = 64) in FP32 vs BF16, the MTTKRP of a
                   128^4 tensor (k = 48, 1.07 GB, beyond the 256 MB Infinity
                   Cache) for the HBM roofline fraction, NTF-HALS on that
                   tensor in FP32 vs FP16, and the hierarchical NMF.

Usage: python bench.py [--gpus 1] [--steps 30] [--warmup 3] [--no-legs]
       [--no-cpu] [--profile-kernel]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "outer iters/sec NMF-MU & NTF-HALS: bf16-MFMA vs fp32 at equal rel. error"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def flat(X):
    """(numel, 1) view of a mode-0-fastest tensor in memory order."""
    return X.permute(*reversed(range(X.dim()))).reshape(-1)[:, None]


def event_time_ms(torch, fn, reps):
    """Average duration of fn() over reps launches on torch's current stream
    (the planc context runs on that stream)."""
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def make_ntf_problem(P, D, ctx, torch, dims, k, rank, stream):
    """CP tensor generated on the device from host-drawn factors."""
    F = [P.colmajor(D.uniform(D.DATA_SEED, 16 * stream + i, (d, rank)).astype(np.float32))
         for i, d in enumerate(dims)]
    # mean of the noiseless tensor = prod_i colmean -> sum over components
    means = np.prod([f.cpu().numpy().astype(np.float64).mean(0) for f in F], axis=0).sum()
    X = ctx.gen_cp_tensor(F, 0.01 * means, D.NOISE_SEED, stream)
    init = [f.astype(np.float32) for f in D.init_factors([(d, k) for d in dims], stream=stream)]
    return X, init


def run_ntf(P, ctx, torch, X, init, iters, dtype, X16=None):
    F = [P.colmajor(f) for f in init]
    lam, rel = ctx.ntf_run(X, F, P.HALS, iters, dtype=dtype, X16=X16)
    return F, rel


def ntf_rate(P, ctx, torch, X, init, warmup, steps, dtype, X16=None):
    F = [P.colmajor(f) for f in init]
    if warmup:
        ctx.ntf_run(X, F, P.HALS, warmup, dtype=dtype, X16=X16)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lam, rel = ctx.ntf_run(X, F, P.HALS, steps, dtype=dtype, X16=X16)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return steps / el, el, rel


def nmf_leg(P, D, ctx, torch, algo, m, n, k, iters, stream=3):
    W0 = P.colmajor(D.uniform(D.DATA_SEED, 16 * stream + 0, (m, k)).astype(np.float32))
    H0 = P.colmajor(D.uniform(D.DATA_SEED, 16 * stream + 1, (k, n)).astype(np.float32))
    mean = float((W0.double().mean(0) * H0.double().mean(1)).sum().item())
    A = ctx.gen_lowrank(W0, H0, 0.01 * mean, D.NOISE_SEED, stream)
    del W0, H0
    Wi = ctx.gen_uniform((m, k), D.INIT_SEED, 16 * stream + 0)
    Hi = ctx.gen_uniform((n, k), D.INIT_SEED, 16 * stream + 1)
    out = {}
    for name, dt in (("fp32", P.F32), ("bf16", P.BF16)):
        A16 = cs = None
        if dt != P.F32:
            A16, cs = ctx.convert(A, dt, scaled=True)
        W, H = Wi.clone(), Hi.clone()
        ctx.nmf_run(A, W, H, algo, 2, dtype=dt, A16=A16, colscale=cs)  # warmup
        W, H = Wi.clone(), Hi.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rel = ctx.nmf_run(A, W, H, algo, iters, dtype=dt, A16=A16, colscale=cs)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[name] = {"iters_per_s": iters / el, "final_relerr": float(rel[-1])}
        del A16, cs
    out["speedup_bf16_over_fp32"] = out["bf16"]["iters_per_s"] / out["fp32"]["iters_per_s"]
    out["relerr_gap"] = abs(out["bf16"]["final_relerr"] - out["fp32"]["final_relerr"])
    out["config"] = f"A {m}x{n} = W0 H0 (rank {k}) + 1% noise, k={k}, {iters} iterations"
    del A
    torch.cuda.empty_cache()
    return out


def hier_leg(P, D, ctx, m=100000, n=500000, density=1e-4, depth=11, iters=100):
    """Batched hierarchical NMF on a power-law term-document CSR (host CSR in,
    one level-batched launch sequence per tree level)."""
    t0 = time.perf_counter()
    indptr, indices, data = D.sparse_powerlaw(m, n, density, 1.0, 7)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    nodes, cols, _, _ = ctx.hiernmf_run(m, n, indptr, indices, data, depth, iters, 193957)
    el = time.perf_counter() - t0
    sizes = [x["ncols"] for x in nodes]
    return {"seconds": round(el, 4), "nodes": len(nodes), "nodes_per_s": round(len(nodes) / el, 1),
            "nnz": int(indptr[-1]), "iters_per_node": iters,
            "node_cols_min": int(min(sizes)), "node_cols_max": int(max(sizes)),
            "generate_s": round(gen_s, 2),
            "config": f"{m}x{n} CSR, density {density}, log-normal values, power-law rows, "
                      f"depth {depth}, {iters} iterations per node (host CSR in, tree out)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-legs", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--profile-kernel", action="store_true",
                    help="only launch the dominant kernel (and a calibration copy) for rocprof")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build "
              "runs on a single GPU; sharding over several GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)

    import torch

    import planc_amd as P
    from planc_amd import datagen as D

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    ctx = P.Context(0)
    dims, k = (256, 256, 256), 32
    log("bench: generating 256^3 CP rank-32 tensor on the device")
    X, init = make_ntf_problem(P, D, ctx, torch, dims, k, 32, stream=2)
    Fd = [P.colmajor(f) for f in init]
    L, R = dims[0] * dims[1], dims[2]
    Pbuf = P.empty_colmajor((L, k))
    Qbuf = P.empty_colmajor((R, k))

    def right():
        ctx.mttkrp_partial(X, Fd, 1, 0, out=Pbuf)

    def left():
        ctx.mttkrp_partial(X, Fd, 1, 1, out=Qbuf)

    if args.profile_kernel:
        cal = torch.empty(64 << 20, dtype=torch.float32, device="cuda")  # 256 MiB
        cal2 = torch.empty_like(cal)
        for _ in range(10):
            right()
        for _ in range(10):
            cal2.copy_(cal)
        torch.cuda.synchronize()
        print(json.dumps({"profile_kernel": "mttkrp_partial right 256^3 k=32 x10; copy 256MiB x10"}))
        ctx.close()
        return

    # ---- the measured workload ------------------------------------------
    log(f"bench: NTF-HALS fp32 warmup {args.warmup}, timed {args.steps}")
    rate, el, rel = ntf_rate(P, ctx, torch, X, init, args.warmup, args.steps, P.F32)
    ms_per_step = 1000.0 * el / args.steps

    # ---- dominant kernel roofline ----------------------------------------
    numel = int(np.prod(dims))
    t_right = event_time_ms(torch, right, 50)
    bytes_right = 4 * numel + 4 * L * k + 4 * R * k
    t_left = event_time_ms(torch, left, 50)
    bytes_left = 4 * numel + 4 * R * k + 4 * (dims[0] + dims[1]) * k
    achieved = bytes_right / (t_right * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "mttkrp_right_256cube_k32_traffic.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    # Achievable read bandwidth of this device (16-byte grid-stride reads of
    # 1 GiB, beyond the Infinity Cache): the ceiling besides the nameplate.
    probe = torch.zeros(1 << 28, dtype=torch.float32, device="cuda")
    read_ceiling = ctx.read_bandwidth(probe, 20)
    del probe
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "nn_f32_kernel<NT=2>, or nn_f32_sh_kernel<NT=2> where the per-shape tuner picks the block-shared variant (right partial MTTKRP, X_(65536x256)*A_2; FP32 MFMA, operands streamed to registers)",
                "algorithmic_bytes": bytes_right, "avg_ms": round(t_right, 5),
                "achievable_read_GBs": round(read_ceiling, 1),
                "frac_of_achievable": round(achieved / read_ceiling, 4),
                "note": "the 67 MB tensor fits the 256 MB Infinity Cache; see legs.mttkrp_128^4 "
                        "for a tensor that does not; achievable_read_GBs is this device's plain "
                        "16-byte read rate over 1 GiB (planc_read_bandwidth)"}
    left_info = {"avg_ms": round(t_left, 5),
                 "achieved_GBs": round(bytes_left / (t_left * 1e-3) / 1e9, 1),
                 "kernels": "tn_f32_kernel (KRP of A_0, A_1 fused) split-K + splitk_reduce_kernel"}

    # ---- parity and CPU baseline over a bounded sample -------------------
    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import oracle as O
        nthreads = min(16, os.cpu_count() or 1)
        O.set_threads(nthreads)
        n = args.cpu_iters
        Xh = X.cpu().numpy().astype(np.float64)
        log(f"bench: oracle NTF-HALS fp64, {n} iterations on {nthreads} threads")
        t0 = time.perf_counter()
        _, _, rel_cpu, rc = O.ntf(1, Xh, [f.astype(np.float64) for f in init], n, dimtree=1)
        el_cpu = time.perf_counter() - t0
        del Xh
        cpu = {"value": round(n / el_cpu, 4), "unit": "outer iters/sec", "cores": nthreads,
               "kind": "port",
               "sample": f"{n} outer iterations of NTF-HALS 256^3 k=32, FP64 oracle restatement "
                         f"(dimension tree, error every iteration), same input and init; "
                         f"{el_cpu:.1f} s"}
        _, rel_gpu = run_ntf(P, ctx, torch, X, init, n, P.F32)
        parity = {"iters": n, "max_abs_relerr_diff_vs_fp64": float(np.abs(rel_gpu - rel_cpu).max()),
                  "final_relerr_gpu_fp32": float(rel_gpu[-1]),
                  "final_relerr_oracle_fp64": float(rel_cpu[-1])}

    legs = {"left_partial_mttkrp": left_info}
    if not args.no_legs:
        total = args.warmup + args.steps
        log("bench: NTF-HALS bf16/fp16 legs")
        ntf_leg = {"fp32": {"iters_per_s": rate, "final_relerr": float(rel[-1])}}
        for name, dt in (("bf16", P.BF16), ("fp16", P.F16)):
            X16, _ = ctx.convert(flat(X), dt, scaled=False)
            r16, _, rel16 = ntf_rate(P, ctx, torch, X, init, args.warmup, args.steps, dt, X16=X16)
            ntf_leg[name] = {"iters_per_s": r16, "final_relerr": float(rel16[-1]),
                             "relerr_gap_vs_fp32": abs(float(rel16[-1]) - float(rel[-1]))}
            del X16
        ntf_leg["iterations"] = total
        legs["ntf_hals_256^3_k32"] = ntf_leg

        log("bench: MTTKRP 128^4 k=48 (HBM-resident)")
        del Pbuf, Qbuf
        dims4, k4 = (128, 128, 128, 128), 48
        X4, init4 = make_ntf_problem(P, D, ctx, torch, dims4, k4, 48, stream=5)
        F4 = [P.colmajor(f) for f in init4]
        L4 = dims4[0] * dims4[1]
        P4 = P.empty_colmajor((L4, k4))
        numel4 = int(np.prod(dims4))
        mt = {}
        for name, dt, Xop in (("fp32", P.F32, X4), ("fp16", P.F16, None)):
            if dt != P.F32:
                Xop, _ = ctx.convert(flat(X4), dt, scaled=False)
            esz = 4 if dt == P.F32 else 2

            def r4(Xop=Xop, dt=dt):
                ctx.mttkrp_partial(Xop, F4, 1, 0, dtype=dt, out=P4)
            t4 = event_time_ms(torch, r4, 10)
            b4 = esz * numel4 + 4 * L4 * k4
            mt[name] = {"avg_ms": round(t4, 4), "achieved_GBs": round(b4 / (t4 * 1e-3) / 1e9, 1),
                        "frac_of_8TBs": round(b4 / (t4 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "frac_of_achievable_read": round(b4 / (t4 * 1e-3) / 1e9 / read_ceiling, 4)}
        legs["mttkrp_128^4_k48_right_partial"] = mt
        del P4, F4
        # NTF-HALS on the 128^4 tensor (beyond the Infinity Cache), rank 48:
        # FP32 against FP16 MFMA in the MTTKRP, 10 timed iterations each after
        # 3 warm-up ones, final relative errors side by side.
        log("bench: NTF-HALS 128^4 k=48 fp32 vs fp16")
        nt = {}
        r32, el32, rel32 = ntf_rate(P, ctx, torch, X4, init4, 3, 10, P.F32)
        nt["fp32"] = {"iters_per_s": r32, "final_relerr": float(rel32[-1])}
        X416, _ = ctx.convert(flat(X4), P.F16, scaled=False)
        r16, el16, rel16 = ntf_rate(P, ctx, torch, X4, init4, 3, 10, P.F16, X16=X416)
        nt["fp16"] = {"iters_per_s": r16, "final_relerr": float(rel16[-1]),
                      "relerr_gap_vs_fp32": abs(float(rel16[-1]) - float(rel32[-1]))}
        nt["speedup_fp16_over_fp32"] = r16 / r32
        nt["config"] = "128^4 CP rank-48 model + 1% noise, k=48, dimension tree, 10 timed iterations"
        legs["ntf_hals_128^4_k48"] = nt
        del X4, X416
        torch.cuda.empty_cache()

        log("bench: NMF 65536x16384 k=64 fp32 vs bf16")
        legs["nmf_mu_65536x16384_k64"] = nmf_leg(P, D, ctx, torch, P.MU, 65536, 16384, 64, 20)
        legs["nmf_hals_65536x16384_k64"] = nmf_leg(P, D, ctx, torch, P.HALS, 65536, 16384, 64, 20)

        log("bench: hierarchical rank-2 NMF, 100000x500000 CSR, depth 11")
        legs["hiernmf_100000x500000_depth11"] = hier_leg(P, D, ctx)

    line = {
        "metric": METRIC,
        "value": round(rate, 3),
        "unit": "outer iters/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (Philox4x64 CP rank-32 model + 1% noise, generated on the device)",
        "config": {"workload": "NTF-HALS 256x256x256 FP32, k=32, dimension-tree MTTKRP, "
                               "relative error every iteration",
                   "dims": list(dims), "k": k, "algorithm": "HALS", "dimtree_split": 1},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "legs": legs,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
