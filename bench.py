#!/usr/bin/env python
"""Benchmark of the MI355X NMF/NTF hot path (one JSON line on stdout).

Workload of the measured `value`: NMF-MU (MUNMF::computeNMF, nmf/mu.hpp:48-91)
on A = W0 H0 + 1 % noise, 65536 x 16384 FP32 (rank-64 U[0,1) factors), rank
k = 64, with the two bulk products W^T A and A H on BF16 MFMA (FP32
accumulation, per-column power-of-two scales) and the Grams and error in
FP64.  A "step" is one outer iteration of computeNMF (H update, W update,
relative error).  W warmup iterations run untimed from the seeded initial
factors; then, from the same initial factors, exactly K iterations are timed
between two device synchronisations in one computeNMF call; value = K /
elapsed.  A and its BF16 copy are resident in HBM before the timed region
(the one-time conversion is timed separately).

Also on the same line:
  fp32 / ratio  -- the same K iterations with FP32 products (the plain GPU
                   form) and the BF16/FP32 ratio of outer iterations/s; both
                   final relative errors next to the FP64 oracle's
                   (tests/golden/nmf_c3a_fp64.json, same input and init) and
                   whether both lie within 1e-3 of it.
  roofline      -- the dominant kernel of the step, the BF16 W^T A product
                   (k x 16384 = W^T (64 x 65536) A), timed with HIP events on
                   the stream it runs on; achieved = algorithmic bytes / time.
  cpu_baseline  -- the FP64 oracle restatement of MUNMF::computeNMF on the
                   same A and init on the host's cores, its GEMMs through a
                   host CBLAS when one loads (MKL), for a bounded number of
                   iterations; plus an FP32 variant.
  legs          -- NMF-HALS and NMF-ANLS-BPP on the same A in FP32 vs BF16,
                   NTF-HALS on a 128^4 tensor (k = 48) in FP32 vs FP16 with
                   its MTTKRP's HBM fraction, NTF-HALS 256^3 (k = 32) in FP32,
                   BF16 and FP16 (both NTF legs next to the FP64 oracle's
                   trajectory, tests/golden/ntf_fp64.json, same input and
                   init), and the hierarchical rank-2 NMF.

Usage: python bench.py [--gpus 1] [--steps 50] [--warmup 3] [--no-legs]
       [--no-cpu] [--cpu-iters 3] [--profile-kernel]
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "outer iters/sec NMF-MU & NTF-HALS: bf16-MFMA vs fp32 at equal rel. error"
HBM_PEAK_GBS = 8000.0
M_C3, N_C3, K_C3, R_C3, STREAM_C3 = 65536, 16384, 64, 64, 3
FP64_FIXTURE = os.path.join(ROOT, "tests", "golden", "nmf_c3a_fp64.json")
NTF_FP64_FIXTURE = os.path.join(ROOT, "tests", "golden", "ntf_fp64.json")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "gemm_tn_bf16_65536x16384_k64_traffic.json")


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


def host_info():
    """lscpu summary and the threads this process may use."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            if key.strip() in ("Model name", "CPU(s)", "Socket(s)", "Core(s) per socket",
                               "Thread(s) per core", "NUMA node(s)"):
                info[key.strip()] = val.strip()
    except Exception as e:  # lscpu missing: record why
        info["lscpu"] = f"unavailable: {e}"
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except Exception:
        info["affinity_cpus"] = os.cpu_count()
    info["OMP_NUM_THREADS"] = os.environ.get("OMP_NUM_THREADS")
    return info


def make_nmf_problem(P, D, ctx, torch, m, n, r, k, stream):
    """A = W0 H0 + 0.01 mean(W0 H0) |N(0,1)| on the device, and the seeded
    U[0,1) initial W (m x k) and H (n x k)."""
    W0 = P.colmajor(D.uniform(D.DATA_SEED, 16 * stream + 0, (m, r)).astype(np.float32))
    H0 = P.colmajor(D.uniform(D.DATA_SEED, 16 * stream + 1, (r, n)).astype(np.float32))
    mean = float((W0.double().mean(0) * H0.double().mean(1)).sum().item())
    A = ctx.gen_lowrank(W0, H0, 0.01 * mean, D.NOISE_SEED, stream)
    del W0, H0
    Wi = ctx.gen_uniform((m, k), D.INIT_SEED, 16 * stream + 0)
    Hi = ctx.gen_uniform((n, k), D.INIT_SEED, 16 * stream + 1)
    return A, Wi, Hi


def norm2_fp64(torch, A, chunk=1024):
    """||A||_F^2 in FP64 (column blocks, so no full FP64 copy is formed)."""
    s = 0.0
    for j in range(0, A.shape[1], chunk):
        s += float((A[:, j:j + chunk].double() ** 2).sum().item())
    return s


def nmf_timed(P, ctx, torch, A, Wi, Hi, algo, warmup, steps, dtype, A16=None, cs=None,
              A16T=None, normA2=0.0):
    """warmup untimed iterations, then `steps` timed iterations from the same
    initial factors in one computeNMF call; returns (iters/s, seconds, relerr).
    A16T / normA2 are the constructor-time At and ||A||^2 (made once)."""
    kw = dict(dtype=dtype, A16=A16, colscale=cs, A16T=A16T, normA2=normA2)
    if warmup:
        W, H = Wi.clone(), Hi.clone()
        ctx.nmf_run(A, W, H, algo, warmup, **kw)
    W, H = Wi.clone(), Hi.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rel = ctx.nmf_run(A, W, H, algo, steps, **kw)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return steps / el, el, rel


def fp64_reference(name, steps):
    """FP64 oracle relative error after `steps` iterations (same input and
    init), from the committed fixture; None when it does not cover them."""
    if not os.path.exists(FP64_FIXTURE):
        return None
    fx = json.load(open(FP64_FIXTURE))
    traj = fx.get(name)
    if not traj or steps > len(traj) or steps < 1:
        return None
    return float(traj[steps - 1]), [float(x) for x in traj[:steps]]


def ntf_fp64_compare(case, steps, runs):
    """Each run's final error and trajectory against the FP64 oracle's
    NTF-HALS trajectory on the same input and init (committed fixture);
    `runs` maps a label to its relative-error trajectory."""
    if not os.path.exists(NTF_FP64_FIXTURE):
        return {}
    fx = json.load(open(NTF_FP64_FIXTURE)).get(case)
    if not fx or len(fx["relerr"]) != steps:
        return {}
    ref = np.array(fx["relerr"])
    out = {"final_relerr_fp64_oracle": float(ref[-1]),
           "fp64_source": "tests/golden/ntf_fp64.json (FP64 oracle, same input and init)"}
    ok = True
    for label, rel in runs.items():
        rel = np.asarray(rel, dtype=np.float64)
        d = float(rel[-1] - ref[-1])
        out[f"{label}_minus_fp64"] = d
        out[f"{label}_max_abs_traj_diff_vs_fp64"] = float(np.abs(rel - ref).max())
        ok = ok and abs(d) < 1e-3
    out["equal_error_within_1e-3"] = bool(ok)
    return out


def compare_precisions(P, ctx, torch, A, Wi, Hi, A16, cs, A16T, normA2, algo, name, warmup, steps):
    out = {}
    ref = fp64_reference(name, steps)
    for label, dt in (("fp32", P.F32), ("bf16", P.BF16)):
        b = dt != P.F32
        rate, el, rel = nmf_timed(P, ctx, torch, A, Wi, Hi, algo, warmup, steps, dt,
                                  A16 if b else None, cs if b else None, A16T if b else None,
                                  normA2)
        out[label] = {"iters_per_s": round(rate, 3), "ms_per_iter": round(1000 * el / steps, 4),
                      "final_relerr": float(rel[-1])}
        if ref is not None:
            out[label]["final_relerr_minus_fp64"] = float(rel[-1]) - ref[0]
            out[label]["max_abs_traj_diff_vs_fp64"] = float(np.abs(rel - np.array(ref[1])).max())
            out[label]["monotone"] = bool(np.all(np.diff(rel) <= 1e-12))
    out["speedup_bf16_over_fp32"] = round(out["bf16"]["iters_per_s"] / out["fp32"]["iters_per_s"], 4)
    if ref is not None:
        out["final_relerr_fp64_oracle"] = ref[0]
        out["equal_error_within_1e-3"] = bool(
            abs(out["bf16"]["final_relerr_minus_fp64"]) < 1e-3 and
            abs(out["fp32"]["final_relerr_minus_fp64"]) < 1e-3)
    out["iterations"] = steps
    return out


def make_ntf_problem(P, D, ctx, torch, dims, k, rank, stream):
    """CP tensor generated on the device from host-drawn factors."""
    F = [P.colmajor(D.uniform(D.DATA_SEED, 16 * stream + i, (d, rank)).astype(np.float32))
         for i, d in enumerate(dims)]
    means = np.prod([f.cpu().numpy().astype(np.float64).mean(0) for f in F], axis=0).sum()
    X = ctx.gen_cp_tensor(F, 0.01 * means, D.NOISE_SEED, stream)
    init = [f.astype(np.float32) for f in D.init_factors([(d, k) for d in dims], stream=stream)]
    return X, init


def ntf_rate(P, ctx, torch, X, init, warmup, steps, dtype, X16=None, X16T=None, reps=5):
    """computeNTF (HALS) timed over `steps` iterations after `warmup`, `reps`
    times from the same initial factors (a run of 20 iterations at 128^4 is
    about 10 ms); returns the median rate, its elapsed time and the
    trajectory (the runs are deterministic).  ||X||^2 is computed once
    before, as AUNTF's constructor computes the norm."""
    normX2 = float((flat(X).double() ** 2).sum().item())
    F = [P.colmajor(f) for f in init]
    if warmup:
        ctx.ntf_run(X, F, P.HALS, warmup, dtype=dtype, X16=X16, X16T=X16T, normX2=normX2)
    els = []
    for _ in range(reps):
        F = [P.colmajor(f) for f in init]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lam, rel = ctx.ntf_run(X, F, P.HALS, steps, dtype=dtype, X16=X16, X16T=X16T, normX2=normX2)
        torch.cuda.synchronize()
        els.append(time.perf_counter() - t0)
    el = float(np.median(els))
    return steps / el, el, rel


def cpu_baseline_leg(O, A, Wi, Hi, iters):
    """FP64 oracle MU (and the FP32 variant) on the host, full shape, `iters`
    timed iterations after the host BLAS has been warmed up."""
    info = host_info()
    nthreads = int(os.environ.get("OMP_NUM_THREADS") or info.get("affinity_cpus") or 1)
    O.set_threads(nthreads)
    blas, has_s = O.use_blas(nthreads)
    t0 = time.perf_counter()
    A32 = A.cpu().numpy()
    W0 = Wi.cpu().numpy().astype(np.float64)
    H0 = Hi.cpu().numpy().astype(np.float64)
    A64 = np.asfortranarray(A32, dtype=np.float64)
    copy_s = time.perf_counter() - t0
    small = np.asfortranarray(np.ones((256, 64)))
    O.gemm("T", "N", small, small)  # brings up the BLAS thread pool
    log(f"bench: CPU oracle MU fp64, {iters} iterations on {nthreads} threads ({blas})")
    _, _, rel64, _, rc = O.nmf(0, A64, W0, H0, iters)
    s64 = O.last_loop_seconds()
    del A64
    f32 = None
    if has_s:
        log(f"bench: CPU MU fp32 variant, {iters} iterations")
        _, _, rel32, rc32 = O.mu_f32(A32, W0.astype(np.float32), H0.astype(np.float32), iters)
        if rc32 == 0:
            s32 = O.last_loop_seconds()
            f32 = {"value": round(iters / s32, 4), "seconds": round(s32, 3),
                   "final_relerr": float(rel32[-1])}
    O.no_blas()
    return {"value": round(iters / s64, 4), "unit": "outer iters/sec", "cores": nthreads,
            "kind": "port",
            "sample": f"{iters} outer iterations of NMF-MU on the full {M_C3}x{N_C3} A, k={K_C3}, "
                      f"same A and init, FP64 oracle restatement of MUNMF::computeNMF "
                      f"(error by computeObjectiveError every iteration), GEMMs through "
                      f"{blas or 'the oracle blocked GEMM'}; loop {s64:.2f} s "
                      f"(steady_clock around the loop; host copy {copy_s:.1f} s outside)",
            "rc": rc, "final_relerr_fp64": float(rel64[-1]),
            "fp32_variant": f32, "host": info}


def hier_leg(P, D, ctx, m=100000, n=500000, density=1e-4, depth=11, iters=100, frontier=1024):
    """Batched hierarchical NMF on a power-law term-document CSR (host CSR in,
    one level-batched launch sequence per tree level)."""
    t0 = time.perf_counter()
    indptr, indices, data = D.sparse_powerlaw(m, n, density, 1.1, 7)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    nodes, cols, _, _ = ctx.hiernmf_run(m, n, indptr, indices, data, depth, iters, 193957)
    el = time.perf_counter() - t0
    sizes = [x["ncols"] for x in nodes]
    # The reference's best-first frontier (--nodes budget) on the same matrix:
    # 1024 accepted nodes report 2047, as many as the complete tree of depth 11.
    t0 = time.perf_counter()
    fnodes, _, _, _ = ctx.hiernmf_run(m, n, indptr, indices, data, 62, iters, 193957, nodes=frontier)
    fel = time.perf_counter() - t0
    return {"seconds": round(el, 4), "nodes": len(nodes), "nodes_per_s": round(len(nodes) / el, 1),
            "nnz": int(indptr[-1]), "iters_per_node": iters,
            "node_cols_min": int(min(sizes)), "node_cols_max": int(max(sizes)),
            "generate_s": round(gen_s, 2),
            "frontier": {"budget": frontier, "seconds": round(fel, 4), "nodes": len(fnodes),
                         "max_level": int(max(x["level"] for x in fnodes)),
                         "nodes_per_s": round(len(fnodes) / fel, 1)},
            "config": f"{m}x{n} CSR, density {density}, log-normal values, power-law rows "
                      f"(alpha 1.1), depth {depth}, {iters} iterations per node (host CSR in, "
                      f"tree out); frontier: budget {frontier}, level cap 62 (a max_level below 61 means the cap never bound)"}


def profile_ntf4(P, D, ctx, torch):
    """The 128^4 k=48 partial MTTKRPs alone (for counter runs): right and left
    contractions in FP32 (left: TN on X) and FP16 (left: NN on the transposed
    copy, as planc_ntf_run runs it), thirteen launches each (the first tunes),
    then ten copies of a 1 GiB buffer for the counter calibration."""
    dims4, k4 = (128, 128, 128, 128), 48
    X4, init4 = make_ntf_problem(P, D, ctx, torch, dims4, k4, 48, stream=5)
    X416, X416T = ctx.ntf_copies16(X4, P.F16)
    F4 = [P.colmajor(f) for f in init4]
    F4T = [F4[2], F4[3], F4[0], F4[1]]
    L4 = dims4[0] * dims4[1]
    P4 = P.empty_colmajor((L4, k4))
    Q4 = P.empty_colmajor((L4, k4))
    for dt, Xop, Xl, Fl, sl in ((P.F32, X4, X4, F4, 1), (P.F16, X416, X416T, F4T, 0)):
        for Xs, Fs, side, out in ((Xop, F4, 0, P4), (Xl, Fl, sl, Q4)):
            torch.zeros(7, device="cuda")  # a fill kernel: marks the start of a phase in the counter lists
            for _ in range(13):
                ctx.mttkrp_partial(Xs, Fs, 1, side, dtype=dt, out=out)
    cal = torch.empty(256 << 20, dtype=torch.float32, device="cuda")  # 1 GiB
    cal2 = torch.empty_like(cal)
    for _ in range(10):
        cal2.copy_(cal)
    torch.cuda.synchronize()
    print(json.dumps({"profile_kernel": "128^4 k=48 partial MTTKRP right/left FP32/FP16 x13; copy 1 GiB x10"}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-legs", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--legs", default="hals,bpp,ntf4,ntf3,hier",
                    help="comma-separated legs after the headline: hals, bpp, ntf4 (NTF-HALS "
                         "128^4), ntf3 (NTF-HALS 256^3), hier (hierarchical NMF)")
    ap.add_argument("--profile-kernel", nargs="?", const="wta", default=None, choices=["wta", "ntf4", "hier"],
                    help="only launch one kernel family (and a calibration copy) for rocprof: "
                         "wta = BF16 W^T A of the headline; ntf4 = the 128^4 k=48 partial "
                         "MTTKRPs (right and left, FP32 and FP16); hier = the hierarchical NMF leg")
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
    if args.profile_kernel == "ntf4":
        profile_ntf4(P, D, ctx, torch)
        ctx.close()
        return
    if args.profile_kernel == "hier":  # the hierarchical NMF leg alone (kernel traces)
        print(json.dumps({"profile_kernel": "hiernmf 100000x500000", "leg": hier_leg(P, D, ctx)}))
        ctx.close()
        return
    legs_on = set(x.strip() for x in args.legs.split(",") if x.strip())
    m, n, k = M_C3, N_C3, K_C3
    log(f"bench: generating A {m}x{n} (rank {R_C3} + 1% noise) on the device")
    A, Wi, Hi = make_nmf_problem(P, D, ctx, torch, m, n, R_C3, k, STREAM_C3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A16, cs = ctx.convert(A, P.BF16, scaled=True)
    A16T = ctx.transpose16(A16)
    torch.cuda.synchronize()
    convert_ms = 1000 * (time.perf_counter() - t0)
    normA2 = norm2_fp64(torch, A)
    Wt = Wi.clone()

    def wta():
        ctx.gemm_tn(A16, Wt, dtype=P.BF16, colscale=cs)

    if args.profile_kernel:
        cal = torch.empty(256 << 20, dtype=torch.float32, device="cuda")  # 1 GiB
        cal2 = torch.empty_like(cal)
        for _ in range(10):
            wta()
        for _ in range(10):
            cal2.copy_(cal)
        torch.cuda.synchronize()
        print(json.dumps({"profile_kernel": f"BF16 W^T A {m}x{n} k={k} x10; copy 1 GiB x10"}))
        del cal, cal2
        ctx.close()
        return

    # ---- the measured workload ------------------------------------------
    log(f"bench: NMF-MU bf16 warmup {args.warmup}, timed {args.steps}")
    rate, el, rel = nmf_timed(P, ctx, torch, A, Wi, Hi, P.MU, args.warmup, args.steps, P.BF16, A16, cs,
                              A16T, normA2)
    ms_per_step = 1000.0 * el / args.steps
    log(f"bench: NMF-MU fp32 warmup {args.warmup}, timed {args.steps}")
    rate32, el32, rel32 = nmf_timed(P, ctx, torch, A, Wi, Hi, P.MU, args.warmup, args.steps, P.F32,
                                    normA2=normA2)
    ref = fp64_reference("mu", args.steps)
    comparison = {
        "fp32_iters_per_s": round(rate32, 3), "fp32_ms_per_iter": round(1000 * el32 / args.steps, 4),
        "speedup_bf16_over_fp32": round(rate / rate32, 4),
        "final_relerr_bf16": float(rel[-1]), "final_relerr_fp32": float(rel32[-1]),
        "bf16_conversion_ms_once": round(convert_ms, 3),
        "note": "A's BF16 copy with column scales, its transpose (the reference's At, formed "
                "once in the MUNMF constructor) and ||A||^2 (computed once in the NMF "
                "constructor) are made before the timed region, as in the reference",
    }
    if ref is not None:
        comparison.update({
            "final_relerr_fp64_oracle": ref[0],
            "bf16_minus_fp64": float(rel[-1]) - ref[0],
            "fp32_minus_fp64": float(rel32[-1]) - ref[0],
            "max_abs_traj_diff_bf16_vs_fp64": float(np.abs(rel - np.array(ref[1])).max()),
            "equal_error_within_1e-3": bool(abs(rel[-1] - ref[0]) < 1e-3 and abs(rel32[-1] - ref[0]) < 1e-3),
            "fp64_source": "tests/golden/nmf_c3a_fp64.json (FP64 oracle, same input and init)",
        })

    # ---- dominant kernel roofline ----------------------------------------
    t_wta = event_time_ms(torch, wta, 20)
    bytes_wta = 2 * m * n + 4 * m * k + 4 * n + 4 * k * n
    achieved = bytes_wta / (t_wta * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        try:
            traffic = json.load(open(TRAFFIC_FILE)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    probe = torch.zeros(1 << 28, dtype=torch.float32, device="cuda")
    read_ceiling = ctx.read_bandwidth(probe, 20)
    del probe
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "BF16 W^T A (thin_gemm TN: tn_h_sh_kernel / tn_h_kernel as tuned; "
                          "v_mfma_f32_16x16x32_bf16, A streamed from HBM into MFMA operands)",
                "algorithmic_bytes": bytes_wta,
                "algorithmic_bytes_formula": "2*m*n (A in BF16) + 4*m*k (W) + 4*n (column "
                                             "scales) + 4*k*n (output)",
                "avg_ms": round(t_wta, 5),
                "achievable_read_GBs": round(read_ceiling, 1),
                "frac_of_achievable": round(achieved / read_ceiling, 4)}

    # ---- CPU baseline ------------------------------------------------------
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        cpu = cpu_baseline_leg(O, A, Wi, Hi, args.cpu_iters)

    legs = {}
    if args.no_legs:
        legs_on = set()
    if "hals" in legs_on:
        log("bench: NMF-HALS fp32 vs bf16")
        legs["nmf_hals_65536x16384_k64"] = compare_precisions(
            P, ctx, torch, A, Wi, Hi, A16, cs, A16T, normA2, P.HALS, "hals", args.warmup, args.steps)
    if "bpp" in legs_on:
        log("bench: NMF-ANLS-BPP fp32 vs bf16")
        legs["nmf_anlsbpp_65536x16384_k64"] = compare_precisions(
            P, ctx, torch, A, Wi, Hi, A16, cs, A16T, normA2, P.ANLSBPP, "anlsbpp", 1, args.steps)
    del A, A16, A16T, cs, Wi, Hi, Wt
    torch.cuda.empty_cache()
    if "ntf4" in legs_on:
        log("bench: NTF-HALS 128^4 k=48 fp32 vs fp16")
        dims4, k4 = (128, 128, 128, 128), 48
        X4, init4 = make_ntf_problem(P, D, ctx, torch, dims4, k4, 48, stream=5)
        nt = {}
        r32, el32, rel32n = ntf_rate(P, ctx, torch, X4, init4, 3, 20, P.F32)
        nt["fp32"] = {"iters_per_s": round(r32, 3), "final_relerr": float(rel32n[-1])}
        # The 16-bit copy and its dimension-tree transpose, made once before
        # the timed region (as the NMF legs make A16 and At).
        X416, X416T = ctx.ntf_copies16(X4, P.F16)
        r16, el16, rel16 = ntf_rate(P, ctx, torch, X4, init4, 3, 20, P.F16, X16=X416, X16T=X416T)
        nt["fp16"] = {"iters_per_s": round(r16, 3), "final_relerr": float(rel16[-1]),
                      "relerr_gap_vs_fp32": abs(float(rel16[-1]) - float(rel32n[-1]))}
        nt["speedup_fp16_over_fp32"] = round(r16 / r32, 4)
        nt.update(ntf_fp64_compare("128^4", 20, {"fp32": rel32n, "fp16": rel16}))
        F4 = [P.colmajor(f) for f in init4]
        L4 = dims4[0] * dims4[1]
        P4 = P.empty_colmajor((L4, k4))
        numel4 = int(np.prod(dims4))
        # Right partial X_(01) KRP(A3, A2); left partial X_(01)^T KRP(A1, A0):
        # FP32 as a TN product on X, FP16 as an NN product on the transposed
        # copy (the right partial of the mode-permuted tensor).
        F4T = [F4[2], F4[3], F4[0], F4[1]]
        for name, dt, Xr, Fr, Xl, Fl, sl in (("fp32", P.F32, X4, F4, X4, F4, 1),
                                             ("fp16", P.F16, X416, F4, X416T, F4T, 0)):
            esz = 4 if dt == P.F32 else 2
            for part, Xop, Fop, side in (("right", Xr, Fr, 0), ("left", Xl, Fl, sl)):
                def r4(Xop=Xop, Fop=Fop, side=side, dt=dt):
                    ctx.mttkrp_partial(Xop, Fop, 1, side, dtype=dt, out=P4)
                t4 = event_time_ms(torch, r4, 10)
                b4 = esz * numel4 + 4 * L4 * k4
                nt[f"{part}_partial_mttkrp_{name}"] = {
                    "avg_ms": round(t4, 4), "achieved_GBs": round(b4 / (t4 * 1e-3) / 1e9, 1),
                    "frac_of_8TBs": round(b4 / (t4 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        nt["config"] = ("128^4 CP rank-48 model + 1% noise, k=48, dimension tree, 20 timed iterations "
                        "(median of 5 timed runs)")
        legs["ntf_hals_128^4_k48"] = nt
        del X4, X416, X416T, F4, P4
        torch.cuda.empty_cache()
    if "ntf3" in legs_on:
        log("bench: NTF-HALS 256^3 k=32")
        dims, k3 = (256, 256, 256), 32
        X, init = make_ntf_problem(P, D, ctx, torch, dims, k3, 32, stream=2)
        r3, el3, rel3 = ntf_rate(P, ctx, torch, X, init, 3, 30, P.F32)
        c2 = {"fp32": {"iters_per_s": round(r3, 3), "final_relerr": float(rel3[-1])}}
        trajs = {"fp32": rel3}
        for name, dt in (("bf16", P.BF16), ("fp16", P.F16)):
            X16, X16T = ctx.ntf_copies16(X, dt)
            r, _, rr = ntf_rate(P, ctx, torch, X, init, 3, 30, dt, X16=X16, X16T=X16T)
            c2[name] = {"iters_per_s": round(r, 3), "final_relerr": float(rr[-1])}
            trajs[name] = rr
            del X16, X16T
        c2.update(ntf_fp64_compare("256^3", 30, trajs))
        c2["config"] = ("256^3 CP rank-32 model + 1% noise, k=32, dimension tree, 30 timed iterations "
                        "(median of 5 timed runs)")
        legs["ntf_hals_256^3_k32"] = c2
        del X
        torch.cuda.empty_cache()
    if "hier" in legs_on:
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
        "dtype": "bf16",
        "data": "synthetic (Philox4x64: A = W0 H0 with rank-64 U[0,1) factors + 1% |N(0,1)| "
                "noise, generated on the device; seeded U[0,1) init)",
        "config": {"workload": f"NMF-MU {m}x{n} FP32 A, k={k}, BF16-MFMA products (FP32 "
                               f"accumulation, per-column scales), FP64 Grams and relative "
                               f"error every iteration, {args.steps} iterations",
                   "m": m, "n": n, "k": k, "algorithm": "MU", "iterations": args.steps},
        "bf16_vs_fp32": comparison,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "legs": legs,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
