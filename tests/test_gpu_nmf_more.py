"""More GPU parity tests of the NMF loop and of kernel-level entries, through
the C ABI against the FP64 oracle:

* regularised MU / HALS / ANLS-BPP (applyReg, common/nmf.hpp:76-90) and the
  symmetric ANLS-BPP regularisation (applySymmetricReg, common/nmf.hpp:
  123-129, nmf/bppnmf.hpp:40-44): trajectories 1e-5 (BPP 1e-4);
* one-iteration factor parity of the 16-bit forms (||dW||/||W|| <= 1e-2) and
  ANLS-BPP with BF16/FP16 products (final relative error within 1e-3 of the
  FP64 oracle);
* the mixed-precision configuration at its full size (A 65536 x 16384, k = 64,
  50 iterations): BF16 MU, HALS and ANLS-BPP against the FP64 oracle's
  trajectories (tests/golden/nmf_c3a_fp64.json, same input and init), the
  BF16 - FP32 gap, monotone error, and the KKT conditions of the last
  ANLS-BPP half-step;
* entries no other test calls directly: planc_normalize_cols,
  planc_hadamard_grams, planc_gen_lowrank, planc_transpose16 (including the
  slab loop beyond 65535 column tiles).
"""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import planc_amd as P  # noqa: E402
from oracle import oracle as O  # noqa: E402
from planc_amd import datagen as D  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(x, ref):
    return np.linalg.norm(np.asarray(x, np.float64) - ref) / max(np.linalg.norm(ref), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = P.Context(0)
    yield c
    c.close()


def _nmf_case(m, n, k, stream):
    A = D.lowrank_matrix(m, n, k, 0.01, stream=stream).astype(np.float32)
    W0, H0 = D.init_factors([(m, k), (n, k)], stream=stream)
    return A, W0.astype(np.float32), H0.astype(np.float32)


@pytest.mark.parametrize("algo", [P.MU, P.HALS, P.ANLSBPP])
def test_nmf_regularised_matches_oracle(ctx, algo):
    """regW = {L2 0.5, L1 0.1}, regH = {L2 0.25, L1 0}: the device applyReg
    (Gram + l2 I + l1 11^T, values rounded to float as the reference's FVEC)
    against the oracle's trajectory and factors."""
    say(f"nmf regularised algo={algo}")
    A, W0, H0 = _nmf_case(500, 300, 10, stream=21)
    regW, regH = (0.5, 0.1), (0.25, 0.0)
    Wr, Hr, rel_r, _, rc = O.nmf(algo, A.astype(np.float64), W0.astype(np.float64),
                                 H0.astype(np.float64), 10, regW=regW, regH=regH)
    assert rc == 0
    dA, dW, dH = P.colmajor(A), P.colmajor(W0), P.colmajor(H0)
    rel = ctx.nmf_run(dA, dW, dH, algo, 10, regW=regW, regH=regH)
    say(f"  gpu {rel[-1]:.10f} oracle {rel_r[-1]:.10f}")
    tol = 1e-4 if algo == P.ANLSBPP else 1e-5
    assert np.abs(rel - rel_r).max() < tol
    ftol = 1e-3 if algo == P.ANLSBPP else 1e-4
    assert rel_err(dW.cpu().numpy(), Wr) < ftol
    assert rel_err(dH.cpu().numpy(), Hr) < ftol
    # the regularisation changes the answer (a dropped term would not pass)
    _, _, rel_0, _, _ = O.nmf(algo, A.astype(np.float64), W0.astype(np.float64),
                              H0.astype(np.float64), 10)
    assert np.abs(rel_0 - rel_r).max() > 10 * tol


def test_nmf_symmetric_bpp_matches_oracle(ctx):
    """Symmetric ANLS-BPP (--symm): Gram + s I, right-hand side + s given^T, on
    a square symmetric A = B B^T, from the reference driver's symmetric init
    (H scaled by 2 sqrt(mean(A)/k), W = H; nmf/nmf.cpp:191-199)."""
    n, k, s = 300, 8, 4.0
    B = D.uniform(D.DATA_SEED, 16 * 22, (n, k))
    A = (B @ B.T).astype(np.float32)
    H0 = D.uniform(D.INIT_SEED, 16 * 22 + 1, (n, k))
    H0 = (2 * np.sqrt(A.astype(np.float64).mean() / k) * H0).astype(np.float32)
    W0 = H0.copy()
    Wr, Hr, rel_r, _, rc = O.nmf(2, A.astype(np.float64), W0.astype(np.float64),
                                 H0.astype(np.float64), 8, symm_reg=s)
    assert rc == 0
    dA, dW, dH = P.colmajor(A), P.colmajor(W0), P.colmajor(H0)
    rel = ctx.nmf_run(dA, dW, dH, P.ANLSBPP, 8, symm_reg=s)
    say(f"  symm gpu {rel[-1]:.10f} oracle {rel_r[-1]:.10f}")
    assert np.abs(rel - rel_r).max() < 1e-4
    assert rel_err(dW.cpu().numpy(), Wr) < 1e-3
    assert rel_err(dH.cpu().numpy(), Hr) < 1e-3
    _, _, rel_0, _, _ = O.nmf(2, A.astype(np.float64), W0.astype(np.float64),
                              H0.astype(np.float64), 8)
    assert np.abs(rel_0 - rel_r).max() > 1e-3
    # a non-square A is refused
    with pytest.raises(P.PlancError):
        ctx.nmf_run(P.colmajor(A[:, :200]), P.colmajor(W0), P.colmajor(H0[:200]), P.ANLSBPP, 1,
                    symm_reg=s)


@pytest.mark.parametrize("dt", [P.BF16, P.F16])
@pytest.mark.parametrize("algo", [P.MU, P.HALS, P.ANLSBPP])
def test_nmf_one_iteration_factors_16bit(ctx, algo, dt):
    """Factors after one iteration with 16-bit products: ||dW||/||W|| and
    ||dH||/||H|| <= 1e-2 against the FP64 oracle.  ANLS-BPP takes the factor
    into its 16-bit products as two planes, so only A is rounded: its factors
    also match the FP64 oracle run on A rounded to the 16-bit format within
    1e-3 (the column scales are powers of two and do not change the
    rounding)."""
    say(f"nmf one iteration 16-bit algo={algo} dt={dt}")
    A, W0, H0 = _nmf_case(2000, 1000, 20, stream=1)
    Wr, Hr, _, _, rc = O.nmf(algo, A.astype(np.float64), W0.astype(np.float64),
                             H0.astype(np.float64), 1)
    assert rc == 0
    dA, dW, dH = P.colmajor(A), P.colmajor(W0), P.colmajor(H0)
    ctx.nmf_run(dA, dW, dH, algo, 1, dtype=dt)
    ew, eh = rel_err(dW.cpu().numpy(), Wr), rel_err(dH.cpu().numpy(), Hr)
    say(f"  dW {ew:.2e} dH {eh:.2e}")
    assert ew < 1e-2 and eh < 1e-2, (ew, eh)
    if algo == P.ANLSBPP:
        Aq = _round16(A.ravel(order="F"), dt).reshape(A.shape, order="F")
        Wq, Hq, _, _, rcq = O.nmf(algo, np.asfortranarray(Aq), W0.astype(np.float64),
                                  H0.astype(np.float64), 1)
        assert rcq == 0
        ewq, ehq = rel_err(dW.cpu().numpy(), Wq), rel_err(dH.cpu().numpy(), Hq)
        say(f"  against the oracle on rounded A: dW {ewq:.2e} dH {ehq:.2e}")
        assert ewq < 1e-3 and ehq < 1e-3, (ewq, ehq)


@pytest.mark.parametrize("dt", [P.BF16, P.F16])
def test_nmf_bpp_16bit_final_error(ctx, dt):
    """ANLS-BPP with 16-bit products ends within 1e-3 of the FP64 oracle's
    final relative error (A 2000 x 1000, k = 20, 20 iterations)."""
    A, W0, H0 = _nmf_case(2000, 1000, 20, stream=1)
    _, _, rel_r, _, rc = O.nmf(2, A.astype(np.float64), W0.astype(np.float64),
                               H0.astype(np.float64), 20)
    assert rc == 0
    dA, dW, dH = P.colmajor(A), P.colmajor(W0), P.colmajor(H0)
    rel = ctx.nmf_run(dA, dW, dH, P.ANLSBPP, 20, dtype=dt)
    say(f"  bpp dt={dt} gpu {rel[-1]:.6f} oracle {rel_r[-1]:.6f}")
    assert abs(rel[-1] - rel_r[-1]) < 1e-3


def _round16(a, dt):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if dt == P.F16:
        return a.astype(np.float16).astype(np.float64)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("dt", [P.F32, P.BF16, P.F16])
@pytest.mark.parametrize("algo", [P.MU, P.HALS, P.ANLSBPP])
def test_ntf_one_iteration_factors(ctx, algo, dt):
    """NTF factors and lambda after one iteration (dimension tree) against the
    FP64 oracle.  FP32: 1e-5 for MU; HALS and BPP 2e-4 -- an FP32 restatement
    of the same iteration (FP32 factors and MTTKRP, FP64 update arithmetic)
    already moves the HALS factors by up to 1.1e-4 on this input, the sweep
    amplifying storage rounding.  16-bit: the MTTKRPs round only X (the
    Khatri-Rao rows enter as two 16-bit planes), so the factors match the FP64
    oracle run on X rounded to the 16-bit format within the FP32 bound (1e-3),
    and the FP64 oracle on the unrounded X within 1e-2 or 1.5 times the
    oracle's own change under that rounding, whichever is larger (with a
    one-plane Khatri-Rao operand BF16 HALS moves mode 2 by 4.4e-2 here)."""
    dims, k = (40, 36, 32), 8
    X, _ = D.cp_tensor(dims, k, 0.01, stream=23)
    X32 = X.astype(np.float32)
    F0 = [f.astype(np.float32) for f in D.init_factors([(d, k) for d in dims], stream=23)]
    F64 = [f.astype(np.float64) for f in F0]
    Fr, lam_r, rel_r, rc = O.ntf(algo, X32.astype(np.float64), F64, 1, dimtree=1)
    assert rc == 0
    dX = P.colmajor(X32)
    dF = [P.colmajor(f) for f in F0]
    lam, rel = ctx.ntf_run(dX, dF, algo, 1, dtype=dt)
    errs = [rel_err(f.cpu().numpy(), fr) for f, fr in zip(dF, Fr)]
    say(f"  ntf one iteration algo={algo} dt={dt}: factor errors {errs}")
    if dt == P.F32:
        tol = 1e-5 if algo == P.MU else 2e-4
    else:
        Xq = np.asfortranarray(_round16(X32.ravel(order="F"), dt).reshape(dims, order="F"))
        Fq, lam_q, _, rcq = O.ntf(algo, Xq, F64, 1, dimtree=1)
        assert rcq == 0
        sens = max(rel_err(a, b) for a, b in zip(Fq, Fr))
        tol = max(1e-2, 1.5 * sens)
        errq = [rel_err(f.cpu().numpy(), fq) for f, fq in zip(dF, Fq)]
        say(f"  oracle change under X rounded: {sens:.2e} -> tolerance {tol:.2e}; "
            f"against the oracle on rounded X: {errq}")
        assert max(errq) < 1e-3
        assert rel_err(lam.cpu().numpy(), lam_q) < 1e-3
    assert max(errs) < tol
    assert rel_err(lam.cpu().numpy(), lam_r) < tol


@pytest.mark.parametrize("dims", [(40, 36, 32), (36, 40, 30)])  # NN right contraction / transposed copy (L % 8 != 0)
@pytest.mark.parametrize("dt", [P.BF16, P.F16])
def test_ntf_16bit_copy_with_column_scales(ctx, dt, dims):
    """The 16-bit copy planc_ntf_run makes itself carries per-column
    power-of-two scales: a CP tensor times 2^20 (far beyond FP16's 65504) with
    its mode-2 slices further scaled by 2^0 .. 2^6 (columns of the tree's
    unfolding six binades apart; still a rank-k CP tensor) runs in FP16 and
    BF16 with finite errors whose final value is within 1e-3 of the FP64
    oracle's on the same tensor."""
    k = 8
    X, _ = D.cp_tensor(dims, k, 0.01, stream=41)
    slab = 2.0 ** (20 + np.arange(dims[2]) % 7)
    Xs = (X * slab[None, None, :]).astype(np.float32)
    assert np.abs(Xs).max() > 65504 * 64
    F0 = [f.astype(np.float32) for f in D.init_factors([(d, k) for d in dims], stream=41)]
    _, _, rel_r, _ = O.ntf(P.HALS, Xs.astype(np.float64), [f.astype(np.float64) for f in F0], 10, dimtree=1)
    dF = [P.colmajor(f) for f in F0]
    lam, rel = ctx.ntf_run(P.colmajor(Xs), dF, P.HALS, 10, dtype=dt)
    say(f"  dt={dt} gpu {rel} oracle {rel_r}")
    assert np.all(np.isfinite(rel)) and np.all(np.isfinite(lam.cpu().numpy()))
    assert abs(rel[-1] - rel_r[-1]) < 1e-3
    assert np.abs(rel - rel_r).max() < 1e-2


# ---- the mixed-precision configuration at full size -------------------------
M3, N3, K3 = 65536, 16384, 64


def _c3a_fixture():
    path = os.path.join(HERE, "golden", "nmf_c3a_fp64.json")
    return json.load(open(path))


@pytest.fixture(scope="module")
def c3a(ctx):
    say("c3a: generating A 65536 x 16384 on the device")
    stream = 3
    W0 = P.colmajor(D.uniform(D.DATA_SEED, 16 * stream + 0, (M3, K3)).astype(np.float32))
    H0 = P.colmajor(D.uniform(D.DATA_SEED, 16 * stream + 1, (K3, N3)).astype(np.float32))
    mean = float((W0.double().mean(0) * H0.double().mean(1)).sum().item())
    A = ctx.gen_lowrank(W0, H0, 0.01 * mean, D.NOISE_SEED, stream)
    del W0, H0
    Wi = ctx.gen_uniform((M3, K3), D.INIT_SEED, 16 * stream + 0)
    Hi = ctx.gen_uniform((N3, K3), D.INIT_SEED, 16 * stream + 1)
    A16, cs = ctx.convert(A, P.BF16, scaled=True)
    A16T = ctx.transpose16(A16)
    yield A, Wi, Hi, A16, cs, A16T
    del A, A16, A16T
    torch.cuda.empty_cache()


def test_c3a_input_matches_fixture_input(c3a):
    """The device-generated A is the input the FP64 fixture was computed on
    (host restatement of planc_gen_lowrank): equal FP64 sums to 1e-9."""
    A = c3a[0]
    fx = _c3a_fixture()
    s = sum(float(A[:, j:j + 2048].double().sum().item()) for j in range(0, N3, 2048))
    say(f"  sum(A) device {s:.6f} fixture {fx['sum_A_fp64']:.6f}")
    assert abs(s - fx["sum_A_fp64"]) / fx["sum_A_fp64"] < 1e-9


@pytest.mark.parametrize("name,algo", [("mu", P.MU), ("hals", P.HALS), ("anlsbpp", P.ANLSBPP)])
def test_c3a_full_size_bf16(ctx, c3a, name, algo):
    fx = _c3a_fixture()
    if name not in fx:
        pytest.skip(f"no FP64 trajectory for {name} in the fixture")
    A, Wi, Hi, A16, cs, A16T = c3a
    iters = int(fx["iters"])
    ref = np.array(fx[name])
    W, H = Wi.clone(), Hi.clone()
    rel16 = ctx.nmf_run(A, W, H, algo, iters, dtype=P.BF16, A16=A16, colscale=cs, A16T=A16T)
    W32, H32 = Wi.clone(), Hi.clone()
    rel32 = ctx.nmf_run(A, W32, H32, algo, iters)
    say(f"  c3a {name}: bf16 {rel16[-1]:.8f} fp32 {rel32[-1]:.8f} fp64 {ref[-1]:.8f}")
    assert abs(rel16[-1] - ref[-1]) < 1e-3
    assert abs(rel32[-1] - ref[-1]) < 1e-3
    assert abs(rel16[-1] - rel32[-1]) < 1e-3
    assert np.abs(rel32 - ref).max() < 1e-4
    assert np.all(np.diff(rel32) <= 1e-7), np.diff(rel32).max()
    if algo != P.HALS:
        # (BF16 HALS is not monotone: its sweeps use rounded products, and
        # near the limit the error moves up and down by ~2e-5)
        assert np.all(np.diff(rel16) <= 1e-6), np.diff(rel16).max()
    if algo == P.ANLSBPP:
        # KKT of the last H half-step (normalize_by_W keeps it: H D solves the
        # NNLS for W D^-1): H >= 0, Y = (W^T W) H^T - W^T A >= 0 where H = 0,
        # Y = 0 where H > 0, with W^T A the BF16 product the solver used.
        Wn = W.cpu().numpy().astype(np.float64)
        Hn = H.cpu().numpy().astype(np.float64)
        Bt = ctx.gemm_tn(A16, W, dtype=P.BF16, colscale=cs).cpu().numpy().astype(np.float64)
        Y = (Wn.T @ Wn) @ Hn.T - Bt
        scale = np.abs(Bt).max(axis=0) + 1e-30
        assert Hn.min() >= 0
        pas = Hn.T > 0
        viol_p = np.abs(Y[pas] / np.broadcast_to(scale, Y.shape)[pas]).max()
        viol_a = (-Y[~pas] / np.broadcast_to(scale, Y.shape)[~pas]).max(initial=0.0)
        say(f"  KKT passive {viol_p:.2e} active {viol_a:.2e}")
        assert viol_p < 1e-3 and viol_a < 1e-3


# ---- kernel-level entries ---------------------------------------------------
def test_normalize_cols_entry(ctx):
    """NCPFactors::normalize(mode): unit columns, lambda = the norms; a zero
    column stays zero with lambda 0."""
    F = D.uniform(D.INIT_SEED, 16 * 24, (300, 7)).astype(np.float32)
    F[:, 3] = 0
    dF = P.colmajor(F)
    lam = ctx.normalize_cols(dF).cpu().numpy()
    nr = np.linalg.norm(F.astype(np.float64), axis=0)
    assert np.allclose(lam, nr, rtol=1e-12, atol=0)
    out = dF.cpu().numpy()
    ref = F.astype(np.float64) / np.where(nr > 0, nr, 1.0)
    assert rel_err(out, ref) < 1e-6
    assert np.all(out[:, 3] == 0)


def test_r2_update_entry(ctx):
    """DistR2::update (distnmf/distr2.hpp:35-61) against the oracle's r2_update
    (itself pinned by brute force over the four active sets): columns with a
    non-negative solution, with either variable negative, and all-zero
    columns."""
    rng = np.random.default_rng(77)
    F = rng.random((50, 2))
    G = F.T @ F
    n = 4000
    R = np.empty((2, n))
    R[:, : n // 2] = rng.random((2, n // 2))                  # mostly feasible
    R[0, n // 2:] = rng.random(n // 2)                         # second variable pushed negative
    R[1, n // 2:] = -rng.random(n // 2) * 0.5
    R[:, -7:] = 0.0
    R[:, 100:200] = R[::-1, 100:200] * np.array([[1.0], [-1.0]])  # first variable negative
    R32 = R.astype(np.float32)
    out = ctx.r2_update(P.colmajor(G, dtype=np.float64), P.colmajor(R32)).cpu().numpy().astype(np.float64)
    ref = O.r2_update(G, R32.astype(np.float64))
    assert np.all(out >= 0)
    assert np.allclose(out, ref, rtol=1e-6, atol=1e-30)
    assert np.all(out[:, -7:] == 0)
    assert np.all((out[0] == 0) | (out[1] == 0) | (ref.min(axis=0) > 0))


def test_rel_error_gram_entry(ctx):
    """computeObjectiveError from the trace identity equals the direct
    ||A - W H^T|| / ||A||; a non-positive fit prints 1 (common/nmf.hpp:462)."""
    torch = pytest.importorskip("torch")
    m, n, k = 700, 500, 12
    A = D.lowrank_matrix(m, n, k, 0.01, stream=31).astype(np.float64)
    W = D.uniform(D.INIT_SEED, 611, (m, k))
    H = D.uniform(D.INIT_SEED, 612, (n, k))
    direct = np.linalg.norm(A - W @ H.T) / np.linalg.norm(A)
    dev = lambda v: torch.tensor(np.atleast_1d(v), dtype=torch.float64, device="cuda")
    WtW, HtH = P.colmajor(W.T @ W, dtype=np.float64), P.colmajor(H.T @ H, dtype=np.float64)
    rel = ctx.rel_error_gram(dev(np.sum(A * A)), dev(np.sum(W * (A @ H))), WtW, HtH)
    assert abs(rel - direct) <= 1e-10 * direct
    tr = float(np.sum((W.T @ W) * (H.T @ H)))
    assert ctx.rel_error_gram(dev(1.0), dev(0.5 * (1.0 + tr) + 1.0), WtW, HtH) == 1.0  # fit^2 = -2


def test_hadamard_grams_entry(ctx):
    """gram_leave_out_one: Hadamard product of the Grams of every mode but one."""
    k = 9
    G = [D.uniform(D.INIT_SEED, 16 * 25 + i, (k, k)) for i in range(4)]
    dG = [P.colmajor(g, dtype=np.float64) for g in G]
    for skip in range(4):
        out = ctx.hadamard_grams(dG, skip).cpu().numpy()
        ref = np.ones((k, k))
        for i in range(4):
            if i != skip:
                ref = ref * G[i]
        assert np.allclose(out, ref, rtol=1e-15, atol=0)


def test_gen_lowrank_entry(ctx):
    """planc_gen_lowrank against its host restatement (FP64 product, Box-Muller
    normal of the same Philox stream), FP32 rounding."""
    m, n, r, stream = 300, 200, 6, 26
    W0 = D.uniform(D.DATA_SEED, 16 * stream, (m, r)).astype(np.float32)
    H0 = D.uniform(D.DATA_SEED, 16 * stream + 1, (r, n)).astype(np.float32)
    A = ctx.gen_lowrank(P.colmajor(W0), P.colmajor(H0), 0.05, D.NOISE_SEED, stream).cpu().numpy()
    ref = W0.astype(np.float64) @ H0.astype(np.float64) + 0.05 * np.abs(
        D.normal(D.NOISE_SEED, stream, (m, n)))
    assert np.abs(A - ref).max() / np.abs(ref).max() < 1e-6


@pytest.mark.parametrize("rows,cols", [(40, 30), (64, 128), (8, 65535 * 64 + 40), (37, 5)])
def test_transpose16_entry(ctx, rows, cols):
    """16-bit transpose, bit-exact, including a matrix wider than 65535 tiles
    of 64 columns (two slabs) and shapes that take the unvectorised kernel."""
    a = (np.arange(rows * cols, dtype=np.int64) % 65521).astype(np.int16).reshape((rows, cols), order="F")
    d = torch.from_numpy(np.asfortranarray(a)).cuda()
    t = ctx.transpose16(d).cpu().numpy()
    assert t.shape == (cols, rows)
    assert np.array_equal(t, a.T)
