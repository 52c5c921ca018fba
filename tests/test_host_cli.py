"""The C++ host API and the dense_nmf / dense_ntf drivers (drop-in CLIs).

CPU: flag parsing (--dry-run prints the parsed configuration without touching
the GPU) and the loud failure when no device is present.
GPU: the drivers' relative-error trajectories against the FP64 oracle on the
same generated input (the drivers draw A and the initial factors from the
Philox streams that planc_amd/datagen.py reproduces).
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "planc_amd", "host", "bin")


def run(args, check=True):
    p = subprocess.run(args, capture_output=True, text=True, timeout=300)
    if check and p.returncode != 0:
        raise AssertionError(f"{args} failed ({p.returncode}):\n{p.stdout}\n{p.stderr}")
    return p


def have_bins():
    return os.path.exists(os.path.join(BIN, "dense_nmf")) and os.path.exists(os.path.join(BIN, "dense_ntf"))


@pytest.mark.skipif(not have_bins(), reason="host drivers not built")
def test_dense_nmf_flags_dry_run():
    p = run([os.path.join(BIN, "dense_nmf"), "-a", "1", "-d", "300 200", "-k", "8", "-t", "5",
             "-i", "rand_lowrank", "-r", "0.5 0.1 0.25 0", "--seed", "7", "--precision", "bf16",
             "--dry-run"])
    out = p.stdout
    assert "a::1::i::rand_lowrank::k::8::m::300::n::200::t::5" in out
    assert "initseed::7" in out and "precision::bf16" in out
    assert "regW::0.5,0.1::regH::0.25,0" in out


@pytest.mark.skipif(not have_bins(), reason="host drivers not built")
def test_dense_ntf_defaults_dry_run():
    p = run([os.path.join(BIN, "dense_ntf"), "-d", "10 11 12", "-i", "rand_uniform", "--dry-run"])
    # reference defaults: algo ANLSBPP (2), k 20, 20 iterations, dimtree on
    assert "a::2::" in p.stdout and "::k::20::" in p.stdout and "::t::20::" in p.stdout
    assert "dimtree::1" in p.stdout


@pytest.mark.skipif(not have_bins(), reason="host drivers not built")
def test_missing_input_is_an_error():
    p = run([os.path.join(BIN, "dense_nmf"), "-d", "3 3"], check=False)
    assert p.returncode != 0 and "Input not given" in p.stderr


def _relerrs(stdout):
    return np.array([float(x) for x in re.findall(r"relative error::([0-9.eE+-]+)", stdout)])


@pytest.mark.gpu
@pytest.mark.parametrize("algo", [0, 1, 2])
def test_dense_nmf_matches_oracle(tmp_path, algo):
    from oracle import oracle as O
    from planc_amd import datagen as D
    m, n, k, t = 300, 200, 8, 10
    prefix = str(tmp_path / "out")
    p = run([os.path.join(BIN, "dense_nmf"), "-a", str(algo), "-d", f"{m} {n}", "-k", str(k), "-t",
             str(t), "-i", "rand_lowrank", "-o", prefix])
    rel = _relerrs(p.stdout)
    assert rel.shape == (t,)
    Wt = D.uniform(1210873, 1, (m, k))
    Ht = D.uniform(1210873, 2, (k, n))
    A = (Wt @ Ht).astype(np.float32).astype(np.float64)
    W0 = D.uniform(193957, 0, (m, k)).astype(np.float32).astype(np.float64)
    H0 = D.uniform(193957, 1, (n, k)).astype(np.float32).astype(np.float64)
    Wr, Hr, rel_r, _, rc = O.nmf(algo, A, W0, H0, t)
    assert np.abs(rel - rel_r).max() < 1e-4
    W = np.loadtxt(prefix + "_W")
    H = np.loadtxt(prefix + "_H")
    assert W.shape == (m, k) and H.shape == (n, k)
    assert np.linalg.norm(W - Wr) / np.linalg.norm(Wr) < 1e-3


@pytest.mark.skipif(not have_bins(), reason="host drivers not built")
def test_dense_nmf_normalization_and_adjrand_parsed():
    p = run([os.path.join(BIN, "dense_nmf"), "-a", "0", "-d", "30 20", "-k", "4", "-i", "rand_lowrank",
             "--normalization", "l2", "--adjrand", "--dry-run"])
    assert "input normalization::1" in p.stdout and "adjrand::1" in p.stdout
    p = run([os.path.join(BIN, "dense_nmf"), "-a", "0", "-d", "30 20", "-k", "4", "-i", "rand_lowrank",
             "--normalization", "max", "--dry-run"])
    assert "input normalization::2" in p.stdout and "adjrand::0" in p.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("norm,adj", [("l2", False), ("max", False), ("max", True)])
def test_dense_nmf_input_normalization_matches_oracle(tmp_path, norm, adj):
    """--normalization l2|max and --adjrand change A before the solve exactly as
    nmf/nmf.cpp:164-185 does (ceil(1*A + 0) for a generated A, then unit
    columns or A / max(A))."""
    from oracle import oracle as O
    from planc_amd import datagen as D
    m, n, k, t = 120, 90, 6, 6
    args = [os.path.join(BIN, "dense_nmf"), "-a", "0", "-d", f"{m} {n}", "-k", str(k), "-t", str(t),
            "-i", "rand_lowrank", "--normalization", norm]
    if adj:
        args.append("--adjrand")
    rel = _relerrs(run(args).stdout)
    Wt = D.uniform(1210873, 1, (m, k))
    Ht = D.uniform(1210873, 2, (k, n))
    A = Wt @ Ht
    if adj:
        A = np.ceil(A)
    if norm == "l2":
        A = A / np.linalg.norm(A, axis=0, keepdims=True)
    else:
        A = A * (1.0 / A.max())
    A = A.astype(np.float32).astype(np.float64)
    W0 = D.uniform(193957, 0, (m, k)).astype(np.float32).astype(np.float64)
    H0 = D.uniform(193957, 1, (n, k)).astype(np.float32).astype(np.float64)
    _, _, rel_r, _, rc = O.nmf(0, A, W0, H0, t)
    assert rc == 0
    assert np.abs(rel - rel_r).max() < 1e-4


@pytest.mark.gpu
def test_dense_ntf_runs_and_error_decreases():
    p = run([os.path.join(BIN, "dense_ntf"), "-a", "1", "-d", "30 31 32", "-k", "5", "-t", "8",
             "-i", "rand_uniform", "-e", "1"])
    rel = np.array([float(x) for x in re.findall(r"relative_error at it::\d+::([0-9.eE+-]+)",
                                                  p.stdout)])
    assert rel.shape == (8,)
    assert np.all(np.diff(rel) <= 1e-6)
