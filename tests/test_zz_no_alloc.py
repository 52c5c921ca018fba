"""GPU test of the allocation-free CP-loss forward (planc_cp_loss_fwd_noalloc)
through the C ABI: two calls in a row, the second on a smaller tensor, each
checked against planc_cp_loss_fwd on the same input (same block partition
and summation order, so the two losses are bit-identical) and against the
FP64 oracle (relative 1e-5, FP32 data)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import planc_amd as P  # noqa: E402
from oracle import oracle as O  # noqa: E402
from planc_amd import datagen as D  # noqa: E402


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _case(dims, k, stream):
    F = [D.uniform(D.DATA_SEED, 16 * stream + i, (d, k)) for i, d in enumerate(dims)]
    subs = "abcdefg"[:len(dims)]
    X = np.einsum(",".join(s + "z" for s in subs) + "->" + subs, *F, optimize=True)
    X = X + 0.01 * X.mean() * np.abs(D.normal(D.NOISE_SEED, stream, dims))
    G = [D.uniform(D.INIT_SEED, 16 * stream + i, (d, k)) for i, d in enumerate(dims)]
    return np.asfortranarray(X.astype(np.float32)), [g.astype(np.float32) for g in G]


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = P.Context(0)
    yield c
    c.close()


def test_noalloc_forward_matches_allocating_forward(ctx):
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    for dims, k, stream in (((40, 36, 32, 28), 48, 311), ((20, 18, 16), 8, 312)):
        say(f"cp_loss_fwd_noalloc dims={dims} k={k}")
        X, F = _case(dims, k, stream)
        dX, dF = P.colmajor(X), [P.colmajor(f) for f in F]
        ref = ctx.cp_loss(dX, dF)
        ctx.cp_loss_noalloc(dX, dF, out)
        got = float(out.item())
        loss_r = O.cp_loss(X.astype(np.float64), [f.astype(np.float64) for f in F])
        say(f"  noalloc {got:.17e} allocating {ref:.17e} oracle {loss_r:.17e}")
        assert got == ref
        assert abs(got - loss_r) / loss_r < 1e-5


def test_noalloc_forward_rejects_oversized_partition(ctx):
    """A two-mode tensor whose right block has 2.2 million rows needs more
    block partials than the fixed array holds: the host check refuses it
    before any launch."""
    dims, k = (1, 2_200_000), 1
    dX = torch.ones(dims[1], dtype=torch.float32, device="cuda")[None, :]
    dF = [torch.ones((d, k), dtype=torch.float32, device="cuda") for d in dims]
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(P.PlancError) as e:
        ctx.cp_loss_noalloc(dX, dF, out)
    assert "exceed" in str(e.value)
    torch.cuda.synchronize()
    # The context still works afterwards.
    X, F = _case((6, 5, 4), 3, 313)
    dX, dF = P.colmajor(X), [P.colmajor(f) for f in F]
    ctx.cp_loss_noalloc(dX, dF, out)
    assert float(out.item()) == ctx.cp_loss(dX, dF)
