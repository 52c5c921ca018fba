"""Deterministic synthetic inputs (Philox4x64-10 through numpy.random.Philox).

Element e of stream (seed, stream) is the same 64-bit word that
planc_amd/csrc/philox.h produces on the device, so host and device
generators agree bit for bit on uniforms.  Matrices are filled in
column-major element order, tensors with mode 0 fastest.
"""
import numpy as np

DATA_SEED = 0x504C414E   # "PLAN"
INIT_SEED = 0x43494E49
NOISE_SEED = 0x45534F4E


def words(seed, stream, count, offset=0):
    bg = np.random.Philox(key=[seed, 0], counter=[0, stream, 0, 0])
    if offset:
        bg = bg.advance(offset // 4) if offset % 4 == 0 else None
        if bg is None:
            raise ValueError("offset must be a multiple of 4")
    return bg.random_raw(count)


def uniform(seed, stream, shape):
    """U[0,1) doubles (53 random bits), column-major fill."""
    n = int(np.prod(shape))
    w = words(seed, stream, n)
    u = (w >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return u.reshape(shape, order="F")


def normal(seed, stream, shape):
    """N(0,1) by Box-Muller from words 2e and 2e+1 (first uniform mapped to (0,1])."""
    n = int(np.prod(shape))
    w = words(seed, stream, 2 * n)
    u = (w >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return z.reshape(shape, order="F")


def lowrank_matrix(m, n, r, sigma, stream, relative=False):
    """A = W0 H0 + sigma*|N(0,1)| (relative=True: sigma * mean(W0 H0)), FP64."""
    W0 = uniform(DATA_SEED, 16 * stream + 0, (m, r))
    H0 = uniform(DATA_SEED, 16 * stream + 1, (r, n))
    A = W0 @ H0
    s = sigma * A.mean() if relative else sigma
    A = A + s * np.abs(normal(NOISE_SEED, stream, (m, n)))
    return np.asfortranarray(A)


def init_factors(shapes, stream):
    return [np.asfortranarray(uniform(INIT_SEED, 16 * stream + i, s)) for i, s in enumerate(shapes)]


def cp_tensor(dims, r, noise_rel, stream):
    """X = [[A_1..A_N]] + noise_rel*mean(X)*|N(0,1)|, mode 0 fastest, FP64."""
    F = [uniform(DATA_SEED, 16 * stream + i, (d, r)) for i, d in enumerate(dims)]
    N = len(dims)
    subs = "abcdefgh"[:N]
    X = np.einsum(",".join(s + "z" for s in subs) + "->" + subs, *F, optimize=True)
    X = X + noise_rel * X.mean() * np.abs(normal(NOISE_SEED, stream, dims))
    return np.asfortranarray(X), F


def gamma2(seed, stream, shape, scale=0.5):
    """Gamma(shape 2, scale) = -scale * (ln U1 + ln U2), from words 2e, 2e+1."""
    n = int(np.prod(shape))
    w = words(seed, stream, 2 * n)
    u = 1.0 - (w >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    g = -scale * (np.log(u[0::2]) + np.log(u[1::2]))
    return g.reshape(shape, order="F")


def sparse_powerlaw(m, n, density, alpha, stream):
    """CSR (indptr int64, indices int32 sorted per row, data float32) of an
    m x n matrix: every column draws round(density*m) rows with probability
    proportional to (i+1)^-alpha (inverse-CDF sampling of Philox uniforms),
    duplicates dropped; values exp(N(0,1)) (log-normal)."""
    per = max(1, int(round(density * m)))
    w = (np.arange(1, m + 1, dtype=np.float64)) ** (-alpha)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    u = uniform(DATA_SEED, 16 * stream + 3, (n * per,))
    rows = np.minimum(np.searchsorted(cdf, u, side="right"), m - 1).astype(np.int64)
    cols = np.repeat(np.arange(n, dtype=np.int64), per)
    key = np.unique(rows * n + cols)
    r = key // n
    c = key % n
    vals = np.exp(normal(NOISE_SEED, stream + 16, (key.size,))).astype(np.float32)
    indptr = np.zeros(m + 1, dtype=np.int64)
    np.add.at(indptr, r + 1, 1)
    indptr = np.cumsum(indptr)
    return indptr, c.astype(np.int32), vals
