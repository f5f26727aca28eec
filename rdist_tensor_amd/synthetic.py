"""Seeded synthetic sparse tensors for the benchmark configurations.

Coordinates are drawn per mode (uniform, or truncated Zipf(s) by inverse CDF
over p_k ~ (k+1)^-s followed by a seeded permutation of the ranks), made
distinct on a uint64 linearisation with top-up rounds until exactly `nnz`
distinct coordinates exist, then shuffled.  numpy.random.Generator(PCG64(seed)).
"""
import numpy as np


def _zipf_sampler(rng, I, s):
    k = np.arange(1, I + 1, dtype=np.float64)
    p = k ** (-s)
    cdf = np.cumsum(p)
    cdf /= cdf[-1]
    perm = rng.permutation(I).astype(np.uint32)

    def draw(n):
        u = rng.random(n)
        r = np.searchsorted(cdf, u, side="left")
        np.minimum(r, I - 1, out=r)
        return perm[r]
    return draw


def _uniform_sampler(rng, I):
    def draw(n):
        return rng.integers(0, I, size=n, dtype=np.int64).astype(np.uint32)
    return draw


def coords(dims, nnz, seed, zipf=None, max_rounds=64):
    """Distinct coordinates (nnz x N uint32).  zipf: per-mode exponent or None."""
    N = len(dims)
    rng = np.random.Generator(np.random.PCG64(seed))
    zipf = zipf or [None] * N
    draws = [(_zipf_sampler(rng, int(d), z) if z else _uniform_sampler(rng, int(d))) for d, z in zip(dims, zipf)]
    strides = np.ones(N, dtype=np.uint64)
    for k in range(N - 2, -1, -1):
        strides[k] = strides[k + 1] * np.uint64(dims[k + 1])
    total = float(np.prod(np.asarray(dims, dtype=np.float64)))
    if nnz > total:
        raise ValueError("more nonzeros than tensor entries")
    have = np.empty(0, dtype=np.uint64)
    need = nnz
    for _ in range(max_rounds):
        n = int(need * 1.15) + 1024
        lin = np.zeros(n, dtype=np.uint64)
        for k in range(N):
            lin += draws[k](n).astype(np.uint64) * strides[k]
        have = np.unique(np.concatenate([have, lin]))
        if len(have) >= nnz:
            break
        need = nnz - len(have)
    else:
        raise RuntimeError("could not draw enough distinct coordinates")
    sel = rng.choice(len(have), size=nnz, replace=False)
    lin = have[np.sort(sel)]
    rng.shuffle(lin)
    out = np.empty((nnz, N), dtype=np.uint32)
    for k in range(N):
        out[:, k] = (lin // strides[k]) % np.uint64(dims[k])
    return out


def values(n, seed, kind):
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    if kind == "uniform":
        return rng.random(n)
    if kind == "lognormal":
        return rng.lognormal(0.0, 1.0, n).astype(np.float32).astype(np.float64)
    if kind == "uniform_f32":
        return rng.random(n).astype(np.float32).astype(np.float64)
    if kind == "poisson":
        return (1.0 + rng.poisson(2.0, n)).astype(np.float32).astype(np.float64)
    raise ValueError(kind)


def factors(dims, R, seed):
    """N(0,1) factors (numpy default_rng(seed), mode by mode), unit columns."""
    rng = np.random.default_rng(seed)
    out = []
    for d in dims:
        u = rng.standard_normal((int(d), R))
        out.append(u / np.linalg.norm(u, axis=0))
    return out


# The benchmark configurations (sizes as specified; values as specified).
CONFIGS = {
    "arls_2000cube": dict(dims=[2000, 2000, 2000], nnz=200_000, zipf=None, values="uniform", R=16, J=32768,
                          sweeps=20, seed=0, sampler="cp_arls_lev"),
    "sts_1e6x1e6x1e5": dict(dims=[1_000_000, 1_000_000, 100_000], nnz=50_000_000, zipf=[1.1, 1.1, 1.1],
                            values="lognormal", R=32, J=65536, sweeps=30, seed=1, sampler="sts_cp"),
    "sts_1e7x1e6x1e4x1e2": dict(dims=[10_000_000, 1_000_000, 10_000, 100], nnz=500_000_000,
                                zipf=[1.2, 1.2, None, None], values="poisson", R=64, J=131072, sweeps=20, seed=2,
                                sampler="sts_cp"),
}


def make_batch(count=256, seed=4, size_range=(500, 50_000), nnz_per_max=20, ranks=(8, 16, 32), scale=1.0):
    """Independent 3-mode tensors of the batched CP-ARLS-LEV workload: mode
    sizes ~ U[size_range], nnz = nnz_per_max x the largest mode size at
    uniform coordinates, values U(0,1) stored FP32, R drawn from `ranks`.
    Returns [(dims, idx, vals, R, seed_t)]; scale shrinks every size."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for t in range(count):
        dims = [max(2, int(d * scale)) for d in rng.integers(size_range[0], size_range[1] + 1, size=3)]
        R = int(rng.choice(ranks))
        nnz = nnz_per_max * max(dims)
        st = seed * 100_003 + t
        idx = coords(dims, nnz, st)
        vals = values(nnz, st, "uniform_f32")
        out.append((dims, idx, vals, R, st))
    return out


def make(config, scale=1.0):
    c = dict(CONFIGS[config])
    nnz = int(c["nnz"] * scale)
    idx = coords(c["dims"], nnz, c["seed"], c["zipf"])
    vals = values(nnz, c["seed"], c["values"])
    return c, idx, vals
