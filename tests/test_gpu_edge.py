"""Edge cases of the randomized ALS path on the GPU against the oracle: rank 1,
a mode of size 1, one sample per solve, mode sizes below R (partial and
single-leaf trees), an empty tensor, and samples whose fibers hold no
nonzero."""
import numpy as np
import pytest

from conftest import normalized_factors, random_tensor

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    """Relative error over the finite entries; inf when the NaN entries of a
    and b differ (a zero column renormalized is NaN on both sides)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if not np.array_equal(np.isnan(a), np.isnan(b)):
        return np.inf
    ok = ~np.isnan(b)
    a, b = a[ok], b[ok]
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def run_pair(ctx, orc, dims, R, nnz, J, sampler, seed, sweeps=2, path=0):
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(seed)
    idx, vals = random_tensor(rng, dims, nnz)
    U = normalized_factors(rng, dims, R)
    kind = abi.EX_SAMPLER_STS_CP if sampler == "sts" else abi.EX_SAMPLER_CP_ARLS_LEV
    T = abi.Tensor(ctx, idx, vals, dims)
    T.mttkrp_path(path)
    F = abi.Factors(ctx, dims, R)
    for k in range(len(dims)):
        F.set(k, U[k])
    S = abi.Sampler(ctx, F, kind, seed=seed)
    A = abi.ALS(ctx, T, F, S)
    A.init()
    oa = orc.ALS(idx, vals, dims, U, sampler, seed=seed)
    worst = 0.0
    for sweep in range(1, sweeps + 1):
        for j in range(len(dims)):
            A.step(j, J, sweep)
            oa.step(j, J, sweep)
            e = rel_err(F.get(j), oa.U[j])
            worst = e if not e <= worst else worst
    return worst, A.fit(), oa.fit()


@pytest.mark.parametrize("path", [0, 3])
@pytest.mark.parametrize("sampler", ["sts", "arls"])
@pytest.mark.parametrize("case", ["rank1", "mode_of_size_1", "one_sample", "modes_below_R"])
def test_edge_cases_match_oracle(ctx, orc, sampler, case, path):
    """path 0: the automatic sampled-MTTKRP choice; 3: bucketed hits."""
    if case == "rank1":
        dims, R, nnz, J = [200, 150, 100], 1, 5000, 512
    elif case == "mode_of_size_1":
        dims, R, nnz, J = [80, 1, 60], 4, 3000, 512
    elif case == "one_sample":
        dims, R, nnz, J = [30, 20, 10], 4, 5000, 1   # dense: the one fiber holds nonzeros
    else:
        dims, R, nnz, J = [5, 7, 3], 8, 90, 256   # every tree one (partial) leaf
    print(f"{sampler}/{case}: dims {dims} R {R} J {J}", flush=True)
    seed = {"rank1": 141, "mode_of_size_1": 142, "one_sample": 143, "modes_below_R": 144}[case]
    worst, fg, fo = run_pair(ctx, orc, dims, R, nnz, J, sampler, seed=seed, path=path)
    print(f"{sampler}/{case}: factor rel err {worst:.2e}, fit {fg:.6f} vs {fo:.6f}")
    assert worst < 1e-9
    if np.isfinite(fo):
        assert abs(fg - fo) < 1e-9


def test_empty_tensor(ctx):
    """No nonzeros: the lookups, the sampled MTTKRP (all zeros, no hits) and
    a sampled step run without error."""
    from rdist_tensor_amd import abi
    dims = [30, 20, 10]
    R = 4
    T = abi.Tensor(ctx, np.zeros((0, 3), np.uint32), np.zeros(0), dims)
    T.build_lookups()
    rng = np.random.default_rng(5)
    s = np.stack([rng.integers(0, d, 50) for d in dims], 1).astype(np.uint32)
    _, first = np.unique(s[:, 1:], axis=0, return_index=True)
    s = s[np.sort(first)]
    Y, hits = T.sampled_mttkrp(0, s, rng.standard_normal((len(s), R)))
    assert hits == 0 and not np.any(Y)
    F = abi.Factors(ctx, dims, R)
    for k, u in enumerate(normalized_factors(rng, dims, R)):
        F.set(k, u)
    S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP, seed=1)
    A = abi.ALS(ctx, T, F, S)
    A.init()
    A.step(0, 256, 1)   # Y = 0: the factor's columns have zero norm (as in the reference)


@pytest.mark.parametrize("path", [0, 1, 2, 3])
def test_samples_without_nonzeros(ctx, orc, path):
    """A very sparse tensor where most sampled fibers are empty: rows of Y
    without hits stay zero, the rest match the oracle (every sampled-MTTKRP
    path)."""
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(6)
    dims = [2000, 3000, 4000]
    idx, vals = random_tensor(rng, dims, 2000)
    T = abi.Tensor(ctx, idx, vals, dims)
    T.mttkrp_path(path)
    T.build_lookups()
    for mode in range(3):
        lk = orc.Lookup(idx, vals, mode, dims)
        s = np.concatenate([idx[rng.integers(0, len(idx), 100)],
                            np.stack([rng.integers(0, d, 5000) for d in dims], 1).astype(np.uint32)])
        sd, _ = orc.dedup(s, np.ones(len(s)), mode, dims)
        design = rng.standard_normal((len(sd), 8))
        Yo, Ho = lk.sampled_mttkrp(sd, design)
        Yg, Hg = T.sampled_mttkrp(mode, sd, design)
        assert Hg == Ho and Ho > 0
        assert rel_err(Yg, Yo) < 1e-12


@pytest.mark.parametrize("case", ["rank1", "mode_of_size_1", "modes_below_R"])
def test_mixed_edge_cases_run(ctx, case):
    """The mixed-precision path (FP32 trees, draws and sampled MTTKRP) on the
    same shapes: finite factors and a fit near the FP64 path's (the two draw
    different samples, so the fits agree to sampling noise, not bitwise)."""
    from rdist_tensor_amd import abi
    dims, R, nnz = {"rank1": ([200, 150, 100], 1, 5000), "mode_of_size_1": ([80, 1, 60], 4, 3000),
                    "modes_below_R": ([5, 7, 3], 8, 90)}[case]
    rng = np.random.default_rng(150)
    idx, vals = random_tensor(rng, dims, nnz)
    U = normalized_factors(rng, dims, R)
    T = abi.Tensor(ctx, idx, vals, dims)
    fits = {}
    for prec in (abi.EX_FP64, abi.EX_MIXED):
        F = abi.Factors(ctx, dims, R)
        for k in range(len(dims)):
            F.set(k, U[k])
        S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP, precision=prec, seed=9)
        A = abi.ALS(ctx, T, F, S, precision=prec)
        A.init()
        for sweep in range(1, 4):
            for j in range(len(dims)):
                A.step(j, 4096, sweep)
        assert all(np.all(np.isfinite(F.get(k))) for k in range(len(dims)))
        fits[prec] = A.fit()
    print(f"mixed/{case}: fit {fits[abi.EX_MIXED]:.6f} vs FP64 {fits[abi.EX_FP64]:.6f}")
    assert abs(fits[abi.EX_MIXED] - fits[abi.EX_FP64]) < 0.05
