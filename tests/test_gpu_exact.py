"""ExactALS on the GPU (exact_als.hpp:16-143, sort_lookup.hpp:440-495) against
the oracle's exact step.

FP64 bar: factors after each step within 1e-9 relative of the oracle's (the
exact MTTKRP sums each row in a different association across 256-nonzero
chunks, so agreement is to rounding, amplified by the R x R solve), fits
within 1e-9.  Mixed bar: final fit within 1e-3 of the FP64 path's.
"""
import json

import numpy as np
import pytest

from conftest import normalized_factors, random_tensor

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def skewed_tensor(rng, dims, nnz):
    """Coordinates with a Zipf-like first mode, so some rows span many
    256-nonzero chunks (multi-level carry reduction) and some rows are empty."""
    total = int(np.prod(dims))
    lin = np.unique(rng.integers(0, total, size=3 * nnz, dtype=np.int64))
    rng.shuffle(lin)
    idx = np.empty((len(lin), len(dims)), np.uint32)
    rem = lin.copy()
    for k in range(len(dims) - 1, -1, -1):
        idx[:, k] = rem % dims[k]
        rem //= dims[k]
    heavy = rng.random(len(lin)) < 0.4
    idx[heavy, 0] = (rng.zipf(1.6, heavy.sum()) - 1) % dims[0]
    key = np.zeros(len(lin), np.int64)
    for k in range(len(dims)):
        key = key * dims[k] + idx[:, k]
    _, first = np.unique(key, return_index=True)
    idx = idx[np.sort(first)][:nnz]
    return idx, rng.random(len(idx)) + 0.1


@pytest.mark.parametrize("case", ["uniform3_R8", "skewed3_R32", "uniform4_R16", "tiny_R64"])
def test_exact_als_matches_oracle(ctx, orc, case):
    from rdist_tensor_amd import abi
    seeds = {"uniform3_R8": 71, "skewed3_R32": 72, "uniform4_R16": 73, "tiny_R64": 74}
    rng = np.random.default_rng(seeds[case])
    if case == "uniform3_R8":
        dims, R = [400, 300, 250], 8
        idx, vals = random_tensor(rng, dims, 30000)
    elif case == "skewed3_R32":
        dims, R = [20000, 200, 150], 32
        idx, vals = skewed_tensor(rng, dims, 200000)
    elif case == "uniform4_R16":
        dims, R = [60, 50, 40, 30], 16
        idx, vals = random_tensor(rng, dims, 40000)
    else:
        dims, R = [5, 4, 3], 64   # R above the row count of every mode: singular Grams
        idx, vals = random_tensor(rng, dims, 37)
    N = len(dims)
    print(f"case {case}: dims {dims} nnz {len(idx)} R {R}", flush=True)
    U = normalized_factors(rng, dims, R)
    T = abi.Tensor(ctx, idx, vals, dims)
    F = abi.Factors(ctx, dims, R)
    for k in range(N):
        F.set(k, U[k])
    A = abi.ALS.exact(ctx, T, F)
    A.init()
    oa = orc.ALS(idx, vals, dims, U, "exact")
    assert abs(A.fit() - oa.fit()) < 1e-12
    for sweep in (1, 2):
        for j in range(N):
            A.step(j, 0, sweep)
            oa.step(j, 0, sweep)
            err = rel_err(F.get(j), oa.U[j])
            serr = rel_err(F.sigma(), oa.sigma)
            print(f"{case} sweep {sweep} mode {j}: factor rel err {err:.3e} sigma {serr:.3e}")
            assert err < 1e-9 and serr < 1e-9
    # fit = 1 - sqrt(|X|^2 - 2<X, Xh> + |Xh|^2) / |X|: when the model
    # reproduces X (tiny_R64: R above every mode size) the sum under the root
    # is rounding noise of order eps |X|^2, and the fit carries sqrt(eps)
    fo = oa.fit()
    assert abs(A.fit() - fo) < (1e-9 if fo < 0.999 else 1e-6)


def test_exact_run_statistics_and_monotone_fit(ctx, orc):
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(61)
    dims = [300, 200, 100]
    R = 8
    idx, vals = random_tensor(rng, dims, 20000)
    U = normalized_factors(rng, dims, R)
    T = abi.Tensor(ctx, idx, vals, dims)
    F = abi.Factors(ctx, dims, R)
    for k in range(3):
        F.set(k, U[k])
    A = abi.ALS.exact(ctx, T, F)
    A.init()
    A.run(5, 1234, 1)
    st = json.loads(A.stats_json())
    assert st["sampler"] == "exact"
    assert st["num_rounds"] == 5 and st["rounds"] == [0, 1, 2, 3, 4, 5]
    assert st["sum_nonzeros_iterated"] == 0
    fits = np.array(st["fits"][1:])
    assert np.all(np.diff(fits) > -1e-10), fits
    oa = orc.ALS(idx, vals, dims, U, "exact")
    ofits, _ = oa.run(5, 0, 1)
    np.testing.assert_allclose(fits, ofits, atol=1e-9)


def test_exact_mixed_fit_within_1e3_of_fp64(ctx):
    from rdist_tensor_amd import abi
    from test_gpu_mixed import planted_tensor
    rng = np.random.default_rng(62)
    dims = [60, 50, 40]
    R = 6
    idx, vals, A0 = planted_tensor(rng, dims, R, 0.05)
    keep = rng.random(len(idx)) < 0.5   # half the entries observed
    idx, vals = idx[keep], vals[keep]
    U = normalized_factors(rng, dims, R)
    fits = {}
    for prec in (abi.EX_FP64, abi.EX_MIXED):
        T = abi.Tensor(ctx, idx, vals, dims)
        F = abi.Factors(ctx, dims, R)
        for k in range(3):
            F.set(k, U[k])
        A = abi.ALS.exact(ctx, T, F, precision=prec)
        A.init()
        A.run(12, 0, 4)
        fits[prec] = json.loads(A.stats_json())["fits"][1:]
    print("fp64", fits[abi.EX_FP64], "mixed", fits[abi.EX_MIXED])
    assert abs(fits[abi.EX_FP64][-1] - fits[abi.EX_MIXED][-1]) < 1e-3
