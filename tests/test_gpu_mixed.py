"""Mixed-precision path (FP32 data, BF16 MFMA Grams with FP32 accumulation,
FP64 R x R solves) against FP64 closed forms and against the FP64 path.

Tolerances: BF16 keeps 8 significand bits, so a Gram entry built from BF16
operands carries a relative error of order 2^-9 / sqrt(terms) after FP32
accumulation; quantities derived from Grams (inverses, leverage scores) are
checked at 2e-2 relative (Frobenius / max-relative), FP32 data paths at 1e-5.
The end-to-end bar is the final fit: within 1e-3 of the FP64 path's.
"""
import numpy as np
import pytest

from conftest import normalized_factors, random_tensor

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def set_factors(ctx, dims, R, U):
    from rdist_tensor_amd import abi
    F = abi.Factors(ctx, dims, R)
    for k in range(len(dims)):
        F.set(k, U[k])
    return F


@pytest.mark.parametrize("R", [8, 16, 32, 64])
def test_mixed_arls_scores(ctx, R):
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(300 + R)
    dims = [5000, 3000, 1001]
    U = normalized_factors(rng, dims, R)
    F = set_factors(ctx, dims, R, U)
    S = abi.Sampler(ctx, F, abi.EX_SAMPLER_CP_ARLS_LEV, precision=abi.EX_MIXED)
    for k in range(3):
        lev = np.einsum("ir,rs,is->i", U[k], np.linalg.pinv(U[k].T @ U[k]), U[k])
        got = S.arls_scores(k)
        err = np.max(np.abs(got / lev - 1.0))
        print(f"R={R} mode {k}: leverage max rel err {err:.2e}, sum {got.sum():.6f} (rank {R})")
        assert err < 2e-2
        assert abs(got.sum() - R) < 1e-2 * R


@pytest.mark.parametrize("R", [8, 16, 32, 64])
def test_mixed_compute_m_factorisations(ctx, R):
    """Slot 0 = (Hadamard of the non-target Grams)^+, slot N rows w_v with
    sum_v w_v w_v^T = slot 0, slot k rows with sum_v w_v w_v^T equal to
    (Hadamard of the Grams of modes >= k, != j) o slot 0, slot N+1 the
    Hadamard product itself (sts_cp.hpp:80-188)."""
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(400 + R)
    dims = [2000, 1500, 1200, 900]
    N = len(dims)
    U = normalized_factors(rng, dims, R)
    F = set_factors(ctx, dims, R, U)
    S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP, precision=abi.EX_MIXED)
    G = [u.T @ u for u in U]
    for j in range(N):
        M = S.sts_m(j)
        H = np.ones((R, R))
        for k in range(N):
            if k != j:
                H = H * G[k]
        Hi = np.linalg.pinv(H)
        assert rel_err(M[N + 1], H) < 2e-2, j
        assert rel_err(M[0], Hi) < 2e-2, j
        assert rel_err(M[N].T @ M[N], M[0]) < 1e-10, j
        for k in range(1, N):
            if k == j:
                continue
            T = np.ones((R, R))
            for m in range(k, N):
                if m != j:
                    T = T * G[m]
            target = T * Hi
            W = M[k]
            assert rel_err(W.T @ W, target) < 2e-2, (j, k)


def planted_tensor(rng, dims, R, noise):
    """A dense tensor (every coordinate a nonzero) of CP rank R plus noise;
    returns the coordinates, the values and the planted factors."""
    A = [rng.standard_normal((d, R)) for d in dims]
    grid = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1).reshape(-1, len(dims))
    vals = np.ones((len(grid), R))
    for k in range(len(dims)):
        vals *= A[k][grid[:, k]]
    vals = vals.sum(1)
    vals += noise * np.linalg.norm(vals) / np.sqrt(len(vals)) * rng.standard_normal(len(vals))
    return grid.astype(np.uint32), vals, A


@pytest.mark.parametrize("kind", ["sts", "arls"])
def test_mixed_fit_within_1e3_of_fp64(ctx, kind):
    """Both precisions start from the planted factors plus a 30% perturbation
    (inside the basin of the planted solution: from a random start, randomized
    ALS on this tensor can sit on plateaus whose level depends on the draws,
    and the two precisions draw differently) and must reach the same final
    fit (the noise floor, 0.95) within 1e-3."""
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(501)
    dims = [120, 100, 80]
    R = 8
    idx, vals, A0 = planted_tensor(rng, dims, R, noise=0.05)
    U0 = [a + 0.3 * np.linalg.norm(a, axis=0) / np.sqrt(len(a)) * rng.standard_normal(a.shape) for a in A0]
    U0 = [u / np.linalg.norm(u, axis=0) for u in U0]
    sk = abi.EX_SAMPLER_STS_CP if kind == "sts" else abi.EX_SAMPLER_CP_ARLS_LEV
    fits = {}
    for prec in (abi.EX_FP64, abi.EX_MIXED):
        T = abi.Tensor(ctx, idx, vals, dims)
        F = set_factors(ctx, dims, R, U0)
        S = abi.Sampler(ctx, F, sk, precision=prec, seed=5)
        A = abi.ALS(ctx, T, F, S, precision=prec)
        A.init()
        hist = []
        for sweep in range(1, 13):
            for j in range(3):
                A.step(j, 8192, sweep)
            hist.append(A.fit())
        fits[prec] = hist
        print(f"{kind} precision {prec}: fits " + " ".join(f"{f:.5f}" for f in hist))
    f64, fmx = fits[abi.EX_FP64][-1], fits[abi.EX_MIXED][-1]
    assert f64 > 0.9
    assert abs(fmx - f64) < 1e-3


def test_mixed_factor_copies_stay_consistent(ctx):
    """Factors written by a mixed step are visible through the FP64 accessors,
    and factors set from the host are seen by the next mixed step."""
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(601)
    dims = [300, 200, 100]
    R = 16
    idx, vals = random_tensor(rng, dims, 20000)
    U = normalized_factors(rng, dims, R)
    T = abi.Tensor(ctx, idx, vals, dims)
    F = set_factors(ctx, dims, R, U)
    S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP, precision=abi.EX_MIXED, seed=3)
    A = abi.ALS(ctx, T, F, S, precision=abi.EX_MIXED)
    A.init()
    A.step(0, 4096, 1)
    U0 = F.get(0)
    # renormalised columns, FP32-representable values
    assert np.allclose(np.linalg.norm(U0, axis=0), 1.0, atol=1e-5)
    assert np.array_equal(U0.astype(np.float32).astype(np.float64), U0)
    # a host write of mode 1 reaches the mixed path: the sampler's FP32 tree
    # for mode 1 is rebuilt from the new factor
    V = normalized_factors(rng, [dims[1]], R)[0]
    F.set(1, V)
    S.update(1)
    A.step(0, 4096, 2)
    assert np.allclose(F.get(1), V)
    assert np.isfinite(A.fit())


@pytest.mark.parametrize("R", [16, 32])
def test_mixed_prepared_draw_setup_equals_fresh(ctx, R):
    """A mixed STS-CP step prepares the next mode's computeM, eigen trees and
    packed weights on the second stream beside its row-tree rebuild.  A draw
    of that mode from the prepared set-up equals, bit for bit, a draw after
    the set-up is recomputed (ex_sts_compute_m clears the prepared mark), and
    a draw of another mode recomputes its own set-up."""
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(700 + R)
    dims = [400, 300, 200, 150]
    idx, vals = random_tensor(rng, dims, 30000)
    U = normalized_factors(rng, dims, R)
    T = abi.Tensor(ctx, idx, vals, dims)
    F = set_factors(ctx, dims, R, U)
    S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP, precision=abi.EX_MIXED, seed=9)
    A = abi.ALS(ctx, T, F, S, precision=abi.EX_MIXED)
    A.init()
    A.step(0, 2048, 1)                       # prepares mode 1's draw set-up
    s_pre, w_pre = S.draw(4096, 1, sweep=5)
    S.sts_m(1)                               # recomputes M, clears the mark
    s_new, w_new = S.draw(4096, 1, sweep=5)  # recomputes the whole set-up
    assert np.array_equal(s_pre, s_new)
    assert np.array_equal(w_pre, w_new)
    assert np.isfinite(w_pre).all() and (w_pre > 0).all()
    A.step(1, 2048, 1)                       # prepares mode 2's
    s_a, w_a = S.draw(4096, 3, sweep=6)      # another mode: its own set-up
    s_b, w_b = S.draw(4096, 3, sweep=6)
    assert np.array_equal(s_a, s_b)
    assert np.array_equal(w_a, w_b)


def test_four_mode_rank64_shape(ctx):
    """The 4-mode R = 64 shape of the at-scale STS-CP workload (Zipf(1.2) on
    modes 0-1, uniform on 2-3, 1 + Poisson(2) values, J = 131,072) with every
    mode size and the nonzero count divided down: both precisions run, the
    counters are consistent and the final fits agree within 1e-3."""
    import json
    from rdist_tensor_amd import abi, synthetic
    dims = [100_000, 10_000, 1_000, 100]
    R, J = 64, 131072
    idx = synthetic.coords(dims, 1_000_000, 2, [1.2, 1.2, None, None])
    vals = synthetic.values(len(idx), 2, "poisson")
    U = synthetic.factors(dims, R, 2)
    fits = {}
    for prec in (abi.EX_FP64, abi.EX_MIXED):
        T = abi.Tensor(ctx, idx, vals, dims)
        F = abi.Factors(ctx, dims, R)
        for k in range(4):
            F.set(k, U[k])
        S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP, precision=prec, seed=2)
        A = abi.ALS(ctx, T, F, S, precision=prec)
        A.init()
        A.run(4, J, 2)
        st = json.loads(A.stats_json())
        assert 0 < st["sum_nonzeros_iterated"] <= 4 * 4 * len(idx)
        fits[prec] = st["fits"][1:]
    print("fp64", fits[abi.EX_FP64], "mixed", fits[abi.EX_MIXED])
    assert abs(fits[abi.EX_FP64][-1] - fits[abi.EX_MIXED][-1]) < 1e-3


def test_batched_run_equals_single_runs():
    """ex_als_run_batch over independent tensors (each optimizer on its own
    context, the streams side by side) gives exactly the factors that each
    optimizer reaches alone with ex_als_run."""
    from rdist_tensor_amd import abi, synthetic
    batch = synthetic.make_batch(count=6, seed=4, size_range=(300, 3000))
    results = {}
    for mode in ("single", "batch"):
        objs = []
        for dims, idx, vals, R, st in batch:
            c = abi.Context(0)
            T = abi.Tensor(c, idx, vals, dims)
            F = abi.Factors(c, dims, R)
            for k, u in enumerate(synthetic.factors(dims, R, st)):
                F.set(k, u)
            S = abi.Sampler(c, F, abi.EX_SAMPLER_CP_ARLS_LEV, precision=abi.EX_MIXED, seed=st)
            A = abi.ALS(c, T, F, S, precision=abi.EX_MIXED)
            A.init()
            objs.append((c, T, F, S, A))
        if mode == "single":
            for c, T, F, S, A in objs:
                A.run(3, 2048, 100)
        else:
            sec = abi.ALS.run_batch([o[4] for o in objs], 3, 2048)
            print(f"batch of {len(objs)}: {sec * 1e3:.1f} ms")
        results[mode] = [[o[2].get(k) for k in range(3)] for o in objs]
        for c, *_ in objs:
            c.close()
    for a, b in zip(results["single"], results["batch"]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", ["sts", "arls"])
def test_graph_replay_equals_eager(ctx, precision, kind):
    """Sweeps replayed from a captured graph (ex_als_use_graph) give the
    eager run's factors and fits bitwise, for both samplers and precisions;
    the batched runner replays too."""
    import json
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(77)
    dims = [400, 300, 200]
    R, J, rounds = 8, 4096, 5
    idx, vals = random_tensor(rng, dims, 20000)
    U = normalized_factors(rng, dims, R)
    prec = abi.EX_MIXED if precision == "mixed" else abi.EX_FP64
    sk = abi.EX_SAMPLER_STS_CP if kind == "sts" else abi.EX_SAMPLER_CP_ARLS_LEV
    out = {}
    for graph in (False, True):
        T = abi.Tensor(ctx, idx, vals, dims)
        F = set_factors(ctx, dims, R, U)
        S = abi.Sampler(ctx, F, sk, precision=prec, seed=5)
        A = abi.ALS(ctx, T, F, S, precision=prec)
        A.init()
        A.use_graph(graph)
        A.run(rounds, J, 1)
        st = json.loads(A.stats_json())
        out[graph] = ([F.get(k) for k in range(3)], st["fits"], st["sum_nonzeros_iterated"])
    for k in range(3):
        assert np.array_equal(out[True][0][k], out[False][0][k]), k
    assert out[True][1] == out[False][1]
    assert out[True][2] == out[False][2]
    print(f"{kind}/{precision}: {rounds} sweeps, fits {out[True][1][-1]:.6f}, graph replay bitwise equal")


def test_graph_with_span_timing_runs_eagerly():
    """With the context's span timing on, an optimizer with graphs enabled
    enqueues every sweep step by step: each mode step's draw span is
    recorded, and the factors equal the graph-replayed run's."""
    import json
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(78)
    dims = [300, 250, 200]
    R, J, rounds = 8, 2048, 4
    idx, vals = random_tensor(rng, dims, 15000)
    U = normalized_factors(rng, dims, R)
    out = {}
    for timing in (False, True):
        c = abi.Context(0)
        try:
            T = abi.Tensor(c, idx, vals, dims)
            F = set_factors(c, dims, R, U)
            S = abi.Sampler(c, F, abi.EX_SAMPLER_STS_CP, precision=abi.EX_MIXED, seed=6)
            A = abi.ALS(c, T, F, S, precision=abi.EX_MIXED)
            A.init()
            A.use_graph(True)
            c.timing(timing)
            A.run(rounds, J, 10 ** 9)
            c.sync()
            if timing:
                spans = json.loads(c.timing_json())["spans"]
                assert spans["sts_draw"]["count"] == rounds * len(dims)
                c.timing(False)
            out[timing] = [F.get(k) for k in range(3)]
        finally:
            c.close()
    for k in range(3):
        assert np.array_equal(out[True][k], out[False][k]), k


def test_batched_graph_replay_equals_eager():
    """ex_als_run_batch with every optimizer on graph replay reaches the
    eager batch's factors bitwise."""
    from rdist_tensor_amd import abi, synthetic
    batch = synthetic.make_batch(count=4, seed=4, size_range=(300, 3000))
    results = {}
    for graph in (False, True):
        objs, opts = [], []
        for dims, idx, vals, R, st in batch:
            c = abi.Context(0)
            T = abi.Tensor(c, idx, vals, dims)
            F = abi.Factors(c, dims, R)
            for k, u in enumerate(synthetic.factors(dims, R, st)):
                F.set(k, u)
            S = abi.Sampler(c, F, abi.EX_SAMPLER_CP_ARLS_LEV, precision=abi.EX_MIXED, seed=st)
            A = abi.ALS(c, T, F, S, precision=abi.EX_MIXED)
            A.init()
            A.use_graph(graph)
            objs.append((c, T, F, S, A))
            opts.append(A)
        abi.ALS.run_batch(opts, 4, 2048)
        results[graph] = [[o[2].get(k) for k in range(3)] for o in objs]
        for c, T, F, S, A in objs:
            c.close()
    for a, b in zip(results[True], results[False]):
        for k in range(3):
            assert np.all(np.isfinite(b[k]))
            assert np.array_equal(a[k], b[k])


def test_mixed_arls_singular_design_gram():
    """A 1524 x 1940 x 1064 tensor with 38,800 uniform nonzeros at R = 16 and
    J = 2048: few samples hit a nonzero, the factors become sparse and the
    design Gram of the fourth sweep is singular.  Gauss-Jordan drops a pivot
    there (a zero column of its inverse, and a 0/0 column norm); for R <= 32
    the mixed path then takes the pivoted-Cholesky pseudo-inverse in the same
    workgroup (the FP64 path's Jacobi pinv for the same matrix), the factors
    stay finite, and the fits of the six sweeps follow the FP64 path's run
    of the same case within 1e-2 (the two paths draw different samples:
    their leverage scores differ at the BF16 Gram's rounding)."""
    from rdist_tensor_amd import abi, synthetic
    dims, idx, vals, R, st = synthetic.make_batch(count=4, seed=4, size_range=(300, 3000))[2]
    assert dims == [1524, 1940, 1064] and R == 16
    fits = {}
    for prec in (abi.EX_FP64, abi.EX_MIXED):
        c = abi.Context(0)
        try:
            T = abi.Tensor(c, idx, vals, dims)
            F = abi.Factors(c, dims, R)
            for k, u in enumerate(synthetic.factors(dims, R, st)):
                F.set(k, u)
            S = abi.Sampler(c, F, abi.EX_SAMPLER_CP_ARLS_LEV, precision=prec, seed=st)
            A = abi.ALS(c, T, F, S, precision=prec)
            A.init()
            A.run(6, 2048, 1)
            import json
            fits[prec] = json.loads(A.stats_json())["fits"][1:]
            if prec == abi.EX_MIXED:
                for k in range(3):
                    assert np.all(np.isfinite(F.get(k)))
                assert np.isfinite(A.fit())
        finally:
            c.close()
    print("fp64", fits[abi.EX_FP64], "mixed", fits[abi.EX_MIXED])
    assert np.max(np.abs(np.subtract(fits[abi.EX_FP64], fits[abi.EX_MIXED]))) < 1e-2


@pytest.mark.parametrize("kind", ["arls", "sts"])
def test_mixed_duplicate_columns_stay_duplicate(ctx, kind):
    """Every factor starts with columns 0 and 1 equal: the design Grams are
    singular in the direction e0 - e1 and both paths take the minimum-norm
    solution, which keeps the two columns equal.  A design Gram built from
    BF16-rounded rows would carry a noise eigenvalue of about 2^-9 relative in
    that direction and invert it; the mixed path builds it from the FP32
    rows (noise below the 1e-11 pivot tolerance).  After six sweeps the two
    columns still agree to 1e-3 on both paths and the fits agree within
    1e-2."""
    import json
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(90)
    dims = [400, 300, 250]
    R = 8
    idx, vals = random_tensor(rng, dims, 30000)
    U = normalized_factors(rng, dims, R)
    for u in U:
        u[:, 1] = u[:, 0]
    sk = abi.EX_SAMPLER_STS_CP if kind == "sts" else abi.EX_SAMPLER_CP_ARLS_LEV
    fits = {}
    for prec in (abi.EX_FP64, abi.EX_MIXED):
        T = abi.Tensor(ctx, idx, vals, dims)
        F = set_factors(ctx, dims, R, U)
        S = abi.Sampler(ctx, F, sk, precision=prec, seed=12)
        A = abi.ALS(ctx, T, F, S, precision=prec)
        A.init()
        A.run(6, 4096, 1)
        fits[prec] = json.loads(A.stats_json())["fits"][1:]
        for k in range(3):
            V = F.get(k)
            assert np.all(np.isfinite(V))
            d = np.linalg.norm(V[:, 0] - V[:, 1]) / np.linalg.norm(V[:, 0])
            print(f"{kind} precision {prec} mode {k}: |u0 - u1| / |u0| = {d:.2e}")
            assert d < 1e-3
    print("fp64", fits[abi.EX_FP64], "mixed", fits[abi.EX_MIXED])
    assert np.max(np.abs(np.subtract(fits[abi.EX_FP64], fits[abi.EX_MIXED]))) < 1e-2


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_recapture_after_host_write(ctx, precision):
    """A replayed optimizer whose buffers moved (the fit enlarges one) captures
    its sweep again; a factor written from the host in between must be read
    by the new capture and by every replay after it, as the eager run reads
    it: bitwise equal factors and fits."""
    import json
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(79)
    dims = [400, 300, 200]
    R, J = 8, 4096
    idx, vals = random_tensor(rng, dims, 20000)
    U = normalized_factors(rng, dims, R)
    V = normalized_factors(rng, [dims[1]], R)[0]
    prec = abi.EX_MIXED if precision == "mixed" else abi.EX_FP64
    out = {}
    for graph in (False, True):
        T = abi.Tensor(ctx, idx, vals, dims)
        F = set_factors(ctx, dims, R, U)
        S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP, precision=prec, seed=8)
        A = abi.ALS(ctx, T, F, S, precision=prec)
        A.init()
        A.use_graph(graph)
        A.run(3, J, 10)
        A.fit()
        F.set(1, V)
        A.run(3, J, 1)
        out[graph] = ([F.get(k) for k in range(3)], json.loads(A.stats_json())["fits"])
    for k in range(3):
        assert np.array_equal(out[True][0][k], out[False][0][k]), k
    assert out[True][1] == out[False][1]
