"""CPU tests pinning the oracle (the checker of the HIP path) against golden
fixtures, closed forms, brute force and statistics.  No GPU needed.

Each check is chosen so that a plausible slip in the oracle (a dropped term,
a transposed operand, a wrong inequality or index) fails it.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest
from scipy.stats import chisquare

from conftest import normalized_factors, random_tensor

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- draws
def test_philox_known_answers(orc):
    with open(os.path.join(GOLDEN, "philox_kat.json")) as f:
        for v in json.load(f):
            assert list(orc.philox4x32(v["ctr"], v["key"])) == v["out"]


def test_uniform_is_generate_canonical_of_philox_words(orc):
    seed, stream = 0x0123456789ABCDEF, orc.stream_id(3, 1, 2, 1)
    u = orc.uniform(seed, stream, 1000, 64)
    for i in range(64):
        idx = 1000 + i
        w = orc.philox4x32([idx & 0xFFFFFFFF, idx >> 32, stream & 0xFFFFFFFF, stream >> 32],
                           [seed & 0xFFFFFFFF, seed >> 32])
        s = float(w[0]) + float(w[1]) * 4294967296.0
        expect = s / 18446744073709551616.0
        if expect >= 1.0:
            expect = np.nextafter(1.0, 0.0)
        assert u[i] == expect
    assert np.all((u >= 0) & (u < 1))


def test_uniform_moments(orc):
    u = orc.uniform(7, orc.stream_id(1, 0, 1, 2), 0, 200000)
    assert abs(u.mean() - 0.5) < 0.005
    assert abs(u.var() - 1 / 12) < 0.002
    assert len(np.unique(u)) == len(u)


# ---------------------------------------------------------------- dense
def test_gram_closed_form(orc):
    rng = np.random.default_rng(1)
    for m, R in [(1, 3), (50, 7), (5000, 16), (4099, 33)]:
        A = rng.standard_normal((m, R))
        G = orc.gram(A)
        np.testing.assert_allclose(G, A.T @ A, rtol=1e-13, atol=1e-12 * m)
        assert np.array_equal(G, G.T)


def test_cos_gram_kat(orc):
    z = np.load(os.path.join(GOLDEN, "cos_gram.npz"))
    np.testing.assert_allclose(orc.gram(z["U"]), z["G"], rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("R", [2, 5, 16, 25, 64])
def test_eigh_closed_form(orc, R):
    rng = np.random.default_rng(R)
    B = rng.standard_normal((2 * R, R))
    A = B.T @ B
    lam, V = orc.eigh(A)
    assert np.all(np.diff(lam) >= 0)
    np.testing.assert_allclose(lam, np.linalg.eigvalsh(A), rtol=1e-12, atol=1e-12 * lam.max())
    np.testing.assert_allclose(V.T @ V, np.eye(R), atol=1e-13)
    np.testing.assert_allclose(V @ np.diag(lam) @ V.T, A, atol=1e-12 * lam.max())


def test_eigh_reads_upper_triangle_only(orc):
    rng = np.random.default_rng(3)
    B = rng.standard_normal((10, 6))
    A = B.T @ B
    garbage = A.copy()
    garbage[np.tril_indices(6, -1)] = 1e9
    l1, V1 = orc.eigh(A)
    l2, V2 = orc.eigh(garbage)
    assert np.array_equal(l1, l2) and np.array_equal(V1, V2)


def test_pinv_tolerance_rule(orc):
    rng = np.random.default_rng(4)
    R = 10
    Q, _ = np.linalg.qr(rng.standard_normal((R, R)))
    lam = np.array([1e-13, 5e-12, 1e-3, 1, 2, 3, 4, 5, 6, 7.0])
    M = Q @ np.diag(lam) @ Q.T
    P = orc.pinv(M, 1e-11)
    keep = lam > 1e-11
    expect = Q[:, keep] @ np.diag(1 / lam[keep]) @ Q[:, keep].T
    # the two eigenvalues below 1e-11 are dropped (keeping either would add ~1e11)
    np.testing.assert_allclose(P, expect, rtol=1e-8, atol=1e-8)
    # exactly the eigenvalue below the tolerance is dropped
    P2 = orc.pinv(M, 1e-2)
    keep2 = lam > 1e-2
    np.testing.assert_allclose(P2, Q[:, keep2] @ np.diag(1 / lam[keep2]) @ Q[:, keep2].T, atol=1e-9)


def test_dagat_dsymm_renormalize(orc):
    rng = np.random.default_rng(5)
    A = rng.standard_normal((300, 9))
    S = rng.standard_normal((9, 9))
    G = S + S.T
    Gu = np.triu(G) + np.triu(G, 1).T   # symmetric defined by the upper triangle
    Gbad = Gu.copy()
    Gbad[np.tril_indices(9, -1)] = 123.0
    np.testing.assert_allclose(orc.dagat(A, Gbad), np.einsum("ij,jk,ik->i", A, Gu, A), rtol=1e-12)
    np.testing.assert_allclose(orc.dsymm(Gbad, A), A @ Gu, rtol=1e-12, atol=1e-12)
    U, sig = orc.renormalize(A)
    np.testing.assert_allclose(sig, np.linalg.norm(A, axis=0), rtol=1e-14)
    np.testing.assert_allclose(U, A / np.linalg.norm(A, axis=0), rtol=1e-14)


# ---------------------------------------------------------------- trees
def _inverse_cdf(p, u):
    cdf = np.cumsum(p) / p.sum()
    return np.searchsorted(cdf, u, side="left")


@pytest.mark.parametrize("n,F,R", [(100, 8, 8), (517, 16, 16), (37, 8, 8), (8, 1, 8), (25, 1, 25), (1000, 32, 32)])
def test_tree_structure_and_descent_is_inverse_cdf(orc, n, F, R):
    rng = np.random.default_rng(n + F)
    U = rng.standard_normal((n, R))
    t = orc.Tree(n, F, R)
    t.build(U)
    G = t.G
    nc = t.t.node_count
    # root holds U^T U (upper), every internal node = left + right child
    np.testing.assert_allclose(np.triu(G[0]), np.triu(U.T @ U), rtol=1e-12, atol=1e-10)
    for c in range(nc):
        if 2 * c + 2 < nc:
            np.testing.assert_allclose(G[c], G[2 * c + 1] + G[2 * c + 2], rtol=1e-13, atol=1e-12)
    # leaves in heap order cover the rows in order
    first = nc - t.t.leaf_count
    idxs = sorted(t.leaf_idx(c) for c in range(first, nc))
    assert idxs == list(range(t.t.leaf_count))
    # descent == inverse CDF over rows with p_i = (u_i . x)^2
    J = 4000
    x = rng.standard_normal(R)
    X = np.tile(x, (J, 1))
    d = rng.random(J)
    smp, h, margin = t.sample(U, X, d)
    expect = _inverse_cdf((U @ x) ** 2, d)
    bad = smp != expect
    assert not np.any(bad & (margin >= 1e-12))
    np.testing.assert_allclose(h, U[smp], rtol=0, atol=0)


def test_tree_boundary_draw_exact_rational(orc):
    """Draws placed exactly on bin boundaries (computed in rationals) go to
    the lower bin at internal nodes ('u <= cutoff') -- checked on a tiny tree
    whose masses are exact binary fractions."""
    R, n, F = 2, 4, 1
    U = np.array([[1.0, 0.0], [1.0, 0.0], [1.0, 0.0], [1.0, 0.0]])
    t = orc.Tree(n, F, R)
    t.build(U)
    x = np.array([[1.0, 0.0]])
    for k in range(1, 4):
        cut = float(Fraction(k, 4))
        smp, _, margin = t.sample(U, x, np.array([cut]))
        assert smp[0] == k - 1
        assert margin[0] == 0.0


def test_multiply_matrices_against_provided(orc):
    rng = np.random.default_rng(9)
    R = 6
    W = rng.standard_normal((R, R))
    Gk = rng.standard_normal((R, R))
    t = orc.Tree(R, 1, R)
    t.build(W)
    before = t.G
    t.multiply(Gk)
    t.multiply(Gk)   # idempotent: multiplies the stored unmultiplied copy
    np.testing.assert_array_equal(t.G, before * Gk)


# ---------------------------------------------------------------- CP-ARLS-LEV
def _discrete_distribution_restated(w, u):
    """libstdc++ discrete_distribution (random.tcc:2655-2712) in Python floats."""
    if len(w) < 2:
        return 0
    s = 0.0
    for x in w:
        s += x
    p = [x / s for x in w]
    cp, run = [], 0.0
    for x in p:
        run = run + x
        cp.append(run)
    cp[-1] = 1.0
    lo, hi = 0, len(cp)
    while lo < hi:
        mid = (lo + hi) // 2
        if cp[mid] < u:
            lo = mid + 1
        else:
            hi = mid
    return lo


def test_arls_draw_bitexact_vs_libstdcxx_restatement(orc):
    rng = np.random.default_rng(11)
    dims = [13, 9, 17]
    U = normalized_factors(rng, dims, 4)
    sc = [orc.arls_scores(u) for u in U]
    J = 300
    draws = rng.random((3, J))
    for j in range(3):
        s, w, _ = orc.arls_draw(sc, j, J, draws)
        for i in range(J):
            logw = -math.log(J)
            for k in range(3):
                if k == j:
                    assert s[i, k] == 0
                    continue
                idx = _discrete_distribution_restated(list(sc[k]), draws[k, i])
                assert s[i, k] == idx
                total = 0.0
                for x in sc[k]:
                    total += x
                logw += math.log(0.0 + total) - math.log(sc[k][idx])
            assert w[i] == math.exp(logw)


def test_arls_scores_are_leverage(orc):
    rng = np.random.default_rng(12)
    A = rng.standard_normal((400, 12))
    Q, _ = np.linalg.qr(A)
    np.testing.assert_allclose(orc.arls_scores(A), (Q ** 2).sum(1), rtol=1e-10)
    assert abs(orc.arls_scores(A).sum() - 12) < 1e-9


def test_arls_chisquare_and_weights(orc):
    rng = np.random.default_rng(13)
    dims = [11, 7, 6]
    U = normalized_factors(rng, dims, 3)
    sc = [orc.arls_scores(u) for u in U]
    J = 200000
    s, w, _ = orc.arls_draw(sc, 1, J, rng.random((3, J)))
    p0, p2 = sc[0] / sc[0].sum(), sc[2] / sc[2].sum()
    p = np.outer(p0, p2).ravel()
    key = s[:, 0].astype(np.int64) * dims[2] + s[:, 2]
    cnt = np.bincount(key, minlength=len(p))
    assert chisquare(cnt, p * J).pvalue > 1e-4
    np.testing.assert_allclose(w * J * p[key], 1.0, rtol=1e-12)


# ---------------------------------------------------------------- STS-CP
def test_sts_compute_m_closed_forms(orc):
    rng = np.random.default_rng(21)
    dims = [40, 30, 20, 10]
    R = 5
    U = normalized_factors(rng, dims, R)
    s = orc.STS(U)
    G = [u.T @ u for u in U]
    for j in range(4):
        s.compute_m(j)
        others = [k for k in range(4) if k != j]
        full = np.ones((R, R))
        for k in others:
            full *= G[k]
        pinv = np.linalg.pinv(full, rcond=1e-14)
        np.testing.assert_allclose(np.triu(s.slot(0)), np.triu(pinv), rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(s.slot(len(dims) + 1), full, rtol=1e-12)
        # rows of slot `off` are scaled eigenvectors: W^T W = later-Grams (.) pinv
        for k in others:
            off = k + 2 if k + 1 == j else k + 1
            later = np.ones((R, R))
            for kk in others:
                if kk > k:
                    later *= G[kk]
            W = s.slot(off)
            np.testing.assert_allclose(W.T @ W, later * pinv, rtol=1e-8, atol=1e-9)


def test_sts_matches_exact_leverage_fixture(orc):
    z = np.load(os.path.join(GOLDEN, "tiny_leverage.npz"))
    U = [z["U0"], z["U1"], z["U2"]]
    j = int(z["j"])
    lev = z["lev"]
    s = orc.STS(U)
    J = 300000
    rng = np.random.default_rng(22)
    smp, w, _ = s.draw(j, J, rng.random((3, J)), rng.random((3, J)))
    p = lev / lev.sum()
    key = smp[:, 0].astype(np.int64) * U[2].shape[0] + smp[:, 2]
    cnt = np.bincount(key, minlength=len(p))
    assert chisquare(cnt, p * J).pvalue > 1e-4
    # weight = 1 / (J p) exactly up to rounding
    np.testing.assert_allclose(w * J * p[key], 1.0, rtol=1e-10)


def test_sts_chisquare_4_modes(orc):
    rng = np.random.default_rng(23)
    dims = [6, 5, 4, 3]
    R = 4
    U = normalized_factors(rng, dims, R)
    s = orc.STS(U)
    J = 300000
    j = 2
    smp, w, _ = s.draw(j, J, rng.random((4, J)), rng.random((4, J)))
    K = np.einsum("ar,br,cr->abcr", U[0], U[1], U[3]).reshape(-1, R)
    Q, _ = np.linalg.qr(K)
    p = (Q ** 2).sum(1)
    p /= p.sum()
    key = (smp[:, 0].astype(np.int64) * dims[1] + smp[:, 1]) * dims[3] + smp[:, 3]
    cnt = np.bincount(key, minlength=len(p))
    assert chisquare(cnt, p * J).pvalue > 1e-4


# ---------------------------------------------------------------- ALS pieces
def test_dedup_vs_numpy_unique(orc):
    rng = np.random.default_rng(31)
    dims = [9, 4, 13]
    s = np.stack([rng.integers(0, d, 3000) for d in dims], 1).astype(np.uint32)
    for j in range(3):
        cols = [k for k in range(3) if k != j]
        key = [tuple(r) for r in s[:, cols]]
        wmap = {}
        for kk in key:
            wmap.setdefault(kk, rng.random())
        w = np.array([wmap[kk] for kk in key])
        sd, wd = orc.dedup(s, w, j, dims)
        u, cnt = np.unique(s[:, cols], axis=0, return_counts=True)
        assert np.array_equal(sd[:, cols], u)
        np.testing.assert_array_equal(wd, np.array([wmap[tuple(r)] for r in u]) * cnt)


def test_sampled_mttkrp_golden(orc):
    z = np.load(os.path.join(GOLDEN, "tiny_mttkrp.npz"))
    dims = list(z["dims"])
    for mode in range(3):
        lk = orc.Lookup(z["idx"], z["vals"], mode, dims)
        Y, H = lk.sampled_mttkrp(z[f"samples{mode}"], z[f"design{mode}"])
        np.testing.assert_allclose(Y, z[f"Y{mode}"], rtol=1e-13, atol=1e-13)
        assert H == int(z[f"hits{mode}"])


def test_exact_mttkrp_and_fit_brute_force(orc):
    rng = np.random.default_rng(32)
    dims = [8, 7, 6]
    idx, vals = random_tensor(rng, dims, 120)
    dense = np.zeros(dims)
    dense[tuple(idx.T.astype(np.int64))] = vals
    R = 3
    U = normalized_factors(rng, dims, R)
    sigma = rng.random(R) + 0.5
    for j in range(3):
        lk = orc.Lookup(idx, vals, j, dims)
        others = [k for k in range(3) if k != j]
        K = np.einsum("ar,br->abr", U[others[0]], U[others[1]])
        X = np.moveaxis(dense, j, 0)
        np.testing.assert_allclose(lk.exact_mttkrp(U), np.einsum("iab,abr->ir", X, K), rtol=1e-12, atol=1e-12)
    lk0 = orc.Lookup(idx, vals, 0, dims)
    xhat = np.einsum("r,ar,br,cr->abc", sigma, *U)
    norm = np.linalg.norm(vals)
    expect = 1 - np.linalg.norm(dense - xhat) / norm
    assert abs(lk0.fit(U, sigma, norm) - expect) < 1e-12


def test_design_closed_form(orc):
    rng = np.random.default_rng(33)
    dims = [10, 9, 8]
    R = 4
    U = normalized_factors(rng, dims, R)
    sd = np.stack([rng.integers(0, d, 50) for d in dims], 1).astype(np.uint32)
    sd, _ = orc.dedup(sd, np.ones(50), 1, dims)
    wd = rng.random(len(sd)) + 0.1
    D, Gi = orc.design(sd, wd, 1, U)
    krp = U[0][sd[:, 0]] * U[2][sd[:, 2]]
    np.testing.assert_allclose(D, krp * wd[:, None], rtol=1e-13)
    np.testing.assert_allclose(Gi, np.linalg.pinv(krp.T @ (wd[:, None] * krp)), rtol=1e-9)


def test_als_recovers_planted_low_rank(orc):
    """Exactly low-rank tensor observed on all entries: randomized ALS fit -> 1."""
    rng = np.random.default_rng(34)
    dims = [12, 11, 10]
    R = 2
    A = [rng.standard_normal((d, R)) for d in dims]
    dense = np.einsum("ar,br,cr->abc", *A)
    idx = np.argwhere(np.ones(dims)).astype(np.uint32)
    vals = dense[tuple(idx.T.astype(np.int64))]
    U0 = normalized_factors(rng, dims, R)
    for sampler in ("sts", "arls"):
        a = orc.ALS(idx, vals, dims, U0, sampler, seed=5)
        # 10 sweeps: later the residual^2 rounds below zero and the unfloored
        # fit formula (sparse_tensor.hpp:351-355) gives NaN, as it does there
        fits, st = a.run(10, 400, 5)
        assert fits[-1] > 0.9999, (sampler, fits)
        assert st["nonzeros_iterated"] > 0


def test_exact_als_step_is_least_squares_optimum(orc):
    """ExactALS step (exact_als.hpp:95-141): U_j sigma is the least-squares
    solution of X_(j) ~ U_j (KRP of the other factors)^T, here solved by
    numpy's SVD-based lstsq on the dense unfolding; columns unit norm."""
    rng = np.random.default_rng(35)
    dims = [9, 8, 7]
    R = 3
    idx, vals = random_tensor(rng, dims, 200)
    dense = np.zeros(dims)
    dense[tuple(idx.T.astype(np.int64))] = vals
    U0 = normalized_factors(rng, dims, R)
    a = orc.ALS(idx, vals, dims, U0, "exact")
    for j in range(3):
        others = [k for k in range(3) if k != j]
        K = np.einsum("ar,br->abr", a.U[others[0]], a.U[others[1]]).reshape(-1, R)
        Xj = np.moveaxis(dense, j, 0).reshape(dims[j], -1)
        sol = np.linalg.lstsq(K, Xj.T, rcond=None)[0].T   # I_j x R
        a.step(j, 0, 1)
        nrm = np.linalg.norm(sol, axis=0)
        np.testing.assert_allclose(a.sigma, nrm, rtol=1e-10)
        np.testing.assert_allclose(a.U[j], sol / nrm, rtol=1e-9, atol=1e-11)


def test_exact_als_fit_is_monotone_and_recovers_planted(orc):
    rng = np.random.default_rng(36)
    dims = [12, 11, 10]
    R = 2
    A = [rng.standard_normal((d, R)) for d in dims]
    dense = np.einsum("ar,br,cr->abc", *A)
    idx = np.argwhere(np.ones(dims)).astype(np.uint32)
    vals = dense[tuple(idx.T.astype(np.int64))]
    U0 = normalized_factors(rng, dims, R)
    a = orc.ALS(idx, vals, dims, U0, "exact")
    # 15 sweeps: later the residual^2 rounds below zero (unfloored fit, NaN)
    fits, st = a.run(15, 0, 1)
    assert np.all(np.diff(fits) > -1e-10), fits
    assert fits[-1] > 0.9999, fits
    assert st["nonzeros_iterated"] == 0   # ExactALS does not count them
    # sparse tensor: exact ALS is monotone as well
    idx, vals = random_tensor(rng, [30, 20, 25], 900)
    a = orc.ALS(idx, vals, [30, 20, 25], normalized_factors(rng, [30, 20, 25], 4), "exact")
    fits, _ = a.run(6, 0, 1)
    assert np.all(np.diff(fits) > -1e-10), fits


def _descend_fsum(G, n, F, lfill, node_count, nodes_upto, cl_off, U, x, u, shift=0):
    """A second evaluation of the PartitionTree descent: every node mass as
    the exactly rounded sum (math.fsum) of its products G_ab x_a x_b, every
    leaf row product u_r . x likewise.  shift moves the chosen leaf row (a
    planted mistake)."""
    R = len(x)
    outer = np.outer(x, x)

    def mass(c):
        Gc = np.triu(G[c]) + np.triu(G[c], 1).T
        return math.fsum((Gc * outer).ravel())

    m = mass(0)
    c, low, high = 0, 0.0, 1.0
    for _ in range(lfill):
        cut = low + mass(2 * c + 1) / m
        c, low, high = (2 * c + 1, low, cut) if u <= cut else (2 * c + 2, cut, high)
    if 2 * c + 1 < node_count:
        cut = low + mass(2 * c + 1) / m
        c, low, high = (2 * c + 1, low, cut) if u <= cut else (2 * c + 2, cut, high)
    leaf = c - nodes_upto if c >= nodes_upto else c - cl_off
    rows = min(F, n - leaf * F)
    q2 = [math.fsum(U[leaf * F + r] * x) ** 2 if r < rows else 0.0 for r in range(F)]
    total = math.fsum(q2)
    frac = (u - low) / (high - low)
    res, run = F - 1, 0.0
    for r in range(F - 1):
        run += q2[r]
        if frac < run / total:
            res = r
            break
    res = min(res, rows - 1)
    return leaf * F + min(res + shift, rows - 1)


def _bound_case(orc, scale, noise, J=3000):
    """Rows (scale x a common column) + noise, queries orthogonal to the
    all-ones direction: u_r . x is the noise part while the Gram entries
    carry the common one, so the node masses cancel by about (scale/noise)^2."""
    rng = np.random.default_rng(44)
    n, R = 3000, 6
    U = scale * rng.standard_normal((n, 1)) + noise * rng.standard_normal((n, R))
    t = orc.Tree(n, R, R)
    t.build(U)
    x = rng.standard_normal((J, R))
    x -= x.mean(axis=1, keepdims=True)
    draws = rng.random(J)
    samples, _, margin = t.sample(U, x, draws)
    tt = t.t
    geo = (n, R, tt.lfill_level, tt.node_count, tt.nodes_upto_lfill, tt.complete_level_offset)
    return t.G, geo, U, x, draws, samples, margin, t.slack


@pytest.mark.parametrize("scale, noise", [(100.0, 1e-3), (10.0, 1e-2)])
def test_boundary_rounding_bound_covers_reordered_sums(orc, scale, noise):
    """The oracle's slack (distance to each compared boundary minus that
    boundary's rounding bound) against a second evaluation of the same
    descent that sums every mass and row product exactly rounded (fsum).
    With masses cancelling by about 1e10 the two evaluations put some draws
    in different bins, and every such draw has slack <= 0 (the bound is
    sound).  With cancellation of about 1e6 few draws are ambiguous, and a
    planted mistake (the chosen row moved by one) disagrees on draws whose
    slack is positive (the bound leaves the check its power)."""
    G, geo, U, x, draws, samples, margin, slack = _bound_case(orc, scale, noise)
    n, F, lfill, nc, nu, cl = geo
    J = len(draws)
    other = np.array([_descend_fsum(G, n, F, lfill, nc, nu, cl, U, x[i], draws[i]) for i in range(J)])
    diff = other != samples
    ambiguous = slack <= 0
    print(f"scale {scale}, noise {noise}: draws in another bin under fsum {diff.sum()}, ambiguous by the "
          f"bound {ambiguous.sum()}, below the 1e-12 margin {(margin < 1e-12).sum()}")
    assert np.all(ambiguous[diff])
    if scale == 100.0:
        assert diff.sum() > 0
    else:
        assert ambiguous.mean() < 0.01
        planted = np.array([_descend_fsum(G, n, F, lfill, nc, nu, cl, U, x[i], draws[i], shift=1)
                            for i in range(300)])
        wrong = (planted != samples[:300]) & ~ambiguous[:300]
        assert wrong.sum() > 150
