"""Sampler distributions on the GPU against exact leverage probabilities.

The samplers draw rows of the Khatri-Rao product of every factor but the
target's.  CP-ARLS-LEV draws each mode's row from that factor's leverage
scores independently (cp_arls_lev.hpp:36-163); STS-CP draws exactly from the
leverage scores of the Khatri-Rao product itself (sts_cp.hpp:203-300).  Both
distributions are computed here in closed form with numpy over every KRP row
of a small problem, and the device's draws (its own Philox stream, no
injected uniforms) are checked with a chi-square goodness-of-fit test and a
total-variation bound.  The weights are checked against 1 / (J p_i).

The mixed-precision samplers build their trees from BF16 Gram products, so
their probabilities carry relative errors around 1e-3: far below what the
chi-square test resolves at these sample counts, and the weight tolerance is
widened to 2e-2 for them.
"""
import numpy as np
import pytest
from scipy import stats

from conftest import normalized_factors

pytestmark = pytest.mark.gpu


def krp_rows(U, j):
    """All rows of the KRP over modes != j, in lexicographic order of the
    remaining coordinates (first remaining mode slowest)."""
    modes = [k for k in range(len(U)) if k != j]
    H = np.ones((1, U[0].shape[1]))
    for k in modes:
        H = (H[:, None, :] * U[k][None, :, :]).reshape(-1, U[0].shape[1])
    return H, modes


def exact_sts_probs(U, j):
    H, modes = krp_rows(U, j)
    G = np.ones((U[0].shape[1],) * 2)
    for k in modes:
        G = G * (U[k].T @ U[k])
    lev = np.einsum("ir,rs,is->i", H, np.linalg.pinv(G), H)
    return lev / lev.sum(), modes


def exact_arls_probs(U, j):
    modes = [k for k in range(len(U)) if k != j]
    p = np.ones(1)
    for k in modes:
        Gk = U[k].T @ U[k]
        lk = np.einsum("ir,rs,is->i", U[k], np.linalg.pinv(Gk), U[k])
        p = (p[:, None] * (lk / lk.sum())[None, :]).reshape(-1)
    return p, modes


def bins_of(samples, modes, dims):
    lin = np.zeros(len(samples), np.int64)
    for k in modes:
        lin = lin * dims[k] + samples[:, k].astype(np.int64)
    return lin


def chi_square_p(counts, p, J):
    E = p * J
    keep = E >= 5.0
    # pool the sparse bins into one so that no probability mass is dropped
    O = np.append(counts[keep], counts[~keep].sum())
    Ek = np.append(E[keep], E[~keep].sum())
    if Ek[-1] < 5.0:
        O, Ek = O[:-1], Ek[:-1]
    chi2 = ((O - Ek) ** 2 / Ek).sum()
    return stats.chi2.sf(chi2, len(O) - 1), chi2, len(O) - 1


def make_factors(ctx, dims, R, seed, skew):
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(seed)
    U = normalized_factors(rng, dims, R)
    # a few heavy rows give the leverage distribution some shape
    for u in U:
        u[: max(1, len(u) // 10)] *= skew
    U = [u / np.linalg.norm(u, axis=0) for u in U]
    F = abi.Factors(ctx, dims, R)
    for k in range(len(dims)):
        F.set(k, U[k])
    return U, F


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", ["sts", "arls"])
def test_draw_distribution_matches_exact_leverage(ctx, kind, precision):
    from rdist_tensor_amd import abi
    dims = [40, 30, 20]
    R = 8
    U, F = make_factors(ctx, dims, R, seed=61, skew=4.0)
    sk = abi.EX_SAMPLER_STS_CP if kind == "sts" else abi.EX_SAMPLER_CP_ARLS_LEV
    prec = abi.EX_MIXED if precision == "mixed" else abi.EX_FP64
    S = abi.Sampler(ctx, F, sk, precision=prec, seed=123)
    J = 200_000
    wtol = 2e-2 if precision == "mixed" else 1e-10
    for j in range(3):
        p, modes = exact_sts_probs(U, j) if kind == "sts" else exact_arls_probs(U, j)
        samples, w = S.draw(J, j, sweep=3 + j)
        b = bins_of(samples, modes, dims)
        assert b.min() >= 0 and b.max() < len(p)
        counts = np.bincount(b, minlength=len(p)).astype(np.float64)
        pval, chi2, df = chi_square_p(counts, p, J)
        tv = 0.5 * np.abs(counts / J - p).sum()
        print(f"{kind}/{precision} mode {j}: chi2 {chi2:.1f} df {df} p {pval:.3g} TV {tv:.4f}")
        assert pval > 1e-4
        # total variation of an exact sampler's histogram: about
        # 0.5 * sum_i sqrt(2 p_i (1 - p_i) / (pi J)) (normal approximation)
        tv_exp = 0.5 * np.sqrt(2.0 * p * (1.0 - p) / (np.pi * J)).sum()
        assert tv < 1.3 * tv_exp
        # weight of a draw = 1 / (J p_i)
        w_exact = 1.0 / (J * p[b])
        assert np.max(np.abs(w / w_exact - 1.0)) < wtol


def test_chi_square_has_power(ctx):
    """The same statistic rejects a plausible mistake: drawing the STS rows
    from the product of per-mode leverages (the CP-ARLS-LEV distribution)."""
    from rdist_tensor_amd import abi
    dims = [40, 30, 20]
    R = 8
    U, F = make_factors(ctx, dims, R, seed=61, skew=4.0)
    S = abi.Sampler(ctx, F, abi.EX_SAMPLER_CP_ARLS_LEV, seed=123)
    J = 200_000
    p_sts, modes = exact_sts_probs(U, 0)
    samples, _ = S.draw(J, 0, sweep=3)
    counts = np.bincount(bins_of(samples, modes, dims), minlength=len(p_sts)).astype(np.float64)
    pval, chi2, df = chi_square_p(counts, p_sts, J)
    print(f"ARLS draws against STS probabilities: chi2 {chi2:.1f} df {df} p {pval:.3g}")
    assert pval < 1e-6


@pytest.mark.parametrize("precision,R", [("mixed", 16), ("mixed", 25), ("mixed", 32), ("mixed", 40),
                                         ("mixed", 64), ("fp64", 40), ("fp64", 64)])
def test_sts_draw_distribution_every_rank_class(ctx, precision, R):
    """STS-CP draws at every lane layout of the draw kernel against exact
    leverage (chi-square, total variation, weights): mixed precision with 16
    lanes per sample (R <= 16), 8 lanes and seventeen float4 per lane (R <=
    32; R = 25 reads leaf rows element by element), 32 lanes (R <= 45) and 64
    lanes (R > 45); FP64 with 64 lanes and nine or seventeen double2 per lane.
    Mode sizes above R keep every factor Gram full rank.  Draws given weight 0
    (a zero-mass leaf reached through FP32 rounding) are counted by the
    library and must stay below one in 10^4."""
    from rdist_tensor_amd import abi
    dims = [80, 70, 60]
    U, F = make_factors(ctx, dims, R, seed=70 + R, skew=4.0)
    prec = abi.EX_MIXED if precision == "mixed" else abi.EX_FP64
    S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP, precision=prec, seed=321 + R)
    J = 200_000
    wtol = 5e-2 if precision == "mixed" else 1e-9
    total = 0
    for j in range(3):
        p, modes = exact_sts_probs(U, j)
        samples, w = S.draw(J, j, sweep=5 + j)
        total += J
        b = bins_of(samples, modes, dims)
        assert b.min() >= 0 and b.max() < len(p)
        counts = np.bincount(b, minlength=len(p)).astype(np.float64)
        pval, chi2, df = chi_square_p(counts, p, J)
        tv = 0.5 * np.abs(counts / J - p).sum()
        tv_exp = 0.5 * np.sqrt(2.0 * p * (1.0 - p) / (np.pi * J)).sum()
        ok = w > 0
        werr = np.max(np.abs(w[ok] * J * p[b[ok]] - 1.0))
        print(f"sts/{precision} R={R} mode {j}: chi2 {chi2:.1f} df {df} p {pval:.3g} TV {tv:.4f} "
              f"(exp {tv_exp:.4f}) max weight rel err {werr:.2e} zero-weight {int((~ok).sum())}")
        assert pval > 1e-4
        assert tv < 1.3 * tv_exp
        assert werr < wtol
    dropped = S.dropped_draws()
    print(f"sts/{precision} R={R}: draws given weight 0: {dropped} of {total}")
    assert dropped <= total * 1e-4
    if precision == "fp64":
        assert dropped == 0
