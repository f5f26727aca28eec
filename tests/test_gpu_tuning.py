"""Every launch variant the autotuner chooses between computes the same
quantities as the default (ex_tuning_*):

* FP64 STS-CP draw variants (lanes per sample, vectors per lane, threads per
  workgroup): index-exact against the oracle with injected uniforms, as the
  default is (draws inside a compared boundary's rounding bound exempt);
* mixed STS-CP draw variants: chi-square, total variation and weights
  against exact KRP leverage, as the default is (sts_first.fp32: the first
  non-target mode drawn from prefix sums of its leaf masses, or by tree
  descents like the other modes);
* row-pass variants (lanes per hit): the FP64 sampled MTTKRP against the
  oracle (1e-12 relative, hits exact); the mixed ALS sweep's fit within
  1e-5 of the default variant's (FP32 sums grouped differently)."""
import json

import numpy as np
import pytest

from conftest import normalized_factors, random_tensor
from test_gpu_sampling_stats import bins_of, chi_square_p, exact_sts_probs, make_factors

pytestmark = pytest.mark.gpu


@pytest.fixture
def knob():
    """Sets knobs for one test and restores every knob to its value before."""
    from rdist_tensor_amd import abi
    before = {k: v for k, (n, v) in abi.tuning().items()}
    yield abi.tuning_set
    for k, v in before.items():
        abi.tuning_set(k, v)


def rel_err(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name,variant,R", [("sts_draw.fp64.r16", v, 16) for v in range(4)] +
                         [("sts_draw.fp64.r32", v, 32) for v in range(6)])
def test_fp64_draw_variants_bitexact(ctx, orc, knob, name, variant, R):
    from rdist_tensor_amd import abi
    knob(name, variant)
    rng = np.random.default_rng(41 + R)
    dims = [3000, 1000, 517]
    U = normalized_factors(rng, dims, R)
    F = abi.Factors(ctx, dims, R)
    for k in range(3):
        F.set(k, U[k])
    S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP)
    so = orc.STS(U)
    J = 20000
    for j in range(3):
        de = rng.random((3, J))
        dr = rng.random((3, J))
        sg, wg = S.draw(J, j, draws_row=dr, draws_eig=de)
        s_o, wo, margin = so.draw(j, J, de, dr)
        near = (margin < 1e-12) | (so.slack <= 0)
        cols = [k for k in range(3) if k != j]
        mismatch = np.any(sg[:, cols] != s_o[:, cols], axis=1)
        print(f"{name}={variant} mode {j}: near-boundary {near.sum()}, mismatches {mismatch.sum()}")
        assert not np.any(mismatch & ~near)
        ok = ~mismatch
        assert rel_err(wg[ok], wo[ok]) < 1e-11


@pytest.mark.parametrize("name,variant,R", [("sts_draw.fp32.r16", v, 16) for v in range(4)] +
                         [("sts_draw.fp32.r32", v, 32) for v in range(7)] +
                         [("sts_draw.fp32.r64", v, 64) for v in range(4)] +
                         [("sts_first.fp32", v, R) for v in range(2) for R in (16, 32, 64)])
def test_mixed_draw_variants_distribution(ctx, knob, name, variant, R):
    from rdist_tensor_amd import abi
    knob(name, variant)
    dims = [80, 70, 60]
    U, F = make_factors(ctx, dims, R, seed=70 + R, skew=4.0)
    S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP, precision=abi.EX_MIXED, seed=321 + R)
    J = 200_000
    for j in range(3):
        p, modes = exact_sts_probs(U, j)
        samples, w = S.draw(J, j, sweep=5 + j)
        b = bins_of(samples, modes, dims)
        assert b.min() >= 0 and b.max() < len(p)
        counts = np.bincount(b, minlength=len(p)).astype(np.float64)
        pval, chi2, df = chi_square_p(counts, p, J)
        tv = 0.5 * np.abs(counts / J - p).sum()
        tv_exp = 0.5 * np.sqrt(2.0 * p * (1.0 - p) / (np.pi * J)).sum()
        ok = w > 0
        werr = np.max(np.abs(w[ok] * J * p[b[ok]] - 1.0))
        print(f"{name}={variant} mode {j}: chi2 {chi2:.1f} df {df} p {pval:.3g} TV {tv:.4f} (exp {tv_exp:.4f}) "
              f"weight rel err {werr:.2e}")
        assert pval > 1e-4
        assert tv < 1.3 * tv_exp
        assert werr < 5e-2
    assert S.dropped_draws() <= 3 * J * 1e-4


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("R", [8, 25, 32])
def test_row_pass_variants_fp64(ctx, orc, knob, variant, R):
    from rdist_tensor_amd import abi
    knob("row_scan.fp64", variant)
    rng = np.random.default_rng(11 + R)
    dims = [300, 200, 100]
    idx, vals = random_tensor(rng, dims, 40000)
    T = abi.Tensor(ctx, idx, vals, dims)
    T.mttkrp_path(abi.EX_MTTKRP_ROW_PASS)
    T.build_lookups()
    for mode in range(3):
        lk = orc.Lookup(idx, vals, mode, dims)
        pick = rng.integers(0, len(idx), 3000)
        s = np.concatenate([idx[pick], np.stack([rng.integers(0, d, 3000) for d in dims], 1).astype(np.uint32)])
        sd, _ = orc.dedup(s, np.ones(len(s)), mode, dims)
        design = rng.standard_normal((len(sd), R))
        Yo, Ho = lk.sampled_mttkrp(sd, design)
        Yg, Hg = T.sampled_mttkrp(mode, sd, design)
        assert Hg == Ho
        assert rel_err(Yg, Yo) < 1e-12


def test_row_pass_variants_mixed_sweep(ctx, knob):
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(5)
    dims = [2000, 1500, 1000]
    R = 32
    idx, vals = random_tensor(rng, dims, 200000)
    U0 = normalized_factors(rng, dims, R)
    fits = []
    for variant in (0, 1):
        knob("row_scan.fp32", variant)
        T = abi.Tensor(ctx, idx, vals, dims)
        T.mttkrp_path(abi.EX_MTTKRP_ROW_PASS)
        F = abi.Factors(ctx, dims, R)
        for k in range(3):
            F.set(k, U0[k])
        S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP, precision=abi.EX_MIXED, seed=3)
        A = abi.ALS(ctx, T, F, S, precision=abi.EX_MIXED)
        A.init()
        A.run(3, 8192, 3)
        fits.append(json.loads(A.stats_json())["fits"][-1])
        for x in (A, S, F, T):
            x.close()
    print(f"mixed fits by row-pass variant: {fits}")
    assert abs(fits[0] - fits[1]) < 1e-5
