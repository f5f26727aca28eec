"""Batched CP-ARLS-LEV, one kernel launch per sweep (ex_batch_*).

Each batch is checked against the same optimizers run alone through
ex_als_run (FP64: fits within 1e-9, the nonzeros iterated equal -- the
draws are the same -- and factors as close to the oracle's as the
optimizer's own) and against the oracle's ALS
(accumulator_stationary.hpp:139-295 with cp_arls_lev.hpp:25-163, FP64),
on tensors whose sampled fibers hit many nonzeros (every row of the
factors stays listed) and on tensors so sparse that a step's MTTKRP reaches
a few rows (the batch's leverage state then covers those rows only).  The
mixed-precision batch is held to the FP64 one at the fit (1e-3)."""
import json

import numpy as np
import pytest

from conftest import normalized_factors, random_tensor

pytestmark = pytest.mark.gpu

# (dims, nnz, R): R = 25 and 5 exercise ranks that are not powers of two
DENSE = [([60, 50, 40], 20000, 8), ([120, 90, 70], 40000, 16), ([200, 150, 30], 30000, 25),
         ([80, 300, 100], 50000, 32), ([40, 40, 40], 10000, 5)]
SPARSE = [([3000, 2500, 2000], 40000, 16), ([5000, 800, 4000], 60000, 32), ([1500, 1500, 1500], 30000, 8)]


def _specs(seed, shapes):
    rng = np.random.default_rng(seed)
    out = []
    for q, (dims, nnz, R) in enumerate(shapes):
        idx, vals = random_tensor(rng, dims, nnz)
        out.append((dims, idx, vals, R, 1000 + q, normalized_factors(rng, dims, R)))
    return out


def _make(ctx, specs, precision):
    from rdist_tensor_amd import abi
    prec = abi.EX_MIXED if precision == "mixed" else abi.EX_FP64
    objs = []
    for dims, idx, vals, R, seed, U0 in specs:
        T = abi.Tensor(ctx, idx, vals, dims)
        F = abi.Factors(ctx, dims, R)
        for k, u in enumerate(U0):
            F.set(k, u)
        S = abi.Sampler(ctx, F, abi.EX_SAMPLER_CP_ARLS_LEV, precision=prec, seed=seed)
        A = abi.ALS(ctx, T, F, S, precision=prec)
        A.init()
        objs.append((T, F, S, A))
    return objs


def _close(objs):
    for o in objs:
        for x in reversed(o):
            x.close()


@pytest.mark.parametrize("hits", ["dense", "sparse"])
def test_fp64_batch_equals_single_runs_and_oracle(ctx, orc, hits):
    from rdist_tensor_amd import abi
    specs = _specs(11 if hits == "dense" else 12, DENSE if hits == "dense" else SPARSE)
    J, rounds = 2048, 3
    single = _make(ctx, specs, "fp64")
    for T, F, S, A in single:
        A.run(rounds, J, 10 ** 9)
    sstats = [json.loads(o[3].stats_json()) for o in single]
    objs = _make(ctx, specs, "fp64")
    B = abi.Batch([o[3] for o in objs], J)
    sec = B.run(rounds)
    nz, ju, rows = B.counters()
    print(f"\n{hits}: {len(objs)} tensors, {rounds} sweeps in {sec * 1e3:.2f} ms; nonzeros iterated {list(nz)}")
    for q, ((dims, idx, vals, R, seed, U0), s1, b1) in enumerate(zip(specs, single, objs)):
        fs, fb = s1[3].fit(), b1[3].fit()
        oa = orc.ALS(idx, vals, dims, U0, "arls", seed=seed)
        ofits, ost = oa.run(rounds, J, rounds)
        print(f"  tensor {q} R={R}: fit batch {fb:.12f} single {fs:.12f} oracle {ofits[-1]:.12f}")
        assert abs(fb - fs) < 1e-9
        assert abs(fb - ofits[-1]) < 1e-9
        assert nz[q] == sstats[q]["sum_nonzeros_iterated"] == ost["nonzeros_iterated"]
        # factors: as close to the oracle's as the FP64 optimizer's own are
        # (1e-8, or 10x the optimizer's distance where few rows carry the
        # factor and its Gram is ill-conditioned)
        eb = max(np.abs(b1[1].get(k) - oa.U[k]).max() for k in range(3))
        es = max(np.abs(s1[1].get(k) - oa.U[k]).max() for k in range(3))
        print(f"    factor error vs oracle: batch {eb:.3e} single {es:.3e}")
        assert eb <= max(1e-8, 10 * es)
    B.close()
    _close(objs)
    _close(single)


def test_mixed_batch_fit_within_1e3_of_fp64(ctx):
    from rdist_tensor_amd import abi
    specs = _specs(13, DENSE)
    J, rounds = 2048, 4
    fits = {}
    for precision in ("fp64", "mixed"):
        objs = _make(ctx, specs, precision)
        B = abi.Batch([o[3] for o in objs], J)
        B.run(rounds)
        fits[precision] = [o[3].fit() for o in objs]
        B.close()
        _close(objs)
    print(f"\nfp64 {fits['fp64']}\nmixed {fits['mixed']}")
    for a, b in zip(fits["fp64"], fits["mixed"]):
        assert abs(a - b) < 1e-3


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_equals_eager(ctx, precision):
    """Sweeps replayed from the captured launch reach the eager batch's
    factors bitwise; a second run restarts the sweep numbering at 1 and
    continues from the factors reached."""
    from rdist_tensor_amd import abi
    specs = _specs(14, DENSE[:3] + SPARSE[:2])
    got = {}
    for graph in (False, True):
        objs = _make(ctx, specs, precision)
        B = abi.Batch([o[3] for o in objs], 1024)
        B.use_graph(graph)
        B.run(2)
        B.run(3)
        got[graph] = [[o[1].get(k) for k in range(3)] for o in objs]
        B.close()
        _close(objs)
    for a, b in zip(got[False], got[True]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_optimizers_continue_after_batch(ctx):
    """After a batch run the optimizers' own sampler state is rebuilt before
    their next draw: ex_als_run continues from the batch's factors exactly
    as it continues from its own (FP64: batch 2 sweeps then 1 alone equals
    alone 2 + 1, within 1e-9 at the fit)."""
    from rdist_tensor_amd import abi
    specs = _specs(15, DENSE[:3])
    J = 2048
    ref = _make(ctx, specs, "fp64")
    for o in ref:
        o[3].run(2, J, 10 ** 9)
        o[3].run(1, J, 10 ** 9)
    objs = _make(ctx, specs, "fp64")
    B = abi.Batch([o[3] for o in objs], J)
    B.run(2)
    for o in objs:
        o[3].run(1, J, 10 ** 9)
    for a, b in zip(ref, objs):
        assert abs(a[3].fit() - b[3].fit()) < 1e-9
    # the factors moved on outside the batch: it refuses to run
    with pytest.raises(abi.ExafacError) as e:
        B.run(1)
    assert e.value.code == abi.EX_ERR_STATE
    B.close()
    _close(objs)
    _close(ref)


def test_batch_envelope_errors(ctx):
    from rdist_tensor_amd import abi
    specs = _specs(16, DENSE[:1])
    objs = _make(ctx, specs, "fp64")
    A = objs[0][3]
    with pytest.raises(abi.ExafacError) as e:
        abi.Batch([A], 8193)
    assert e.value.code == abi.EX_ERR_UNSUPPORTED
    with pytest.raises(abi.ExafacError) as e:
        abi.Batch([A, A], 1024)
    assert e.value.code == abi.EX_ERR_ARG
    dims, idx, vals, R, seed, U0 = specs[0]
    T = abi.Tensor(ctx, idx, vals, dims)
    F = abi.Factors(ctx, dims, 40)
    F.init_deterministic()
    S = abi.Sampler(ctx, F, abi.EX_SAMPLER_CP_ARLS_LEV, seed=1)
    A40 = abi.ALS(ctx, T, F, S)
    A40.init()
    with pytest.raises(abi.ExafacError) as e:
        abi.Batch([A40], 1024)
    assert e.value.code == abi.EX_ERR_UNSUPPORTED
    F2 = abi.Factors(ctx, dims, 8)
    F2.init_deterministic()
    S2 = abi.Sampler(ctx, F2, abi.EX_SAMPLER_STS_CP, seed=1)
    A2 = abi.ALS(ctx, T, F2, S2)
    A2.init()
    with pytest.raises(abi.ExafacError) as e:
        abi.Batch([A2], 1024)
    assert e.value.code == abi.EX_ERR_UNSUPPORTED
    for x in (A2, S2, F2, A40, S, F, T):
        x.close()
    _close(objs)
