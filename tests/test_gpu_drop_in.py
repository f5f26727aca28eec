"""The drop-in surface on the GPU: decompose.py's call sequence
(decompose.py:9-115) through `exafac.cpp_ext.py_module` as registered by
install_exafac_alias(), checked against the oracle started from the same
factors (FP64: fits within 1e-9, the nonzeros iterated equal)."""
import json

import numpy as np
import pytest

from conftest import random_tensor

pytestmark = pytest.mark.gpu


def _module():
    import rdist_tensor_amd
    rdist_tensor_amd.install_exafac_alias()
    from exafac.cpp_ext import py_module
    return py_module


@pytest.mark.parametrize("algorithm", ["sts", "cp_arls_lev", "exact"])
@pytest.mark.parametrize("distribution", ["accumulator_stationary", "tensor_stationary"])
def test_decompose_sequence_matches_oracle(orc, algorithm, distribution):
    m = _module()
    if algorithm == "exact" and distribution == "tensor_stationary":
        pytest.skip("decompose.py builds ExactALS for either distribution")
    rng = np.random.default_rng(81)
    dims = [300, 250, 200]
    R, J, rounds = 8, 2048, 3
    idx, vals = random_tensor(rng, dims, 30000)
    grid = m.Grid(np.ones(3, dtype=np.int32))
    tg = m.TensorGrid(np.array(dims, dtype=np.int32), grid)
    st = m.SparseTensor(tg, idx, vals, "none")
    lrt = m.LowRankTensor(R, tg)
    lrt.initialize_factors_gaussian_random(5)
    U0 = [lrt.get_factor(k).copy() for k in range(3)]
    if algorithm == "exact":
        opt = m.ExactALS(st, lrt)
    else:
        sampler = m.STS_CP(lrt, seed=11) if algorithm == "sts" else m.CP_ARLS_LEV(lrt, seed=11)
        cls = m.AccumulatorStationary if distribution == "accumulator_stationary" else m.TensorStationary
        opt = cls(st, lrt, sampler)
    opt.initialize_ground_truth_for_als()
    fit0 = opt.compute_exact_fit()
    opt.execute_ALS_rounds(rounds, J, 1)
    stats = json.loads(opt.get_statistics_json())
    for key in ["num_rounds", "als_total_time", "fit_computation_time", "leverage_sampling_time",
                "row_gather_time", "design_matrix_prep_time", "spmm_time", "avg_nonzeros_iterated",
                "sum_nonzeros_iterated", "dense_reduce_time", "postprocessing_time", "sampler_update_time",
                "rounds", "fits", "als_times", "fit_computation_times"]:
        assert key in stats, key
    kind = {"sts": "sts", "cp_arls_lev": "arls", "exact": "exact"}[algorithm]
    oa = orc.ALS(idx, vals, dims, U0, kind, seed=11)
    assert abs(fit0 - oa.fit()) < 1e-12
    ofits, ost = oa.run(rounds, J, 1)
    print(f"{algorithm}/{distribution}: fits {stats['fits'][1:]} oracle {list(ofits)}")
    np.testing.assert_allclose(stats["fits"][1:], ofits, atol=1e-9)
    assert stats["sum_nonzeros_iterated"] == ost["nonzeros_iterated"]
    for k in range(3):
        np.testing.assert_allclose(lrt.get_factor(k), oa.U[k], rtol=0, atol=1e-9)
