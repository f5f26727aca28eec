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


def test_tensor_file_reader_and_factor_checkpoint(tmp_path, orc):
    """A FROSTT .tns file (1-based, minimum above 1) through the converter and
    DistSparseTensorE (coordinates shifted by the minimum, log_count
    preprocessing, sparse_tensor_e.py:63-117) into the drop-in SparseTensor;
    its ALS fits equal the oracle's on the shifted arrays.  Then a factor
    checkpoint written and read back restores factors and column norms
    bitwise."""
    from rdist_tensor_amd import tensor_io
    m = _module()
    rng = np.random.default_rng(90)
    dims = [120, 90, 60]
    idx, vals = random_tensor(rng, dims, 6000)
    raw = idx.astype(np.int64) + np.array([3, 1, 7])      # on disk: 1-based, shifted minima
    vals = rng.integers(1, 20, len(vals)).astype(np.float64)
    p = tmp_path / "x.tns"
    with open(p, "w") as f:
        for c, v in zip(raw, vals):
            f.write(f"{c[0]} {c[1]} {c[2]} {v:.1f}\n")
    conv = tensor_io.tns_to_tensor_file(p)
    t = tensor_io.DistSparseTensorE(conv, preprocessing="log_count")
    shift = raw - raw.min(0)
    assert np.array_equal(t.tensor_idxs, shift.astype(np.uint32))
    assert list(t.tensor_dims) == list(raw.max(0) - raw.min(0) + 1)
    lv = np.log(vals + 1)
    assert abs(t.sparse_tensor.tensor_norm - np.sqrt(np.sum(lv * lv))) < 1e-12 * np.sqrt(np.sum(lv * lv))
    lrt = m.LowRankTensor(6, t.tensor_grid)
    lrt.initialize_factors_gaussian_random(3)
    U0 = [lrt.get_factor(k).copy() for k in range(3)]
    opt = m.AccumulatorStationary(t.sparse_tensor, lrt, m.STS_CP(lrt, seed=2))
    opt.initialize_ground_truth_for_als()
    opt.execute_ALS_rounds(2, 1024, 1)
    fits = json.loads(opt.get_statistics_json())["fits"][1:]
    oa = orc.ALS(shift.astype(np.uint32), lv, [int(d) for d in t.tensor_dims], U0, "sts", seed=2)
    ofits, _ = oa.run(2, 1024, 1)
    np.testing.assert_allclose(fits, ofits, atol=1e-9)
    ck = tensor_io.write_factors(lrt, tmp_path / "factors.npz")
    other = m.LowRankTensor(6, t.tensor_grid)
    other.initialize_factors_gaussian_random(9)
    tensor_io.read_factors(other, ck)
    for k in range(3):
        assert np.array_equal(other.get_factor(k), lrt.get_factor(k))
    assert np.array_equal(other.get_sigma(), lrt.get_sigma())


def test_synthetic_block_tensor():
    """SparseTensor(tensor_grid, I, N, Q, seed) (sparse_tensor.hpp:102-215):
    Q disjoint all-ones blocks, so the norm is sqrt(nnz), nnz is near its
    expectation I^3 / Q^2, a seed reproduces the tensor, and exact ALS at
    rank Q fits it closely."""
    m = _module()
    I, Q = 60, 4
    grid = m.Grid(np.ones(3, dtype=np.int32))
    tg = m.TensorGrid(np.array([I, I, I], dtype=np.int32), grid)
    st = m.SparseTensor(tg, I, 3, Q, seed=5)
    again = m.SparseTensor(tg, I, 3, Q, seed=5)
    assert st.nnz == again.nnz > 0
    assert abs(st.tensor_norm - np.sqrt(st.nnz)) < 1e-12 * np.sqrt(st.nnz)
    assert 0.3 * I ** 3 / Q ** 2 < st.nnz < 3.0 * I ** 3 / Q ** 2
    with pytest.raises(ValueError):
        m.SparseTensor(tg, I, 4, Q, 5)
    lrt = m.LowRankTensor(Q, tg)
    lrt.initialize_factors_gaussian_random(1)
    opt = m.ExactALS(st, lrt)
    opt.initialize_ground_truth_for_als()
    opt.execute_ALS_rounds(30, 0, 10)
    fits = json.loads(opt.get_statistics_json())["fits"]
    print(f"block tensor nnz {st.nnz}, exact-ALS fits {fits}")
    assert fits[-1] > 0.5
