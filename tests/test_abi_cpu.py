"""CPU tests of the boundary: the C-ABI library loads and exports every
symbol declared in include/exafac_hip.h; errors come back as status codes
(no crash) without a GPU; the pybind11 module exposes the reference's Python
surface (exafac/cpp_ext/py_module.cpp:37-77).  No compute runs here.
"""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "exafac_hip.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ex_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from rdist_tensor_amd import abi
    L = abi.lib()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    # the ctypes binding covers the header exactly
    assert set(abi.SIGNATURES) | {"ex_last_error"} == set(names)


def test_exported_dynamic_symbols_match_header():
    from rdist_tensor_amd import abi
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (ex_[a-z0-9_]+)$", out, flags=re.M))
    assert exported == set(header_functions())


def test_errors_are_status_codes_without_gpu():
    from rdist_tensor_amd import abi
    L = abi.lib()
    h = C.c_void_p()
    rc = L.ex_ctx_create(0, C.byref(h))
    if rc == abi.EX_OK:
        L.ex_ctx_destroy(h)
        pytest.skip("a GPU is visible")
    assert rc == abi.EX_ERR_HIP
    assert L.ex_last_error()
    # null arguments are rejected, not dereferenced
    assert L.ex_ctx_create(0, None) == abi.EX_ERR_ARG
    assert L.ex_tensor_norm(None, None) == abi.EX_ERR_ARG
    assert L.ex_als_run(None, 1, 1, 1) == abi.EX_ERR_ARG
    assert b"null" in L.ex_last_error()
    assert L.ex_batch_run(None, 1, None) == abi.EX_ERR_ARG
    assert L.ex_batch_create(None, 0, 1024, C.byref(h)) == abi.EX_ERR_ARG
    assert L.ex_batch_destroy(None) == abi.EX_OK


def test_pymodule_surface():
    import rdist_tensor_amd
    m = rdist_tensor_amd.py_module()
    for name in ["Grid", "TensorGrid", "DistMat1D", "LowRankTensor", "SparseTensor", "ALS_Optimizer", "ExactALS",
                 "TensorStationary", "AccumulatorStationary", "Sampler", "CP_ARLS_LEV", "STS_CP"]:
        assert hasattr(m, name), name
    for meth in ["initialize_ground_truth_for_als", "execute_ALS_rounds", "get_statistics_json",
                 "compute_exact_fit", "deinitialize"]:
        assert hasattr(m.ALS_Optimizer, meth)
    for meth in ["test_gram_matrix_computation", "initialize_factors_deterministic",
                 "initialize_factors_gaussian_random"]:
        assert hasattr(m.LowRankTensor, meth)
    assert issubclass(m.STS_CP, m.Sampler) and issubclass(m.CP_ARLS_LEV, m.Sampler)
    assert issubclass(m.AccumulatorStationary, m.ALS_Optimizer)


def test_grid_is_single_rank():
    import numpy as np
    import rdist_tensor_amd
    m = rdist_tensor_amd.py_module()
    g = m.Grid(np.array([1, 1, 1], dtype=np.uint64))   # forcecast, as get_best_mpi_dim returns uint64
    assert g.get_dimension() == 3
    with pytest.raises(RuntimeError):
        m.Grid(np.array([2, 1, 1]))
    tg = m.TensorGrid(np.array([10, 20, 30], dtype=np.int32), g)
    assert list(tg.tensor_dims) == [10, 20, 30]
    with pytest.raises(RuntimeError):
        m.TensorGrid(np.array([10, 20]), g)
    d = m.DistMat1D(4, tg, 1)
    assert d.padded_rows == 21 and d.data.shape == (21, 4)


def test_exafac_alias():
    import rdist_tensor_amd
    rdist_tensor_amd.install_exafac_alias()
    from exafac.cpp_ext.py_module import STS_CP, AccumulatorStationary  # noqa: F401


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_tuning_knobs_and_file(tmp_path):
    """Launch-variant knobs: names and variant counts, set / load with
    errors for unknown names, out-of-range values and malformed lines (host
    only: no GPU call)."""
    from rdist_tensor_amd import abi
    t = abi.tuning()
    assert {"sts_draw.fp64.r32", "sts_draw.fp32.r32", "row_scan.fp32"} <= set(t)
    before = {k: v for k, (n, v) in t.items()}
    try:
        abi.tuning_set("sts_draw.fp32.r32", 2)
        assert abi.tuning()["sts_draw.fp32.r32"][1] == 2
        for bad in [("no_such_knob", 0), ("sts_draw.fp32.r32", 99), ("sts_draw.fp32.r32", -1)]:
            with pytest.raises(abi.ExafacError) as e:
                abi.tuning_set(*bad)
            assert e.value.code == abi.EX_ERR_ARG
        f = tmp_path / "tuning.txt"
        f.write_text("# comment\nsts_draw.fp64.r32 3   # trailing comment\n\nrow_scan.fp64 1\n")
        abi.tuning_load(f)
        assert abi.tuning()["sts_draw.fp64.r32"][1] == 3 and abi.tuning()["row_scan.fp64"][1] == 1
        for text in ["row_scan.fp64 7\n", "bogus 0\n", "row_scan.fp64\n"]:
            f.write_text(text)
            with pytest.raises(abi.ExafacError) as e:
                abi.tuning_load(f)
            assert e.value.code == abi.EX_ERR_ARG
        with pytest.raises(abi.ExafacError):
            abi.tuning_load(tmp_path / "missing.txt")
    finally:
        for k, v in before.items():
            abi.tuning_set(k, v)
