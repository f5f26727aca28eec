import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def ctx():
    import torch  # noqa: F401  (tests hand torch device tensors to the library: its runtime starts first)
    from rdist_tensor_amd import abi
    c = abi.Context(0)
    print(f"\n[gpu] device: {c.device_name()}", flush=True)
    yield c
    c.close()


def random_tensor(rng, dims, nnz, values="uniform"):
    """Distinct uniform-random coordinates (dedup on a linear index)."""
    total = int(np.prod(np.asarray(dims, dtype=np.float64)))
    nnz = min(nnz, total)
    lin = np.unique(rng.integers(0, total, size=int(nnz * 1.3) + 16, dtype=np.int64))
    rng.shuffle(lin)
    lin = lin[:nnz]
    idx = np.empty((len(lin), len(dims)), np.uint32)
    rem = lin.copy()
    for k in range(len(dims) - 1, -1, -1):
        idx[:, k] = rem % dims[k]
        rem //= dims[k]
    if values == "uniform":
        vals = rng.random(len(lin))
    else:
        vals = rng.lognormal(0, 1, len(lin))
    return idx, vals


def normalized_factors(rng, dims, R):
    U = [rng.standard_normal((d, R)) for d in dims]
    return [u / np.linalg.norm(u, axis=0) for u in U]
