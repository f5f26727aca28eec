"""The Gram entry point that allocates nothing (ex_gram_dev_static, device
pointers, partial sums in the library's static per-GPU scratch array)
against ex_gram, which allocates its buffers: three calls in a row with
shrinking inputs (the first needs the block count capped to fit the
scratch), each within 1e-12 relative of ex_gram's result."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_gram_dev_static_matches_gram(ctx):
    import torch
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(7)
    for m, R in ((70001, 64), (5000, 32), (1000, 16)):
        A = rng.standard_normal((m, R))
        want = ctx.gram(A)
        dA = torch.from_numpy(A).to("cuda")
        dG = torch.full((R, R), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        abi.check(abi.lib().ex_gram_dev_static(ctx.h, C.c_void_p(dA.data_ptr()), m, R, C.c_void_p(dG.data_ptr())))
        ctx.sync()
        got = dG.cpu().numpy()
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"m={m} R={R}: rel err vs ex_gram {err:.2e}", flush=True)
        assert err < 1e-12
