"""Differentiable sampled MTTKRP (rdist_tensor_amd.autograd): the forward
against the oracle's sampled MTTKRP on the oracle's design rows, the
hand-written adjoint against the FP64 gradient of a dense torch restatement
and against central finite differences (eps = 1e-3, FP64), and first-order
fitting (Adam, lr 1e-2, 200 steps) of the sampled least-squares objective.
There is no reference counterpart to the adjoint: the forward is pinned by
the sampled MTTKRP (sort_lookup.hpp:238-375), the adjoint by the dense
gradient and by finite differences."""
import numpy as np
import pytest

from conftest import normalized_factors, random_tensor

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def make_samples(orc, rng, idx, dims, mode, n_hit, n_rand):
    """Deduplicated sample tuples: rows of nonzeros (their fibers exist) and
    random tuples; positive weights."""
    s1 = idx[rng.integers(0, len(idx), n_hit)]
    s2 = np.stack([rng.integers(0, d, n_rand) for d in dims], 1).astype(np.uint32)
    s = np.concatenate([s1, s2]).astype(np.uint32)
    sd, _ = orc.dedup(s, np.ones(len(s)), mode, dims)
    wd = rng.random(len(sd)) + 0.5
    return sd, wd


def hits_of(idx, vals, sd, mode):
    """(output row, sample row, value) of every nonzero in a sampled fiber."""
    cols = [k for k in range(idx.shape[1]) if k != mode]
    lut = {tuple(r): c for c, r in enumerate(sd[:, cols].tolist())}
    rows, cs, xs = [], [], []
    for i, key in enumerate(map(tuple, idx[:, cols].tolist())):
        c = lut.get(key)
        if c is not None:
            rows.append(idx[i, mode])
            cs.append(c)
            xs.append(vals[i])
    return np.array(rows, np.int64), np.array(cs, np.int64), np.array(xs)


def dense_forward(torch, U, sd, wd, mode, hits, Ij):
    """The same Y in torch CPU FP64 (differentiable): design rows
    (prod U * sqrt(w)) * sqrt(w), summed over the hits."""
    rows, cs, xs = hits
    sdt = torch.as_tensor(sd.astype(np.int64))
    K = None
    for k, u in enumerate(U):
        if k == mode:
            continue
        r = u[sdt[:, k]]
        K = r if K is None else K * r
    sw = torch.sqrt(torch.as_tensor(wd))
    D = (K * sw[:, None]) * sw[:, None]
    contrib = torch.as_tensor(xs)[:, None] * D[torch.as_tensor(cs)]
    return torch.zeros((Ij, U[0].shape[1]), dtype=torch.float64).index_add(0, torch.as_tensor(rows), contrib)


def setup_case(ctx, orc, N, R, seed):
    import torch
    from rdist_tensor_amd import abi
    rng = np.random.default_rng(seed)
    dims = [400, 300, 200] if N == 3 else [60, 50, 40, 30]
    idx, vals = random_tensor(rng, dims, 30000)
    T = abi.Tensor(ctx, idx, vals, dims)
    T.build_lookups()
    U = normalized_factors(rng, dims, R)
    Ut = [torch.tensor(u, device="cuda") for u in U]
    return rng, dims, idx, vals, T, U, Ut


@pytest.mark.parametrize("N,R", [(3, 8), (3, 32), (4, 16)])
def test_forward_matches_oracle(ctx, orc, N, R):
    import torch
    from rdist_tensor_amd.autograd import SampledMTTKRP
    rng, dims, idx, vals, T, U, Ut = setup_case(ctx, orc, N, R, 90 + N + R)
    for mode in range(N):
        sd, wd = make_samples(orc, rng, idx, dims, mode, 3000, 1000)
        D, _ = orc.design(sd, wd, mode, U)
        Yo, _ = orc.Lookup(idx, vals, mode, dims).sampled_mttkrp(sd, D)
        Y = SampledMTTKRP.apply(T, mode, torch.tensor(sd.astype(np.int32), device="cuda"),
                                torch.tensor(wd, device="cuda"), *Ut)
        err = rel_err(Y.cpu().numpy(), Yo)
        print(f"N={N} R={R} mode {mode}: forward rel err {err:.2e}")
        assert err < 1e-12


@pytest.mark.parametrize("N,R", [(3, 8), (4, 16)])
def test_backward_matches_dense_gradient(ctx, orc, N, R):
    import torch
    from rdist_tensor_amd.autograd import SampledMTTKRP
    rng, dims, idx, vals, T, U, _ = setup_case(ctx, orc, N, R, 100 + N + R)
    for mode in range(N):
        sd, wd = make_samples(orc, rng, idx, dims, mode, 2000, 500)
        C = rng.standard_normal((dims[mode], R))
        Ug = [torch.tensor(u, device="cuda", requires_grad=True) for u in U]
        Y = SampledMTTKRP.apply(T, mode, torch.tensor(sd.astype(np.int32), device="cuda"),
                                torch.tensor(wd, device="cuda"), *Ug)
        (Y * torch.tensor(C, device="cuda")).sum().backward()
        Uc = [torch.tensor(u, requires_grad=True) for u in U]
        Yd = dense_forward(torch, Uc, sd, wd, mode, hits_of(idx, vals, sd, mode), dims[mode])
        (Yd * torch.tensor(C)).sum().backward()
        assert Ug[mode].grad is None or float(Ug[mode].grad.abs().max()) == 0.0
        for k in range(N):
            if k == mode:
                continue
            err = rel_err(Ug[k].grad.cpu().numpy(), Uc[k].grad.numpy())
            print(f"N={N} R={R} mode {mode}: dL/dU_{k} rel err {err:.2e}")
            assert err < 1e-12


def test_stream_ordered_entries_match_synchronous(ctx, orc):
    """ex_smttkrp_forward_async / _backward_async on a side torch stream (the
    autograd path, nothing waited for) give the synchronous entries' Y bitwise
    and their gradients within 1e-12 (atomic adds in the adjoint), over three
    modes in a row on one tensor (the temporaries it keeps are reused)."""
    import torch
    from rdist_tensor_amd import abi
    from rdist_tensor_amd.autograd import _ptr_array
    N, R = 3, 16
    rng, dims, idx, vals, T, U, Ut = setup_case(ctx, orc, N, R, 77)
    side = torch.cuda.Stream()
    for mode in range(N):
        sd, wd = make_samples(orc, rng, idx, dims, mode, 3000, 1000)
        sdt = torch.tensor(sd.astype(np.int32), device="cuda")
        wdt = torch.tensor(wd, device="cuda")
        G = torch.tensor(rng.standard_normal((dims[mode], R)), device="cuda")
        Ys = torch.empty((dims[mode], R), dtype=torch.float64, device="cuda")
        gs = [torch.zeros_like(u) if k != mode else None for k, u in enumerate(Ut)]
        torch.cuda.synchronize()
        abi.check(abi.lib().ex_smttkrp_forward(T.h, mode, len(sd), sdt.data_ptr(), wdt.data_ptr(), _ptr_array(Ut),
                                               R, Ys.data_ptr()))
        abi.check(abi.lib().ex_smttkrp_backward(T.h, mode, len(sd), sdt.data_ptr(), wdt.data_ptr(), _ptr_array(Ut),
                                                R, G.data_ptr(), _ptr_array(gs)))
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            Ya = torch.empty((dims[mode], R), dtype=torch.float64, device="cuda")
            ga = [torch.zeros_like(u) if k != mode else None for k, u in enumerate(Ut)]
            abi.check(abi.lib().ex_smttkrp_forward_async(T.h, side.cuda_stream, mode, len(sd), sdt.data_ptr(),
                                                         wdt.data_ptr(), _ptr_array(Ut), R, Ya.data_ptr()))
            abi.check(abi.lib().ex_smttkrp_backward_async(T.h, side.cuda_stream, mode, len(sd), sdt.data_ptr(),
                                                          wdt.data_ptr(), _ptr_array(Ut), R, G.data_ptr(),
                                                          _ptr_array(ga)))
            Yh = Ya.cpu()
            gh = [g.cpu() if g is not None else None for g in ga]
        assert torch.equal(Yh, Ys.cpu())
        for k in range(N):
            if k != mode:
                err = rel_err(gh[k].numpy(), gs[k].cpu().numpy())
                print(f"mode {mode}: async dL/dU_{k} rel err {err:.2e}")
                assert err < 1e-12


def test_duplicate_samples_rejected(ctx, orc):
    """Samples sharing every non-target coordinate: the synchronous entries
    return EX_ERR_ARG (outputs NaN), the stream-ordered ones fill their
    outputs with NaN; a deduplicated call on the same tensor afterwards is
    exact again."""
    import torch
    from rdist_tensor_amd import abi
    from rdist_tensor_amd.autograd import _ptr_array
    N, R = 3, 8
    rng, dims, idx, vals, T, U, Ut = setup_case(ctx, orc, N, R, 78)
    mode = 1
    sd, wd = make_samples(orc, rng, idx, dims, mode, 500, 100)
    dup = np.concatenate([sd, sd[:1]]).copy()
    dup[-1, mode] = (dup[-1, mode] + 1) % dims[mode]   # the target column differs, the fiber is the same
    wdup = np.concatenate([wd, wd[:1]])
    sdt = torch.tensor(dup.astype(np.int32), device="cuda")
    wdt = torch.tensor(wdup, device="cuda")
    Y = torch.zeros((dims[mode], R), dtype=torch.float64, device="cuda")
    rc = abi.lib().ex_smttkrp_forward(T.h, mode, len(dup), sdt.data_ptr(), wdt.data_ptr(), _ptr_array(Ut), R,
                                      Y.data_ptr())
    assert rc == abi.EX_ERR_ARG and "deduplicated" in abi.lib().ex_last_error().decode()
    assert bool(torch.isnan(Y).all())
    G = torch.ones((dims[mode], R), dtype=torch.float64, device="cuda")
    ga = [torch.zeros_like(u) if k != mode else None for k, u in enumerate(Ut)]
    abi.check(abi.lib().ex_smttkrp_backward_async(T.h, None, mode, len(dup), sdt.data_ptr(), wdt.data_ptr(),
                                                  _ptr_array(Ut), R, G.data_ptr(), _ptr_array(ga)))
    torch.cuda.synchronize()
    assert all(bool(torch.isnan(g).all()) for g in ga if g is not None)
    # deduplicated again: matches the oracle
    D, _ = orc.design(sd, wd, mode, U)
    Yo, _ = orc.Lookup(idx, vals, mode, dims).sampled_mttkrp(sd, D)
    s2 = torch.tensor(sd.astype(np.int32), device="cuda")
    w2 = torch.tensor(wd, device="cuda")
    abi.check(abi.lib().ex_smttkrp_forward(T.h, mode, len(sd), s2.data_ptr(), w2.data_ptr(), _ptr_array(Ut), R,
                                           Y.data_ptr()))
    assert rel_err(Y.cpu().numpy(), Yo) < 1e-12


def test_backward_central_differences(ctx, orc):
    import torch
    from rdist_tensor_amd.autograd import SampledMTTKRP
    N, R = 3, 8
    rng, dims, idx, vals, T, U, _ = setup_case(ctx, orc, N, R, 120)
    mode = 1
    sd, wd = make_samples(orc, rng, idx, dims, mode, 2000, 500)
    sdt = torch.tensor(sd.astype(np.int32), device="cuda")
    wdt = torch.tensor(wd, device="cuda")
    C = torch.tensor(rng.standard_normal((dims[mode], R)), device="cuda")
    Ug = [torch.tensor(u, device="cuda", requires_grad=True) for u in U]
    (SampledMTTKRP.apply(T, mode, sdt, wdt, *Ug) * C).sum().backward()

    def L(Us):
        with torch.no_grad():
            return float((SampledMTTKRP.apply(T, mode, sdt, wdt, *Us) * C).sum())

    eps = 1e-3
    worst = 0.0
    # entries in rows that the samples touch (elsewhere the gradient is 0)
    for t in range(64):
        k = [0, 2][t % 2]
        i = int(sd[rng.integers(0, len(sd)), k])
        r = int(rng.integers(0, R))
        Up = [u.detach().clone() for u in Ug]
        Um = [u.detach().clone() for u in Ug]
        Up[k][i, r] += eps
        Um[k][i, r] -= eps
        fd = (L(Up) - L(Um)) / (2 * eps)
        g = float(Ug[k].grad[i, r])
        worst = max(worst, abs(fd - g) / max(1.0, abs(g)))
    print(f"central differences: worst |fd - grad| / max(1, |grad|) = {worst:.2e}")
    assert worst < 1e-8


def test_adam_fits_sampled_objective(ctx, orc):
    """Planted rank-6 tensor (unit-column factors) observed on 40% of its
    entries; leverage samples
    drawn once per mode from the STS-CP sampler, then 200 Adam steps on the
    sum of the modes' sampled least-squares objectives."""
    import torch
    from rdist_tensor_amd import abi
    from rdist_tensor_amd.autograd import sampled_lsq_loss
    rng = np.random.default_rng(130)
    dims = [60, 50, 40]
    R = 6
    A = normalized_factors(rng, dims, R)   # the same scale as the unit-column starting factors
    grid = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1).reshape(-1, 3)
    keep = rng.random(len(grid)) < 0.4
    grid = grid[keep].astype(np.uint32)
    vals = np.einsum("ir,ir,ir->i", A[0][grid[:, 0]], A[1][grid[:, 1]], A[2][grid[:, 2]])
    T = abi.Tensor(ctx, grid, vals, dims)
    T.build_lookups()
    U0 = normalized_factors(rng, dims, R)
    F = abi.Factors(ctx, dims, R)
    for k in range(3):
        F.set(k, U0[k])
    S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP, seed=13)
    samples = []
    const = 0.0   # 1/2 sum_c w_c |X[:, s_c]|^2 over the modes: objective + const = sampled residual
    for mode in range(3):
        s, w = S.draw(8192, mode)
        sd, wd = orc.dedup(s, w, mode, dims)
        _, cs, xs = hits_of(grid, vals, sd, mode)
        const += 0.5 * float(np.sum(wd[cs] * xs * xs))
        samples.append((torch.tensor(sd.astype(np.int32), device="cuda"), torch.tensor(wd, device="cuda")))
    Ut = [torch.tensor(u, device="cuda", requires_grad=True) for u in U0]
    opt = torch.optim.Adam(Ut, lr=1e-2)
    resid = []
    for it in range(200):
        opt.zero_grad()
        loss = sum(sampled_lsq_loss(T, m, samples[m][0], samples[m][1], Ut) for m in range(3))
        loss.backward()
        opt.step()
        resid.append(loss.detach().item() + const)
    print(f"sampled residual: {resid[0]:.4e} -> {resid[-1]:.4e} (const {const:.4e}); "
          f"steps 0/50/100/150/199: {[round(resid[i] / resid[0], 4) for i in (0, 50, 100, 150, 199)]}")
    assert np.all(np.isfinite(resid))
    assert min(resid) > -1e-6 * const   # a residual norm
    assert resid[-1] < 0.5 * resid[0]


def test_full_size_fd_and_adam(ctx):
    """The full differentiable-MTTKRP workload: (5e5)^3 tensor, 2e7 uniform
    nonzeros, U(0,1) values, R = 32, J = 65,536 leverage samples per mode
    (STS-CP on the starting factors, fixed), forward plus adjoint for all three
    factors, central differences (eps = 1e-3, FP64) on 64 random entries of
    sampled rows, then 200 Adam steps (lr 1e-2) on the sampled objective."""
    import time
    import torch
    from rdist_tensor_amd import abi, synthetic
    from rdist_tensor_amd.autograd import SampledMTTKRP, sampled_lsq_loss
    dims = [500_000] * 3
    R, J = 32, 65536
    t0 = time.time()
    idx = synthetic.coords(dims, 20_000_000, 3)
    vals = synthetic.values(len(idx), 3, "uniform_f32")
    U0 = synthetic.factors(dims, R, 3)
    print(f"generated in {time.time() - t0:.1f} s", flush=True)
    T = abi.Tensor(ctx, idx, vals, dims)
    T.build_lookups()
    F = abi.Factors(ctx, dims, R)
    for k in range(3):
        F.set(k, U0[k])
    S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP, seed=3)
    samples = []
    for mode in range(3):
        s, w = S.draw(J, mode)
        sd, wd = ctx.dedup(s, w, mode, dims)
        samples.append((torch.tensor(sd.astype(np.int32), device="cuda"), torch.tensor(wd, device="cuda")))
    Ut = [torch.tensor(u, device="cuda", requires_grad=True) for u in U0]
    rng = np.random.default_rng(3)
    # forward + adjoint for every mode, timed, and central differences
    worst = 0.0
    nonzero_grads = 0
    for mode in range(3):
        sdt, wdt = samples[mode]
        C = torch.tensor(rng.standard_normal((dims[mode], R)), device="cuda")
        for u in Ut:
            u.grad = None
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        Y = SampledMTTKRP.apply(T, mode, sdt, wdt, *Ut)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        (Y * C).sum().backward()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        print(f"mode {mode}: J_u {sdt.shape[0]}, forward {1e3 * (t2 - t1):.2f} ms, "
              f"adjoint {1e3 * (t3 - t2):.2f} ms", flush=True)

        def L(Us):
            with torch.no_grad():
                return float((SampledMTTKRP.apply(T, mode, sdt, wdt, *Us) * C).sum())
        # FD entries in rows of samples whose fibers hold nonzeros (elsewhere
        # the gradient is zero); uniform leverage samples of this tensor
        # reach few nonzero fibers
        sd_host = sdt.cpu().numpy().astype(np.int64)
        others = [k for k in range(3) if k != mode]
        skey = sd_host[:, others[0]] * dims[others[1]] + sd_host[:, others[1]]
        nkey = idx[:, others[0]].astype(np.int64) * dims[others[1]] + idx[:, others[1]]
        hit = np.flatnonzero(np.isin(skey, nkey))
        print(f"mode {mode}: samples with nonzero fibers {len(hit)}", flush=True)
        assert len(hit) > 0
        for t in range(64 // 3 + 1):
            k = others[t % 2]
            i = int(sd_host[hit[rng.integers(0, len(hit))], k])
            r = int(rng.integers(0, R))
            Up = [u.detach().clone() for u in Ut]
            Um = [u.detach().clone() for u in Ut]
            Up[k][i, r] += 1e-3
            Um[k][i, r] -= 1e-3
            fd = (L(Up) - L(Um)) / 2e-3
            g = float(Ut[k].grad[i, r])
            nonzero_grads += g != 0.0
            worst = max(worst, abs(fd - g) / max(1.0, abs(g)))
    print(f"central differences: worst {worst:.2e}", flush=True)
    assert worst < 1e-8 and nonzero_grads > 0
    opt = torch.optim.Adam(Ut, lr=1e-2)
    losses = []
    t1 = time.perf_counter()
    for it in range(200):
        opt.zero_grad()
        loss = sum(sampled_lsq_loss(T, m, samples[m][0], samples[m][1], Ut) for m in range(3))
        loss.backward()
        opt.step()
        losses.append(loss.detach().item())
    el = time.perf_counter() - t1
    print(f"200 Adam steps in {el:.1f} s: objective {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < losses[0]
