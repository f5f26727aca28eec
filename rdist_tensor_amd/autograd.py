"""Differentiable sampled MTTKRP for first-order CP fitting.

`SampledMTTKRP` is a torch.autograd.Function over the C-ABI entry points
ex_smttkrp_forward_async / ex_smttkrp_backward_async (HIP kernels enqueued
on torch's current stream, nothing waited for; PyTorch supplies the device
memory and the stream).  Duplicate sample tuples are caught on the device
and turn the outputs into NaN.  For a target mode j, deduplicated sample
tuples s_c (c < J_u) with weights w_c and factors U_0 .. U_{N-1}:

    D_c = (prod_{k != j} U_k[s_c,k] * sqrt(w_c)) * sqrt(w_c)        (design row)
    Y[i] = sum over the nonzeros x at (i, s_c) of x * D_c           (I_j x R)

which is the reference's sampled MTTKRP (sort_lookup.hpp:238-375 on the
design of accumulator_stationary.hpp:210-249).  The backward returns
dL/dU_k for every k != j: the output-gradient rows are gathered over each
sampled fiber's nonzeros and scattered back through the Khatri-Rao product,
whose rows are recomputed rather than stored.

`sampled_lsq_loss` builds the sampled least-squares objective of mode j,
    1/2 sum_c w_c |U_j (prod_{k != j} U_k[s_c,k])^T|^2 - <U_j, Y>,
(the sampled residual norm without its constant sum_c w_c |X[:, s_c]|^2 / 2),
whose gradient drives the factors by first-order methods.

Tensors: factors float64 (I_k x R) on the library context's device; samples
int32 (J_u x N) holding coordinates below 2^31 (the bits of the library's
uint32 tuples); weights float64 (J_u).
"""
import ctypes as C

import torch

from . import abi


def _ptr_array(ts):
    return (C.c_void_p * len(ts))(*[(t.data_ptr() if t is not None else None) for t in ts])


def _check_inputs(tensor, mode, samples, weights, factors):
    N = len(tensor.dims)
    if len(factors) != N:
        raise ValueError(f"expected {N} factors, got {len(factors)}")
    dev = samples.device
    if dev.type != "cuda":
        raise ValueError("samples, weights and factors must be on the GPU")
    if samples.dtype != torch.int32 or samples.dim() != 2 or samples.shape[1] != N:
        raise ValueError("samples must be an int32 tensor of shape (J_u, N)")
    if weights.dtype != torch.float64 or weights.shape != (samples.shape[0],):
        raise ValueError("weights must be a float64 tensor of shape (J_u,)")
    R = factors[0].shape[1]
    for k, f in enumerate(factors):
        if f.dtype != torch.float64 or f.device != dev or tuple(f.shape) != (int(tensor.dims[k]), R):
            raise ValueError(f"factor {k} must be float64 of shape ({int(tensor.dims[k])}, {R}) on {dev}")
    if not 0 <= mode < N:
        raise ValueError("mode out of range")
    return R


class SampledMTTKRP(torch.autograd.Function):
    """Y = SampledMTTKRP.apply(tensor, mode, samples, weights, *factors)."""

    @staticmethod
    def forward(ctx, tensor, mode, samples, weights, *factors):
        R = _check_inputs(tensor, mode, samples, weights, factors)
        samples = samples.contiguous()
        weights = weights.contiguous()
        U = [f.detach().contiguous() for f in factors]
        Y = torch.empty((int(tensor.dims[mode]), R), dtype=torch.float64, device=samples.device)
        # enqueued on torch's current stream, after the work that produced
        # the inputs; nothing is waited for (duplicate samples give NaN)
        stream = torch.cuda.current_stream(samples.device).cuda_stream
        abi.check(abi.lib().ex_smttkrp_forward_async(tensor.h, stream, mode, samples.shape[0], samples.data_ptr(),
                                                     weights.data_ptr(), _ptr_array(U), R, Y.data_ptr()))
        ctx.tensor = tensor
        ctx.mode = mode
        ctx.save_for_backward(samples, weights, *U)
        return Y

    @staticmethod
    def backward(ctx, G):
        samples, weights, *U = ctx.saved_tensors
        mode = ctx.mode
        R = U[0].shape[1]
        G = G.contiguous()
        grads = [torch.zeros_like(u) if (k != mode and ctx.needs_input_grad[4 + k]) else None
                 for k, u in enumerate(U)]
        if any(g is not None for g in grads):
            stream = torch.cuda.current_stream(G.device).cuda_stream
            abi.check(abi.lib().ex_smttkrp_backward_async(ctx.tensor.h, stream, mode, samples.shape[0],
                                                          samples.data_ptr(), weights.data_ptr(), _ptr_array(U), R,
                                                          G.data_ptr(), _ptr_array(grads)))
        return (None, None, None, None, *grads)


def krp_rows(mode, samples, factors):
    """prod_{k != mode} U_k[s_c,k] for every sample row c (torch autograd)."""
    K = None
    for k, f in enumerate(factors):
        if k == mode:
            continue
        rows = f[samples[:, k].long()]
        K = rows if K is None else K * rows
    return K


def sampled_lsq_loss(tensor, mode, samples, weights, factors):
    """Sampled least-squares objective of mode `mode` (see the module text)."""
    K = krp_rows(mode, samples, factors)
    Q = (K * weights[:, None]).T @ K                  # sum_c w_c K_c^T K_c
    Uj = factors[mode]
    quad = 0.5 * ((Uj.T @ Uj) * Q).sum()
    Y = SampledMTTKRP.apply(tensor, mode, samples, weights, *factors)
    return quad - (Uj * Y).sum()
