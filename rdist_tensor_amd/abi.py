"""ctypes binding of the C-ABI in include/exafac_hip.h (libexafac_hip.so).

This is the binding a Python caller (or a test) uses to reach the HIP library
without pybind11.  Raises RuntimeError if the library is missing: there is no
CPU fallback.
"""
import ctypes as C
import os
import sys
import weakref

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libexafac_hip.so")

EX_OK, EX_ERR_ARG, EX_ERR_HIP, EX_ERR_STATE, EX_ERR_UNSUPPORTED, EX_ERR_NOMEM = range(6)
EX_FP64, EX_MIXED = 0, 1
EX_SAMPLER_CP_ARLS_LEV, EX_SAMPLER_STS_CP = 0, 1
EX_DIST_ACCUMULATOR_STATIONARY, EX_DIST_TENSOR_STATIONARY = 0, 1
EX_PRE_NONE, EX_PRE_LOG_COUNT, EX_PRE_ONES = 0, 1, 2
EX_MTTKRP_AUTO, EX_MTTKRP_ROW_PASS, EX_MTTKRP_HIT_LIST, EX_MTTKRP_BUCKETS = 0, 1, 2, 3

_vp = C.c_void_p
_d = C.POINTER(C.c_double)
_u32 = C.POINTER(C.c_uint32)
_i64 = C.POINTER(C.c_int64)
_pp = C.POINTER(C.c_void_p)

# name -> argtypes (all return int status, except ex_last_error)
SIGNATURES = {
    "ex_ctx_create": [C.c_int, _pp],
    "ex_ctx_destroy": [_vp],
    "ex_ctx_sync": [_vp],
    "ex_ctx_device_name": [_vp, C.c_char_p, C.c_size_t],
    "ex_ctx_timing": [_vp, C.c_int],
    "ex_ctx_timing_json": [_vp, C.c_char_p, C.POINTER(C.c_size_t)],
    "ex_tensor_create": [_vp, C.c_int, _i64, C.c_int64, _u32, _d, C.c_int, _pp],
    "ex_tensor_destroy": [_vp],
    "ex_tensor_norm": [_vp, _d],
    "ex_tensor_build_lookups": [_vp],
    "ex_sampled_mttkrp": [_vp, C.c_int, C.c_int64, _u32, _d, C.c_int64, _d, _i64],
    "ex_factors_create": [_vp, C.c_int, _i64, C.c_int64, _pp],
    "ex_factors_destroy": [_vp],
    "ex_factors_set": [_vp, C.c_int, _d],
    "ex_factors_get": [_vp, C.c_int, _d],
    "ex_factors_get_sigma": [_vp, _d],
    "ex_factors_set_sigma": [_vp, _d],
    "ex_factors_renormalize": [_vp, C.c_int, _d],
    "ex_factors_init_deterministic": [_vp],
    "ex_factors_gram": [_vp, C.c_int, _d],
    "ex_sampler_create": [_vp, _vp, C.c_int, C.c_int, C.c_uint64, _pp],
    "ex_sampler_destroy": [_vp],
    "ex_sampler_update": [_vp, C.c_int],
    "ex_sampler_draw": [_vp, C.c_int64, C.c_int, C.c_uint64, _d, _d, _u32, _d],
    "ex_sts_compute_m": [_vp, C.c_int, _d],
    "ex_sampler_dropped_draws": [_vp, _i64],
    "ex_arls_scores": [_vp, C.c_int, _d],
    "ex_als_create": [_vp, _vp, _vp, _vp, C.c_int, C.c_int, _pp],
    "ex_exact_als_create": [_vp, _vp, _vp, C.c_int, _pp],
    "ex_tensor_mttkrp_path": [_vp, C.c_int],
    "ex_als_run_batch": [_pp, C.c_int, C.c_int64, C.c_int64, _d],
    "ex_als_use_graph": [_vp, C.c_int],
    "ex_batch_create": [_pp, C.c_int, C.c_int64, _pp],
    "ex_batch_run": [_vp, C.c_int64, _d],
    "ex_batch_use_graph": [_vp, C.c_int],
    "ex_batch_counters": [_vp, _i64, _i64, _i64],
    "ex_batch_destroy": [_vp],
    "ex_tuning_count": [C.POINTER(C.c_int)],
    "ex_tuning_info": [C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "ex_tuning_set": [C.c_char_p, C.c_int],
    "ex_tuning_load": [C.c_char_p],
    "ex_smttkrp_forward": [_vp, C.c_int, C.c_int64, _vp, _vp, _vp, C.c_int64, _vp],
    "ex_smttkrp_backward": [_vp, C.c_int, C.c_int64, _vp, _vp, _vp, C.c_int64, _vp, _vp],
    "ex_smttkrp_forward_async": [_vp, _vp, C.c_int, C.c_int64, _vp, _vp, _vp, C.c_int64, _vp],
    "ex_smttkrp_backward_async": [_vp, _vp, C.c_int, C.c_int64, _vp, _vp, _vp, C.c_int64, _vp, _vp],
    "ex_als_destroy": [_vp],
    "ex_als_init_ground_truth": [_vp],
    "ex_als_step": [_vp, C.c_int, C.c_int64, C.c_uint64],
    "ex_als_run": [_vp, C.c_int64, C.c_int64, C.c_int64],
    "ex_als_fit": [_vp, _d],
    "ex_als_stats_json": [_vp, C.c_char_p, C.POINTER(C.c_size_t)],
    "ex_gram": [_vp, _d, C.c_int64, C.c_int64, _d],
    "ex_gram_dev_static": [_vp, _vp, C.c_int64, C.c_int64, _vp],
    "ex_eigh": [_vp, _d, C.c_int64, _d, _d],
    "ex_pinv": [_vp, _d, C.c_int64, C.c_double, _d],
    "ex_spd_pinv": [_vp, _d, C.c_int64, C.c_double, _d],
    "ex_dsymm": [_vp, C.c_int, _d, _d, C.c_int64, C.c_int64, _d],
    "ex_tns_scan": [C.c_char_p, _i64, C.POINTER(C.c_int)],
    "ex_tns_read": [C.c_char_p, C.c_int64, C.c_int, C.POINTER(C.c_uint64), _d, C.POINTER(C.c_uint64),
                    C.POINTER(C.c_uint64)],
    "ex_dedup": [_vp, _u32, _d, C.c_int64, C.c_int, C.c_int, _i64, _u32, _d, _i64],
}

_LIB = None


class ExafacError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m rdist_tensor_amd.build` "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.ex_last_error.argtypes = []
        L.ex_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def check(rc):
    if rc != EX_OK:
        raise ExafacError(rc, lib().ex_last_error().decode())


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def tuning():
    """{knob name: (variants, current value)} of the launch variants."""
    n = C.c_int()
    check(lib().ex_tuning_count(C.byref(n)))
    out = {}
    for q in range(n.value):
        buf = C.create_string_buffer(64)
        nv, v = C.c_int(), C.c_int()
        check(lib().ex_tuning_info(q, buf, 64, C.byref(nv), C.byref(v)))
        out[buf.value.decode()] = (nv.value, v.value)
    return out


def tuning_set(name, value):
    check(lib().ex_tuning_set(name.encode(), value))


def tuning_load(path):
    check(lib().ex_tuning_load(os.fsencode(path)))


class _Owned:
    """A library object of one context: destroyed once, by close() or by
    garbage collection, and always before its context (Context.close()
    closes the live objects it created, newest first)."""
    _destroy = None

    def _register(self, ctx):
        self.ctx = ctx
        ctx._children.append(weakref.ref(self))

    def close(self):
        h = getattr(self, "h", None)
        if h:
            getattr(lib(), self._destroy)(h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    def __init__(self, device=0):
        self._children = []
        # PyTorch-ROCm carries its own HIP runtime.  In a process that also
        # uses PyTorch on the GPU, PyTorch's runtime has to start first (the
        # other order leaves PyTorch without devices), so it is started here
        # when the caller has imported torch.
        torch = sys.modules.get("torch")
        if torch is not None and torch.cuda.is_available():
            torch.cuda.init()
        self.h = C.c_void_p()
        check(lib().ex_ctx_create(device, C.byref(self.h)))

    def close(self):
        if self.h:
            for ref in reversed(self._children):
                obj = ref()
                if obj is not None:
                    obj.close()
            self._children = []
            lib().ex_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def device_name(self):
        buf = C.create_string_buffer(256)
        check(lib().ex_ctx_device_name(self.h, buf, 256))
        return buf.value.decode()

    def sync(self):
        check(lib().ex_ctx_sync(self.h))

    def timing(self, enable=True, clear=True):
        """Span timing on (clearing the totals unless clear=False) or off
        (the totals stay readable)."""
        check(lib().ex_ctx_timing(self.h, (1 if clear else 2) if enable else 0))

    def timing_json(self):
        n = C.c_size_t(0)
        check(lib().ex_ctx_timing_json(self.h, None, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(lib().ex_ctx_timing_json(self.h, buf, C.byref(n)))
        return buf.value.decode()

    # dense hooks
    def gram(self, A):
        A = f64(A)
        m, R = A.shape
        G = np.empty((R, R))
        check(lib().ex_gram(self.h, _p(A), m, R, _p(G)))
        return G

    def eigh(self, A):
        A = f64(A)
        R = A.shape[0]
        lam = np.empty(R)
        V = np.empty((R, R))
        check(lib().ex_eigh(self.h, _p(A), R, _p(lam), _p(V)))
        return lam, V

    def pinv(self, M, tol=1e-11):
        M = f64(M)
        R = M.shape[0]
        out = np.empty((R, R))
        check(lib().ex_pinv(self.h, _p(M), R, tol, _p(out)))
        return out

    def spd_pinv(self, M, tol_rel=1e-11):
        """The mixed path's pseudo-inverse of a PSD matrix (Gauss-Jordan, or
        the Moore-Penrose pseudo-inverse when a pivot is dropped)."""
        M = f64(M)
        R = M.shape[0]
        out = np.empty((R, R))
        check(lib().ex_spd_pinv(self.h, _p(M), R, tol_rel, _p(out)))
        return out

    def dsymm(self, G, Y, precision=EX_FP64):
        """out = Y G (the factor update's product) at the given precision."""
        G = f64(G)
        Y = f64(Y)
        m, R = Y.shape
        out = np.empty((m, R))
        check(lib().ex_dsymm(self.h, precision, _p(G), _p(Y), m, R, _p(out)))
        return out

    def dedup(self, samples, weights, mode, dims):
        samples = u32(samples)
        weights = f64(weights)
        J, N = samples.shape
        dims = np.asarray(dims, np.int64)
        sd = np.empty((max(J, 1), N), np.uint32)
        wd = np.empty(max(J, 1))
        ju = C.c_int64()
        check(lib().ex_dedup(self.h, _p(samples, _u32), _p(weights), J, N, mode, _p(dims, _i64),
                             _p(sd, _u32), _p(wd), C.byref(ju)))
        return sd[:ju.value].copy(), wd[:ju.value].copy()


class Tensor(_Owned):
    _destroy = "ex_tensor_destroy"

    def __init__(self, ctx, idx, vals, dims, preprocessing=EX_PRE_NONE):
        self.idx = u32(idx)
        self.vals = f64(vals)
        self.dims = np.asarray(dims, np.int64)
        nnz, N = self.idx.shape
        self.N = N
        self.h = C.c_void_p()
        check(lib().ex_tensor_create(ctx.h, N, _p(self.dims, _i64), nnz, _p(self.idx, _u32), _p(self.vals),
                                     preprocessing, C.byref(self.h)))
        self._register(ctx)

    def norm(self):
        v = C.c_double()
        check(lib().ex_tensor_norm(self.h, C.byref(v)))
        return v.value

    def mttkrp_path(self, path):
        check(lib().ex_tensor_mttkrp_path(self.h, path))

    def build_lookups(self):
        check(lib().ex_tensor_build_lookups(self.h))

    def sampled_mttkrp(self, mode, samples, design):
        samples = u32(samples)
        design = f64(design)
        Ju = samples.shape[0]
        R = design.shape[1]
        Y = np.empty((int(self.dims[mode]), R))
        hits = C.c_int64()
        check(lib().ex_sampled_mttkrp(self.h, mode, Ju, _p(samples, _u32), _p(design), R, _p(Y), C.byref(hits)))
        return Y, hits.value


class Factors(_Owned):
    _destroy = "ex_factors_destroy"

    def __init__(self, ctx, dims, R):
        self.dims = np.asarray(dims, np.int64)
        self.N = len(dims)
        self.R = R
        self.h = C.c_void_p()
        check(lib().ex_factors_create(ctx.h, self.N, _p(self.dims, _i64), R, C.byref(self.h)))
        self._register(ctx)

    def set(self, mode, U):
        U = f64(U)
        assert U.shape == (self.dims[mode], self.R)
        check(lib().ex_factors_set(self.h, mode, _p(U)))

    def get(self, mode):
        U = np.empty((int(self.dims[mode]), self.R))
        check(lib().ex_factors_get(self.h, mode, _p(U)))
        return U

    def sigma(self):
        s = np.empty(self.R)
        check(lib().ex_factors_get_sigma(self.h, _p(s)))
        return s

    def set_sigma(self, s):
        s = f64(s)
        check(lib().ex_factors_set_sigma(self.h, _p(s)))

    def gram(self, mode):
        G = np.empty((self.R, self.R))
        check(lib().ex_factors_gram(self.h, mode, _p(G)))
        return G

    def init_deterministic(self):
        check(lib().ex_factors_init_deterministic(self.h))


class Sampler(_Owned):
    _destroy = "ex_sampler_destroy"

    def __init__(self, ctx, factors, kind, precision=EX_FP64, seed=0):
        self.f = factors
        self.kind = kind
        self.h = C.c_void_p()
        check(lib().ex_sampler_create(ctx.h, factors.h, kind, precision, seed, C.byref(self.h)))
        self._register(ctx)

    def update(self, mode):
        check(lib().ex_sampler_update(self.h, mode))

    def draw(self, J, mode, sweep=1, draws_row=None, draws_eig=None):
        N = self.f.N
        samples = np.empty((J, N), np.uint32)
        w = np.empty(J)
        dr_arr = f64(draws_row) if draws_row is not None else None
        de_arr = f64(draws_eig) if draws_eig is not None else None
        dr = _p(dr_arr) if dr_arr is not None else None
        de = _p(de_arr) if de_arr is not None else None
        check(lib().ex_sampler_draw(self.h, J, mode, sweep, dr, de, _p(samples, _u32), _p(w)))
        return samples, w

    def dropped_draws(self):
        """STS-CP draws given weight 0 since creation (ex_sampler_dropped_draws)."""
        n = C.c_int64()
        check(lib().ex_sampler_dropped_draws(self.h, C.byref(n)))
        return n.value

    def sts_m(self, mode):
        N, R = self.f.N, self.f.R
        M = np.empty((N + 2, R, R))
        check(lib().ex_sts_compute_m(self.h, mode, _p(M)))
        return M

    def arls_scores(self, mode):
        s = np.empty(int(self.f.dims[mode]))
        check(lib().ex_arls_scores(self.h, mode, _p(s)))
        return s


class ALS(_Owned):
    _destroy = "ex_als_destroy"

    def __init__(self, ctx, tensor, factors, sampler, distribution=EX_DIST_ACCUMULATOR_STATIONARY,
                 precision=EX_FP64):
        self.keep = (tensor, factors, sampler)
        self.h = C.c_void_p()
        check(lib().ex_als_create(ctx.h, tensor.h, factors.h, sampler.h, distribution, precision, C.byref(self.h)))
        self._register(ctx)

    @classmethod
    def exact(cls, ctx, tensor, factors, precision=EX_FP64):
        """ExactALS (exact_als.hpp:16-143): J and sweep arguments are ignored."""
        self = cls.__new__(cls)
        self.keep = (tensor, factors)
        self.h = C.c_void_p()
        check(lib().ex_exact_als_create(ctx.h, tensor.h, factors.h, precision, C.byref(self.h)))
        self._register(ctx)
        return self

    def init(self):
        check(lib().ex_als_init_ground_truth(self.h))

    @staticmethod
    def run_batch(optimizers, rounds, J):
        """`rounds` sweeps of every optimizer (each on its own context), the
        contexts' streams running side by side; returns the wall seconds."""
        arr = (C.c_void_p * len(optimizers))(*[o.h.value for o in optimizers])
        sec = C.c_double()
        check(lib().ex_als_run_batch(arr, len(optimizers), rounds, J, C.byref(sec)))
        return sec.value

    def use_graph(self, on=True):
        """Replay captured sweeps in run / run_batch (ex_als_use_graph)."""
        check(lib().ex_als_use_graph(self.h, 1 if on else 0))

    def step(self, mode, J, sweep):
        check(lib().ex_als_step(self.h, mode, J, sweep))

    def run(self, rounds, J, epoch=1):
        check(lib().ex_als_run(self.h, rounds, J, epoch))

    def fit(self):
        v = C.c_double()
        check(lib().ex_als_fit(self.h, C.byref(v)))
        return v.value

    def stats_json(self):
        n = C.c_size_t(0)
        check(lib().ex_als_stats_json(self.h, None, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(lib().ex_als_stats_json(self.h, buf, C.byref(n)))
        return buf.value.decode()


class Batch(_Owned):
    """CP-ARLS-LEV optimizers advanced together by one kernel launch per sweep
    (ex_batch_create): same device and precision, R <= 32, J <= 8192."""
    _destroy = "ex_batch_destroy"

    def __init__(self, optimizers, J):
        self.keep = tuple(optimizers)
        arr = (C.c_void_p * len(optimizers))(*[o.h.value for o in optimizers])
        self.h = C.c_void_p()
        check(lib().ex_batch_create(arr, len(optimizers), J, C.byref(self.h)))
        self._register(optimizers[0].ctx)

    def run(self, rounds):
        """`rounds` sweeps of every optimizer; returns the wall seconds."""
        sec = C.c_double()
        check(lib().ex_batch_run(self.h, rounds, C.byref(sec)))
        return sec.value

    def use_graph(self, on=True):
        check(lib().ex_batch_use_graph(self.h, 1 if on else 0))

    def counters(self):
        """(nonzeros iterated, unique samples, factor rows written) per
        optimizer, summed over the batch's mode steps."""
        n = len(self.keep)
        nz = np.zeros(n, dtype=np.int64)
        ju = np.zeros(n, dtype=np.int64)
        rows = np.zeros(n, dtype=np.int64)
        check(lib().ex_batch_counters(self.h, _p(nz, _i64), _p(ju, _i64), _p(rows, _i64)))
        return nz, ju, rows
