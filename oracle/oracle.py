"""ctypes wrapper for the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product path
(rdist_tensor_amd/) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = C.POINTER(C.c_double)
_u32 = C.POINTER(C.c_uint32)
_i64 = C.POINTER(C.c_int64)


def build(force=False):
    so = os.path.join(_HERE, "liborc.so")
    src = os.path.join(_HERE, "exafac_oracle.c")
    if force or not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return so


class Stats(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "t_sampling", "t_gather", "t_design", "t_spmm", "t_reduce", "t_post", "t_update",
        "als_time", "fit_time", "nonzeros_iterated")]


class PTree(C.Structure):
    _fields_ = [("n", C.c_int64), ("F", C.c_int64), ("R", C.c_int64),
                ("leaf_count", C.c_uint32), ("node_count", C.c_uint32),
                ("lfill_level", C.c_uint32), ("lfill_count", C.c_uint32),
                ("nodes_upto_lfill", C.c_uint32), ("nodes_before_lfill", C.c_uint32),
                ("complete_level_offset", C.c_uint32),
                ("G", _d), ("G_unmult", _d)]


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        sig = {
            "orc_philox4x32": (None, [_u32, _u32, _u32]),
            "orc_uniform": (C.c_double, [C.c_uint64, C.c_uint64, C.c_uint64]),
            "orc_uniform_fill": (None, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int64, _d]),
            "orc_stream_id": (C.c_uint64, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
            "orc_gram": (None, [_d, C.c_int64, C.c_int64, _d]),
            "orc_eigh": (C.c_int, [_d, C.c_int64, _d, _d]),
            "orc_pinv": (None, [_d, C.c_int64, C.c_double, _d]),
            "orc_dagat": (None, [_d, C.c_int64, C.c_int64, _d, _d]),
            "orc_dsymm": (None, [_d, _d, C.c_int64, C.c_int64, _d]),
            "orc_renormalize": (None, [_d, C.c_int64, C.c_int64, _d]),
            "orc_ptree_init": (None, [C.POINTER(PTree), C.c_int64, C.c_int64, C.c_int64]),
            "orc_ptree_free": (None, [C.POINTER(PTree)]),
            "orc_ptree_build": (None, [C.POINTER(PTree), _d]),
            "orc_ptree_multiply": (None, [C.POINTER(PTree), _d]),
            "orc_ptree_leaf_idx": (C.c_int64, [C.POINTER(PTree), C.c_int64]),
            "orc_ptree_sample": (None, [C.POINTER(PTree), _d, C.c_int64, _d, _d, _d, _u32,
                                        C.c_int64, C.c_int64, _d, _d]),
            "orc_arls_scores": (None, [_d, C.c_int64, C.c_int64, _d]),
            "orc_arls_draw": (None, [C.c_int, _i64, C.POINTER(_d), C.c_int, C.c_int64, _d, _u32, _d, _d]),
            "orc_sts_create": (C.c_void_p, [C.c_int, _i64, C.c_int64, C.POINTER(_d)]),
            "orc_sts_free": (None, [C.c_void_p]),
            "orc_sts_update": (None, [C.c_void_p, C.c_int, _d]),
            "orc_sts_compute_m": (None, [C.c_void_p, C.c_int]),
            "orc_sts_slot": (_d, [C.c_void_p, C.c_int]),
            "orc_sts_gram": (_d, [C.c_void_p, C.c_int]),
            "orc_sts_tree": (C.POINTER(PTree), [C.c_void_p, C.c_int, C.c_int]),
            "orc_sts_draw": (None, [C.c_void_p, C.POINTER(_d), C.c_int, C.c_int64, _d, _d, _u32, _d, _d, _d]),
            "orc_dedup": (C.c_int64, [_u32, _d, C.c_int64, C.c_int, C.c_int, _i64, _u32, _d]),
            "orc_design": (None, [_u32, _d, C.c_int64, C.c_int, C.c_int, C.POINTER(_d), C.c_int64, _d, _d]),
            "orc_lookup_create": (C.c_void_p, [C.c_int, C.c_int, _i64, C.c_int64, _u32, _d]),
            "orc_lookup_free": (None, [C.c_void_p]),
            "orc_sampled_mttkrp": (C.c_int64, [C.c_void_p, _u32, C.c_int64, _d, C.c_int64, _d, C.c_int64]),
            "orc_exact_mttkrp": (None, [C.c_void_p, C.POINTER(_d), C.c_int64, _d, C.c_int64]),
            "orc_fit": (C.c_double, [C.c_void_p, C.c_int, _i64, C.POINTER(_d), _d, C.c_int64, C.c_double]),
            "orc_als_create": (C.c_void_p, [C.c_int, _i64, C.c_int64, _u32, _d, C.c_int64, C.POINTER(_d),
                                            _d, C.c_int, C.c_uint64]),
            "orc_als_free": (None, [C.c_void_p]),
            "orc_als_fit": (C.c_double, [C.c_void_p]),
            "orc_als_step": (None, [C.c_void_p, C.c_int, C.c_int64, C.c_uint64]),
            "orc_als_run": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_int64, _d, C.POINTER(Stats)]),
            "orc_num_threads": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ptr_array(arrs):
    return (_d * len(arrs))(*[_p(a) for a in arrs])


# ---------------------------------------------------------------- draws
def uniform(seed, stream, first, count):
    out = np.empty(count, np.float64)
    lib().orc_uniform_fill(seed, stream, first, count, _p(out))
    return out


def philox4x32(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.empty(4, np.uint32)
    lib().orc_philox4x32(_p(c, _u32), _p(k, _u32), _p(out, _u32))
    return out


def stream_id(sweep, target_mode, draw_mode, kind):
    return lib().orc_stream_id(sweep, target_mode, draw_mode, kind)


# ---------------------------------------------------------------- dense
def gram(A):
    A = f64(A)
    m, R = A.shape
    G = np.empty((R, R))
    lib().orc_gram(_p(A), m, R, _p(G))
    return G


def eigh(A):
    A = f64(A)
    R = A.shape[0]
    lam = np.empty(R)
    V = np.empty((R, R))
    lib().orc_eigh(_p(A), R, _p(lam), _p(V))
    return lam, V


def pinv(M, tol=1e-11):
    M = f64(M)
    R = M.shape[0]
    out = np.empty((R, R))
    lib().orc_pinv(_p(M), R, tol, _p(out))
    return out


def dagat(A, G):
    A = f64(A)
    G = f64(G)
    m, R = A.shape
    res = np.empty(m)
    lib().orc_dagat(_p(A), m, R, _p(G), _p(res))
    return res


def dsymm(G, Y):
    G = f64(G)
    Y = f64(Y)
    m, R = Y.shape
    out = np.empty((m, R))
    lib().orc_dsymm(_p(G), _p(Y), m, R, _p(out))
    return out


def renormalize(U):
    U = f64(U).copy()
    m, R = U.shape
    sigma = np.empty(R)
    lib().orc_renormalize(_p(U), m, R, _p(sigma))
    return U, sigma


# ---------------------------------------------------------------- trees
class Tree:
    """A PartitionTree(n, F, R) restated in the oracle."""

    def __init__(self, n, F, R):
        self.t = PTree()
        lib().orc_ptree_init(C.byref(self.t), n, F, R)

    def __del__(self):
        try:
            lib().orc_ptree_free(C.byref(self.t))
        except Exception:
            pass

    def build(self, U):
        U = f64(U)
        lib().orc_ptree_build(C.byref(self.t), _p(U))

    def multiply(self, mat):
        mat = f64(mat)
        lib().orc_ptree_multiply(C.byref(self.t), _p(mat))

    def leaf_idx(self, c):
        return lib().orc_ptree_leaf_idx(C.byref(self.t), c)

    @property
    def G(self):
        t = self.t
        n = t.node_count * t.R * t.R
        return np.ctypeslib.as_array(t.G, shape=(n,)).reshape(t.node_count, t.R, t.R).copy()

    def sample(self, U, x, draws, h=None):
        U = f64(U)
        x = f64(x)
        J, R = x.shape
        h = np.ones((J, R)) if h is None else f64(h).copy()
        draws = f64(draws)
        samples = np.zeros((J, 1), np.uint32)
        margin = np.full(J, np.inf)
        slack = np.full(J, np.inf)
        lib().orc_ptree_sample(C.byref(self.t), _p(U), J, _p(x), _p(h), _p(draws),
                               _p(samples, _u32), 1, 0, _p(margin), _p(slack))
        self.slack = slack
        return samples[:, 0].copy(), h, margin


# ---------------------------------------------------------------- samplers
def arls_scores(U):
    U = f64(U)
    I, R = U.shape
    s = np.empty(I)
    lib().orc_arls_scores(_p(U), I, R, _p(s))
    return s


def arls_draw(scores, j, J, draws):
    """draws: (N, J) array, row k used for mode k != j."""
    N = len(scores)
    dims = np.array([len(s) for s in scores], np.int64)
    sc = [f64(s) for s in scores]
    draws = f64(draws)
    samples = np.empty((J, N), np.uint32)
    w = np.empty(J)
    margin = np.empty(J)
    lib().orc_arls_draw(N, _p(dims, _i64), _ptr_array(sc), j, J, _p(draws), _p(samples, _u32),
                        _p(w), _p(margin))
    return samples, w, margin


class STS:
    """STS-CP sampler state (sts_cp.hpp) restated in the oracle."""

    def __init__(self, U):
        self.U = [f64(u) for u in U]
        self.N = len(U)
        self.R = self.U[0].shape[1]
        self.dims = np.array([u.shape[0] for u in self.U], np.int64)
        self.h = lib().orc_sts_create(self.N, _p(self.dims, _i64), self.R, _ptr_array(self.U))

    def __del__(self):
        try:
            lib().orc_sts_free(self.h)
        except Exception:
            pass

    def update(self, j, Uj):
        self.U[j] = f64(Uj)
        lib().orc_sts_update(self.h, j, _p(self.U[j]))

    def compute_m(self, j):
        lib().orc_sts_compute_m(self.h, j)

    def slot(self, k):
        R = self.R
        return np.ctypeslib.as_array(lib().orc_sts_slot(self.h, k), shape=(R * R,)).reshape(R, R).copy()

    def gram(self, k):
        R = self.R
        return np.ctypeslib.as_array(lib().orc_sts_gram(self.h, k), shape=(R * R,)).reshape(R, R).copy()

    def tree_G(self, k, eigen):
        t = lib().orc_sts_tree(self.h, k, 1 if eigen else 0).contents
        n = t.node_count * t.R * t.R
        return np.ctypeslib.as_array(t.G, shape=(n,)).reshape(t.node_count, t.R, t.R).copy()

    def draw(self, j, J, draws_eig, draws_row):
        N = self.N
        de = f64(draws_eig)
        dr = f64(draws_row)
        samples = np.empty((J, N), np.uint32)
        w = np.empty(J)
        margin = np.empty(J)
        slack = np.empty(J)
        lib().orc_sts_draw(self.h, _ptr_array(self.U), j, J, _p(de), _p(dr), _p(samples, _u32),
                           _p(w), _p(margin), _p(slack))
        # per sample: the smallest distance to a compared boundary minus that
        # boundary's rounding bound (orc_ptree_sample); <= 0: the bin depends
        # on the summation order
        self.slack = slack
        return samples, w, margin


# ---------------------------------------------------------------- ALS pieces
def dedup(samples, weights, j, dims):
    samples = u32(samples)
    weights = f64(weights)
    J, N = samples.shape
    dims = np.asarray(dims, np.int64)
    sd = np.empty((J, N), np.uint32)
    wd = np.empty(J)
    n = lib().orc_dedup(_p(samples, _u32), _p(weights), J, N, j, _p(dims, _i64), _p(sd, _u32), _p(wd))
    if n < 0:
        raise ValueError("coordinate key does not fit 64 bits")
    return sd[:n].copy(), wd[:n].copy()


def design(sd, wd, j, U):
    sd = u32(sd)
    wd = f64(wd)
    Ju, N = sd.shape
    Us = [f64(u) for u in U]
    R = Us[0].shape[1]
    D = np.empty((max(Ju, 1), R))
    Gi = np.empty((R, R))
    lib().orc_design(_p(sd, _u32), _p(wd), Ju, N, j, _ptr_array(Us), R, _p(D), _p(Gi))
    return D[:Ju].copy(), Gi


class Lookup:
    def __init__(self, idx, vals, j, dims):
        self.idx = u32(idx)
        self.vals = f64(vals)
        nnz, N = self.idx.shape
        self.N = N
        self.j = j
        self.dims = np.asarray(dims, np.int64)
        self.h = lib().orc_lookup_create(N, j, _p(self.dims, _i64), nnz, _p(self.idx, _u32), _p(self.vals))
        if not self.h:
            raise ValueError("coordinate key does not fit 64 bits")

    def __del__(self):
        try:
            lib().orc_lookup_free(self.h)
        except Exception:
            pass

    def sampled_mttkrp(self, sd, design):
        sd = u32(sd)
        design = f64(design)
        Ju = sd.shape[0]
        R = design.shape[1]
        Ij = int(self.dims[self.j])
        Y = np.zeros((Ij, R))
        H = lib().orc_sampled_mttkrp(self.h, _p(sd, _u32), Ju, _p(design), R, _p(Y), Ij)
        return Y, H

    def exact_mttkrp(self, U):
        Us = [f64(u) for u in U]
        R = Us[0].shape[1]
        Ij = int(self.dims[self.j])
        Y = np.zeros((Ij, R))
        lib().orc_exact_mttkrp(self.h, _ptr_array(Us), R, _p(Y), Ij)
        return Y

    def fit(self, U, sigma, tensor_norm):
        Us = [f64(u) for u in U]
        sigma = f64(sigma)
        R = Us[0].shape[1]
        return lib().orc_fit(self.h, self.N, _p(self.dims, _i64), _ptr_array(Us), _p(sigma), R, tensor_norm)


class ALS:
    """Whole CP-ALS (accumulator-stationary, one rank) in the oracle: sampler
    "arls" (CP-ARLS-LEV), "sts" (STS-CP) or "exact" (ExactALS, no sampling;
    J is ignored).  Factors are updated in place in self.U / self.sigma."""

    def __init__(self, idx, vals, dims, U, sampler="sts", seed=0):
        self.idx = u32(idx)
        self.vals = f64(vals)
        self.dims = np.asarray(dims, np.int64)
        self.U = [f64(u).copy() for u in U]
        self.R = self.U[0].shape[1]
        self.sigma = np.ones(self.R)
        self._uptr = _ptr_array(self.U)
        kind = {"arls": 0, "sts": 1, "exact": 2}[sampler]
        nnz, N = self.idx.shape
        self.N = N
        self.h = lib().orc_als_create(N, _p(self.dims, _i64), nnz, _p(self.idx, _u32), _p(self.vals),
                                      self.R, self._uptr, _p(self.sigma), kind, seed)
        if not self.h:
            raise ValueError("oracle ALS creation failed")

    def __del__(self):
        try:
            lib().orc_als_free(self.h)
        except Exception:
            pass

    def fit(self):
        return lib().orc_als_fit(self.h)

    def step(self, j, J, sweep):
        lib().orc_als_step(self.h, j, J, sweep)

    def run(self, rounds, J, epoch=1):
        fits = np.zeros(max(rounds // max(epoch, 1), 1))
        st = Stats()
        n = lib().orc_als_run(self.h, rounds, J, epoch, _p(fits), C.byref(st))
        return fits[:n], {f: getattr(st, f) for f, _ in Stats._fields_}
