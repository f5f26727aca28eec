"""Tensor files: FROSTT .tns ingest and the reference's tensor-file layout.

The reference converts a FROSTT .tns text file (one nonzero per line: N
1-based coordinates, then the value) into an HDF5 file holding one column per
mode (`MODE_0` .. `MODE_{N-1}`), the values (`VALUES`, float64) and the
per-mode minimum and maximum coordinates (`MIN_MODE_SET`, `MAX_MODE_SET`)
(tensor_io/process_frostt_tensor.cpp:19-172).  Its reader shifts each mode's
coordinates by that minimum, takes dims = max - min + 1 and builds the
SparseTensor with the chosen preprocessing (sparse_tensor_e.py:63-117).

The datasets are written and read as HDF5 through the system libhdf5 (the
`h5` module, the C API over ctypes), so files from the reference's converter
load directly and files written here load in the reference.  Where no
libhdf5 is found the same datasets go to a NumPy .npz container (read with
allow_pickle=False); the reader takes either, by the file's signature.  The
text parser is native (`ex_tns_scan` / `ex_tns_read` in the C-ABI library;
host code only, no GPU call).
"""
import ctypes as C
import os

import numpy as np

from . import abi, h5


def read_tns(path):
    """Parse a .tns file: (coords uint64 (nnz, N) as written, values float64,
    min_idx uint64 (N), max_idx uint64 (N))."""
    bpath = os.fsencode(path)
    nnz = C.c_int64(0)
    n = C.c_int(0)
    abi.check(abi.lib().ex_tns_scan(bpath, C.byref(nnz), C.byref(n)))
    nnz, n = nnz.value, n.value
    if n == 0:
        raise ValueError(f"{path}: no nonzeros")
    coords = np.empty((max(nnz, 1), n), np.uint64)
    vals = np.empty(max(nnz, 1), np.float64)
    mins = np.empty(n, np.uint64)
    maxs = np.empty(n, np.uint64)
    u64 = C.POINTER(C.c_uint64)
    abi.check(abi.lib().ex_tns_read(bpath, nnz, n, coords.ctypes.data_as(u64),
                                    vals.ctypes.data_as(C.POINTER(C.c_double)),
                                    mins.ctypes.data_as(u64), maxs.ctypes.data_as(u64)))
    return coords[:nnz], vals[:nnz], mins, maxs


def _write_datasets(out, data, container):
    if container == "hdf5":
        h5.write(out, data)
    else:
        with open(out, "wb") as f:
            np.savez(f, **data)


def _container(container):
    if container == "auto":
        return "hdf5" if h5.available() else "npz"
    if container not in ("hdf5", "npz"):
        raise ValueError("container must be 'auto', 'hdf5' or 'npz'")
    return container


def tns_to_tensor_file(path, out=None, index_dtype=np.uint32, container="auto"):
    """The converter: .tns -> the tensor-file datasets (MODE_i, VALUES,
    MIN_MODE_SET, MAX_MODE_SET) in `out` (default: path + "_converted.hdf5",
    the reference's name, or "_converted.npz" without libhdf5).  Indices are
    stored as uint32 by default, as the reference's converter stores them
    (process_frostt_tensor.cpp:174-176), or uint64.  Returns the output
    path."""
    coords, vals, mins, maxs = read_tns(path)
    index_dtype = np.dtype(index_dtype)
    if index_dtype not in (np.dtype(np.uint32), np.dtype(np.uint64)):
        raise ValueError("index_dtype must be uint32 or uint64")
    if coords.size and int(maxs.max()) > np.iinfo(index_dtype).max:
        raise ValueError("coordinates do not fit the index type")
    container = _container(container)
    out = out or (os.fspath(path) + ("_converted.hdf5" if container == "hdf5" else "_converted.npz"))
    data = {f"MODE_{k}": coords[:, k].astype(index_dtype) for k in range(coords.shape[1])}
    data["VALUES"] = vals
    data["MIN_MODE_SET"] = mins.astype(index_dtype)
    data["MAX_MODE_SET"] = maxs.astype(index_dtype)
    _write_datasets(out, data, container)
    return out


def load_tensor_file(path):
    """Read a tensor file (HDF5 or .npz datasets above, or a .tns directly): returns
    (idx uint32 (nnz, N) with each mode shifted by its minimum, values
    float64, dims int64 = max - min + 1)."""
    if os.fspath(path).endswith(".tns"):
        coords, vals, mins, maxs = read_tns(path)
    elif h5.is_hdf5(path):
        f = h5.read(path, ["MAX_MODE_SET", "MIN_MODE_SET"])
        maxs = f["MAX_MODE_SET"].astype(np.uint64)
        mins = f["MIN_MODE_SET"].astype(np.uint64)
        n = len(maxs)
        f = h5.read(path, [f"MODE_{k}" for k in range(n)] + ["VALUES"])
        coords = np.stack([f[f"MODE_{k}"].astype(np.uint64) for k in range(n)], 1)
        vals = f["VALUES"].astype(np.float64)
    else:
        with np.load(path, allow_pickle=False) as f:
            maxs = f["MAX_MODE_SET"].astype(np.uint64)
            mins = f["MIN_MODE_SET"].astype(np.uint64)
            n = len(maxs)
            coords = np.stack([f[f"MODE_{k}"].astype(np.uint64) for k in range(n)], 1)
            vals = f["VALUES"].astype(np.float64)
    dims = (maxs - mins + 1).astype(np.int64)
    if np.any(dims > np.iinfo(np.uint32).max):
        raise ValueError("a mode is larger than 2^32 - 1")
    idx = (coords - mins[None, :]).astype(np.uint32)
    return idx, vals, dims


class DistSparseTensorE:
    """Reader of a tensor file into the drop-in module's SparseTensor, as
    the reference's DistSparseTensorE(filename, grid, preprocessing) does at
    one rank (sparse_tensor_e.py:63-117): attributes tensor_dims, dim, nnz,
    grid, tensor_grid, tensor_idxs, values, sparse_tensor."""

    def __init__(self, filename, grid=None, preprocessing="None"):
        from .cpp_ext import py_module as pm
        self.tensor_idxs, self.values, dims = load_tensor_file(filename)
        self.tensor_dims = np.asarray(dims, dtype=np.int32)
        self.dim = len(self.tensor_dims)
        self.nnz = len(self.values)
        self.grid = grid if grid is not None else pm.Grid(np.ones(self.dim, dtype=np.int32))
        if self.grid.get_dimension() != self.dim:
            raise ValueError("Grid dimension must match tensor dimension")
        self.tensor_grid = pm.TensorGrid(self.tensor_dims, self.grid)
        self.sparse_tensor = pm.SparseTensor(self.tensor_grid, self.tensor_idxs, self.values, preprocessing)


def write_factors(low_rank_tensor, filename, container="auto"):
    """Factor checkpoint: one dataset per mode, `FACTOR_MODE_i` (I_i x R,
    float64), the names LowRankTensor.write_to_file gives them
    (exafac/low_rank_tensor.py:186-192), plus the column norms as `SIGMA`
    so that read_factors restores the decomposition exactly.  HDF5 (each
    factor flattened row-major into a 1-D dataset, with `FACTOR_SHAPE_i`
    holding I_i and R) or .npz.  Works with the drop-in module's
    LowRankTensor (get_factor / get_sigma)."""
    n = low_rank_tensor.dim
    container = _container(container)
    data = {}
    for i in range(n):
        U = np.asarray(low_rank_tensor.get_factor(i), np.float64)
        if container == "hdf5":
            data[f"FACTOR_MODE_{i}"] = U.ravel()
            data[f"FACTOR_SHAPE_{i}"] = np.asarray(U.shape, np.uint64)
        else:
            data[f"FACTOR_MODE_{i}"] = U
    data["SIGMA"] = np.asarray(low_rank_tensor.get_sigma(), np.float64)
    _write_datasets(filename, data, container)
    return filename


def read_factors(low_rank_tensor, filename):
    """Restores factors and column norms written by write_factors into a
    LowRankTensor of the same shape (set_factor / set_sigma)."""
    n = low_rank_tensor.dim
    if h5.is_hdf5(filename):
        names = [f"FACTOR_MODE_{i}" for i in range(n)] + [f"FACTOR_SHAPE_{i}" for i in range(n)] + ["SIGMA"]
        f = h5.read(filename, names)
        for i in range(n):
            low_rank_tensor.set_factor(i, f[f"FACTOR_MODE_{i}"].reshape(tuple(int(x) for x in f[f"FACTOR_SHAPE_{i}"])))
        low_rank_tensor.set_sigma(f["SIGMA"])
        return
    with np.load(filename, allow_pickle=False) as f:
        for i in range(n):
            low_rank_tensor.set_factor(i, f[f"FACTOR_MODE_{i}"])
        if "SIGMA" in f:
            low_rank_tensor.set_sigma(f["SIGMA"])

