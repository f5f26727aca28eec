"""CPU tests of the tensor-file ingest (host code in the C-ABI library; no
GPU call): the .tns parser against NumPy's own text parsing of the same
file, the per-mode minimum / maximum and the shift by the minimum
(process_frostt_tensor.cpp:115-165, sparse_tensor_e.py:68-103), the
converter's datasets, and errors that name the offending line."""
import os

import numpy as np
import pytest

from rdist_tensor_amd import abi, tensor_io


def write_tns(path, coords, vals, fmt="%.17g", sep=" "):
    with open(path, "w") as f:
        for c, v in zip(coords, vals):
            f.write(sep.join(str(int(x)) for x in c) + sep + (fmt % v) + "\n")


def test_read_tns_matches_numpy(tmp_path):
    rng = np.random.default_rng(7)
    nnz, dims = 5000, [40, 3000, 7, 100000]
    coords = np.stack([rng.integers(1, d + 1, nnz) for d in dims], 1)
    vals = rng.standard_normal(nnz) * 10.0 ** rng.integers(-5, 6, nnz)
    p = tmp_path / "t.tns"
    write_tns(p, coords, vals)
    c, v, mn, mx = tensor_io.read_tns(p)
    ref = np.loadtxt(p)
    assert c.dtype == np.uint64 and c.shape == (nnz, 4)
    assert np.array_equal(c, ref[:, :4].astype(np.uint64))
    assert np.array_equal(v, ref[:, 4])             # both round correctly
    assert np.array_equal(mn, coords.min(0)) and np.array_equal(mx, coords.max(0))
    idx, vals2, d = tensor_io.load_tensor_file(p)
    assert np.array_equal(idx, (coords - coords.min(0)).astype(np.uint32))
    assert np.array_equal(d, coords.max(0) - coords.min(0) + 1)
    assert np.array_equal(vals2, v)


def test_tns_layout_variants(tmp_path):
    """Tabs, carriage returns, comments, blank lines, a missing final newline,
    exponent notation, a leading '+' and 64-bit coordinates."""
    p = tmp_path / "v.tns"
    p.write_text("# a comment\n\n1\t2\t3\t1.5e-3\r\n  4 5 6   -2\n\n# tail\n7 8 18446744073709551615 +3.25")
    c, v, mn, mx = tensor_io.read_tns(p)
    assert c.tolist() == [[1, 2, 3], [4, 5, 6], [7, 8, 18446744073709551615]]
    assert v.tolist() == [1.5e-3, -2.0, 3.25]
    assert mn.tolist() == [1, 2, 3] and mx.tolist() == [7, 8, 18446744073709551615]


def test_converter_datasets(tmp_path):
    rng = np.random.default_rng(8)
    coords = np.stack([rng.integers(5, 60, 300) for _ in range(3)], 1)
    vals = rng.random(300)
    p = tmp_path / "c.tns"
    write_tns(p, coords, vals)
    out = tensor_io.tns_to_tensor_file(p, index_dtype=np.uint32, container="npz")
    assert out == str(p) + "_converted.npz"
    with np.load(out, allow_pickle=False) as f:
        assert sorted(f.files) == ["MAX_MODE_SET", "MIN_MODE_SET", "MODE_0", "MODE_1", "MODE_2", "VALUES"]
        for k in range(3):
            assert f[f"MODE_{k}"].dtype == np.uint32
            assert np.array_equal(f[f"MODE_{k}"], coords[:, k])
        assert np.array_equal(f["MIN_MODE_SET"], coords.min(0))
        assert np.array_equal(f["MAX_MODE_SET"], coords.max(0))
        assert np.array_equal(f["VALUES"], vals)
    idx, v, d = tensor_io.load_tensor_file(out)
    idx2, v2, d2 = tensor_io.load_tensor_file(p)
    assert np.array_equal(idx, idx2) and np.array_equal(v, v2) and np.array_equal(d, d2)


@pytest.mark.parametrize("text, where", [
    ("1 2 3 4.0\n1 2 3.0\n", "line 2"),      # fewer fields than the first line
    ("1 2 3 4.0\n1 x 3 4.0\n", "line 2"),    # non-integer coordinate
    ("1 2 3 4.0\n1 2 3 4.0zz\n", "line 2"),  # malformed value
    ("1 -2 3 4.0\n", "line 1"),              # negative coordinate
])
def test_malformed_lines_are_reported(tmp_path, text, where):
    p = tmp_path / "bad.tns"
    p.write_text(text)
    with pytest.raises(abi.ExafacError, match=where):
        tensor_io.read_tns(p)


def test_missing_file(tmp_path):
    with pytest.raises(abi.ExafacError, match="cannot open"):
        tensor_io.read_tns(tmp_path / "absent.tns")


needs_hdf5 = pytest.mark.skipif(not __import__("rdist_tensor_amd.h5").h5.available(), reason="no libhdf5")


@needs_hdf5
@pytest.mark.parametrize("index_dtype", [np.uint32, np.uint64])
def test_converter_hdf5_round_trip(tmp_path, index_dtype):
    """The converter's HDF5 file: the reference's name, the HDF5 signature,
    every dataset back with its stored type, and the reader's shifted
    coordinates equal to the .tns path's."""
    from rdist_tensor_amd import h5
    rng = np.random.default_rng(9)
    coords = np.stack([rng.integers(3, 5000, 2000) for _ in range(4)], 1)
    vals = rng.standard_normal(2000)
    p = tmp_path / "h.tns"
    write_tns(p, coords, vals)
    out = tensor_io.tns_to_tensor_file(p, index_dtype=index_dtype)
    assert out == str(p) + "_converted.hdf5" and h5.is_hdf5(out)
    f = h5.read(out, ["MODE_0", "MODE_1", "MODE_2", "MODE_3", "VALUES", "MIN_MODE_SET", "MAX_MODE_SET"])
    for k in range(4):
        assert f[f"MODE_{k}"].dtype == index_dtype
        assert np.array_equal(f[f"MODE_{k}"], coords[:, k])
    assert np.array_equal(f["VALUES"], vals)
    assert np.array_equal(f["MIN_MODE_SET"], coords.min(0)) and np.array_equal(f["MAX_MODE_SET"], coords.max(0))
    idx, v, d = tensor_io.load_tensor_file(out)
    idx2, v2, d2 = tensor_io.load_tensor_file(p)
    assert np.array_equal(idx, idx2) and np.array_equal(v, v2) and np.array_equal(d, d2)


@needs_hdf5
def test_hdf5_file_read_by_h5dump(tmp_path):
    """An independent reader (the HDF5 distribution's h5dump) sees the
    datasets this package writes."""
    import shutil
    import subprocess
    from rdist_tensor_amd import h5
    exe = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
    if exe is None:
        pytest.skip("h5dump is not installed")
    out = tmp_path / "d.hdf5"
    h5.write(out, {"MODE_0": np.array([5, 7, 11], np.uint32), "VALUES": np.array([0.5, -2.0, 3.25])})
    text = subprocess.run([exe, str(out)], capture_output=True, text=True, check=True).stdout
    assert "H5T_STD_U32LE" in text and "5, 7, 11" in text
    assert "H5T_IEEE_F64LE" in text and "0.5, -2, 3.25" in text


REF_CONVERTER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                             "process_frostt_tensor")


@needs_hdf5
@pytest.mark.skipif(not os.path.exists(REF_CONVERTER), reason="the reference converter is not built (oracle/_ref)")
def test_reads_the_reference_converters_file(tmp_path):
    """The reference's own converter (tensor_io/process_frostt_tensor.cpp,
    built from its source into oracle/_ref) turns a .tns into HDF5; this
    package's reader returns the same shifted coordinates, values and mode
    sizes as its .tns parser."""
    import subprocess
    rng = np.random.default_rng(10)
    coords = np.stack([rng.integers(2, 9000, 3000) for _ in range(3)], 1)
    vals = rng.standard_normal(3000)
    p = tmp_path / "r.tns"
    write_tns(p, coords, vals)
    subprocess.run([REF_CONVERTER, str(p), str(len(vals))], capture_output=True, check=True)
    ref_file = str(p) + "_converted.hdf5"
    idx, v, d = tensor_io.load_tensor_file(ref_file)
    idx2, v2, d2 = tensor_io.load_tensor_file(p)
    assert np.array_equal(idx, idx2) and np.array_equal(v, v2) and np.array_equal(d, d2)
