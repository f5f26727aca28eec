"""Builds the native parts of rdist_tensor_amd in-tree.

* libexafac_hip.so  -- HIP kernels for gfx950 + the C-ABI (include/exafac_hip.h)
* cpp_ext/py_module*.so -- pybind11 module with rdist_tensor's Python surface,
  bound over the C-ABI

Usage: python -m rdist_tensor_amd.build   (or rdist_tensor_amd.build.build()).
Every call recompiles what is out of date; force=True rebuilds everything.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libexafac_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["api.hip", "kernels_dense.hip", "kernels_sample.hip", "kernels_sparse.hip", "kernels_mixed.hip",
               "kernels_batch.hip", "sort.hip", "tensor_io.hip"]
HEADERS = ["common.hpp", "device.hpp", "dsymm.hpp", "spd_block.hpp", "jacobi_block.hpp"]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions (v_pk_*_f32) stay off: some of their forms have
    # given wrong results on this hardware without an error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_hip(force=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    hdrs = [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(ROOT, "include", "exafac_hip.h")]
    objs, todo = [], []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(o)
        if force or _stale(o, [s] + hdrs):
            todo.append([HIPCC] + HIP_FLAGS + ["-c", s, "-o", o])
    if todo:
        with ThreadPoolExecutor(max_workers=jobs or min(len(todo), 8)) as ex:
            list(ex.map(_run, todo))
    if force or todo or _stale(LIB, objs):
        tmp = LIB + ".tmp"
        _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
        os.replace(tmp, LIB)
    return LIB


def pymodule_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(PKG, "cpp_ext", "py_module" + suffix)


def build_pymodule(force=False):
    import pybind11
    src = os.path.join(PKG, "cpp_ext", "py_module.cpp")
    out = pymodule_path()
    deps = [src, LIB, os.path.join(ROOT, "include", "exafac_hip.h")]
    if force or _stale(out, deps):
        cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
               "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
               "-I", os.path.join(ROOT, "include"),
               src, "-o", out + ".tmp",
               "-L", PKG, "-lexafac_hip", "-Wl,-rpath,$ORIGIN/.."]
        _run(cmd)
        os.replace(out + ".tmp", out)
    return out


def build(force=False):
    build_hip(force=force)
    build_pymodule(force=force)


if __name__ == "__main__":
    build(force="--force" in sys.argv)
    print("built", LIB, pymodule_path())
