"""rdist_tensor_amd: rdist_tensor's randomized CP-ALS inner loop (CP-ARLS-LEV,
STS-CP) on one AMD MI355X.

* ``rdist_tensor_amd.cpp_ext.py_module`` -- the pybind11 module with the Python
  surface of rdist_tensor's ``exafac.cpp_ext.py_module``.
* ``rdist_tensor_amd.abi`` -- ctypes binding of the C-ABI (include/exafac_hip.h).
* ``rdist_tensor_amd.install_exafac_alias()`` -- makes
  ``from exafac.cpp_ext.py_module import ...`` resolve to this build, so
  rdist_tensor's decompose.py runs unchanged apart from its MPI imports.
"""
import importlib
import os
import sys

_PKG = os.path.dirname(os.path.abspath(__file__))


def native_paths():
    from . import build as _b
    return {"libexafac_hip": _b.LIB, "py_module": _b.pymodule_path()}


def py_module():
    """Import the pybind11 module; raises ImportError if it has not been built."""
    return importlib.import_module("rdist_tensor_amd.cpp_ext.py_module")


def install_exafac_alias():
    """Register ``exafac``, ``exafac.cpp_ext`` and ``exafac.cpp_ext.py_module``
    as aliases of this build in sys.modules."""
    import types
    mod = py_module()
    pkg = sys.modules.get("exafac") or types.ModuleType("exafac")
    pkg.__path__ = []
    sub = sys.modules.get("exafac.cpp_ext") or types.ModuleType("exafac.cpp_ext")
    sub.__path__ = []
    sub.py_module = mod
    pkg.cpp_ext = sub
    sys.modules["exafac"] = pkg
    sys.modules["exafac.cpp_ext"] = sub
    sys.modules["exafac.cpp_ext.py_module"] = mod
    return mod
