"""Python view of the ITensor-compatible C++ host layer (itensor_amd/host).

Loads the pybind11 module built in-tree (itensor_amd/lib/_itensor_host*.so),
which links libitensor_host.so and libitx.so.  A missing build raises
ImportError: there is no Python or CPU implementation behind these names.
"""
from __future__ import annotations

import glob
import importlib.util
import os

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")


def _load():
    cands = sorted(glob.glob(os.path.join(_LIB, "_itensor_host*.so")))
    if not cands:
        raise ImportError(f"_itensor_host is not built in {_LIB}; run __graft_entry__.build()")
    spec = importlib.util.spec_from_file_location("_itensor_host", cands[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


_h = _load()

# every public name of the module: the ITensor-compatible classes (Index,
# ITensor, LocalOp, SiteSet, AutoMPO), the free functions of the path
# (contraction helpers, decompositions, davidson, dmrg, svdBond, position,
# binary read/write) and the H*psi execution modes
__all__ = [n for n in dir(_h) if not n.startswith("_")]
globals().update({n: getattr(_h, n) for n in __all__})

PREC_F64, PREC_F32, PREC_BF16X3 = 0, 1, 2
