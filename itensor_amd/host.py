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

Index = _h.Index
ITensor = _h.ITensor
ITError = _h.ITError
dense = _h.dense
qdense = _h.qdense
norm = _h.norm
dot = _h.dot
remove_qns = _h.remove_qns
combiner = _h.combiner
diag_hermitian = _h.diag_hermitian
denmat_decomp = _h.denmat_decomp
svd = _h.svd
svd_spec = _h.svd_spec
factor = _h.factor
truncate = _h.truncate
davidson_bond = _h.davidson_bond
local_product = _h.local_product
dmrg = _h.dmrg
set_stream = _h.set_stream
synchronize = _h.synchronize
set_contract_precision = _h.set_contract_precision
set_autotune = _h.set_autotune
LocalOp = _h.LocalOp
random_itensor = _h.random_itensor
block_offsets = _h.block_offsets
qn_task_list = _h.qn_task_list
from_dense = _h.from_dense
SiteSet = _h.SiteSet
AutoMPO = _h.AutoMPO
spin_half_sites = _h.spin_half_sites
electron_sites = _h.electron_sites
square_lattice = _h.square_lattice
HPSI_EAGER, HPSI_PLANNED, HPSI_GRAPH, HPSI_PER_BLOCK = (_h.HPSI_EAGER, _h.HPSI_PLANNED, _h.HPSI_GRAPH,
                                                        _h.HPSI_PER_BLOCK)

PREC_F64, PREC_F32, PREC_BF16X3 = 0, 1, 2
