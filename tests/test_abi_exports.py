"""CPU: the C-ABI library loads and exports every symbol include/itx.h declares."""
import ctypes
import os

import pytest

from itensor_amd import itx


def test_header_parses():
    syms = itx.header_symbols()
    assert "itx_contract_dense" in syms and "itx_contract_qn" in syms and "itx_syev" in syms


def test_library_exports_all_declared_symbols():
    if not os.path.exists(itx.LIB_PATH):
        pytest.fail("libitx.so not built: run __graft_entry__.build()")
    L = ctypes.CDLL(itx.LIB_PATH)
    missing = [s for s in itx.header_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_ctypes_prototypes_cover_header():
    assert set(itx.header_symbols()) == set(itx._PROTOS)


def test_create_without_gpu_fails_cleanly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(itx.ItxError):
        itx.Context(0)
