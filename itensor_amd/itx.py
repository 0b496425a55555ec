"""ctypes binding of the libitx C ABI (include/itx.h).

This is the Python view of the drop-in boundary: every function here is a
thin call into libitx.so, the library that the ITensor-compatible C++ host
layer links against.  Device buffers are passed as raw addresses (for
example torch tensors' data_ptr()); nothing here computes on the host.

The library must have been built (itensor_amd/build.py); a missing library
raises ImportError -- there is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import re
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libitx.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "itx.h")

ITX_OK = 0
PREC = {"f64": 0, "f32": 1, "bf16x3": 2}

_lib = None

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p

_PROTOS = {
    "itx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "itx_destroy": (C.c_int, [_vp]),
    "itx_set_stream": (C.c_int, [_vp, _vp]),
    "itx_get_stream": (_vp, [_vp]),
    "itx_synchronize": (C.c_int, [_vp]),
    "itx_last_error": (C.c_char_p, [_vp]),
    "itx_device_cus": (C.c_int, [_vp]),
    "itx_malloc": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "itx_free": (C.c_int, [_vp, _vp]),
    "itx_trim": (C.c_int, [_vp]),
    "itx_memcpy_htod": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "itx_memcpy_dtoh": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "itx_memcpy_dtod": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "itx_fill": (C.c_int, [_vp, _vp, C.c_size_t, C.c_double]),
    "itx_host_alloc": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "itx_host_free": (C.c_int, [_vp, _vp]),
    "itx_memcpy_async": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "itx_contract_dense": (C.c_int, [_vp, _vp, C.c_int, _i64p, _i32p, _vp, C.c_int, _i64p, _i32p,
                                     _vp, C.c_int, _i64p, _i32p, C.c_double, C.c_double, C.c_int]),
    "itx_contract_qn": (C.c_int, [_vp, _vp, C.c_int, _i32p, _vp, C.c_int, _i32p, _vp, C.c_int, _i32p,
                                  C.c_int, _i64p, _i64p, _i64p, C.c_double, C.c_int]),
    "itx_contract_qn_noalloc": (C.c_int, [_vp, _vp, C.c_int, _i32p, _vp, C.c_int, _i32p, _vp, C.c_int, _i32p,
                                          C.c_int, _i64p, _i64p, _i64p, C.c_double, C.c_int]),
    "itx_split_bf16": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "itx_plan_set_split": (C.c_int, [_vp, _vp, C.c_int]),
    "itx_plan_dense": (C.c_int, [_vp, C.c_int, _i64p, _i32p, C.c_int, _i64p, _i32p, C.c_int, _i64p,
                                 _i32p, C.c_int, C.POINTER(_vp)]),
    "itx_plan_qn": (C.c_int, [_vp, C.c_int, _i32p, C.c_int, _i32p, C.c_int, _i32p, C.c_int, _i64p,
                              _i64p, _i64p, C.c_int, C.POINTER(_vp)]),
    "itx_plan_execute": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_double, C.c_double]),
    "itx_plan_flops": (C.c_double, [_vp]),
    "itx_plan_info": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_int)]),
    "itx_plan_destroy": (C.c_int, [_vp, _vp]),
    "itx_dot": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _dp]),
    "itx_nrm2": (C.c_int, [_vp, _vp, C.c_size_t, _dp]),
    "itx_dot_dev": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _vp]),
    "itx_multi_dot": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, _vp, C.c_size_t, _dp]),
    "itx_axpy": (C.c_int, [_vp, C.c_double, _vp, _vp, C.c_size_t]),
    "itx_scal": (C.c_int, [_vp, C.c_double, _vp, C.c_size_t]),
    "itx_lincomb": (C.c_int, [_vp, C.POINTER(_vp), _dp, C.c_int, _vp, C.c_size_t]),
    "itx_permute": (C.c_int, [_vp, _vp, _vp, C.c_int, _i64p, _i32p, C.c_double, C.c_double]),
    "itx_syev": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "itx_syev_batched": (C.c_int, [_vp, C.c_int, _i32p, _i64p, _vp, _i64p, _vp]),
    "itx_syev_batched_host": (C.c_int, [_vp, C.c_int, _i32p, _i64p, _vp, _i64p, _vp, _vp]),
    "itx_block_copy": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _i64p, C.c_double, C.c_double]),
    "itx_set_autotune": (C.c_int, [_vp, C.c_int]),
    "itx_tune_save": (C.c_int, [_vp, C.c_char_p]),
    "itx_tune_load": (C.c_int, [_vp, C.c_char_p]),
    "itx_tune_size": (C.c_int, [_vp]),
    "itx_chain_create": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp)]),
    "itx_chain_launch": (C.c_int, [_vp, _vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), _dp, _dp]),
    "itx_chain_destroy": (C.c_int, [_vp, _vp]),
    "itx_gesvd": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "itx_gesvd_batched": (C.c_int, [_vp, C.c_int, _i32p, _i32p, _i64p, _vp, _i64p, _vp, _i64p, _vp, _i64p, _vp]),
}


def header_symbols() -> list[str]:
    """Names of all functions declared in include/itx.h."""
    with open(HEADER) as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(itx_[a-z0-9_]+)\s*\(", txt)))


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libitx.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _arr(t, vals):
    vals = list(vals)
    return (t * max(1, len(vals)))(*vals)


class ItxError(RuntimeError):
    pass


_live_contexts: "weakref.WeakSet[Context]" = weakref.WeakSet()


@atexit.register
def _close_all_contexts():
    """Destroy every open context at interpreter exit, while the HIP runtime
    is still up (itx_destroy also releases the plans and chains a context
    still holds); Plan / Chain objects freed later make no library call."""
    for c in list(_live_contexts):
        c.close()


class Context:
    """An itx_ctx bound to one device and one HIP stream."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self.L = lib()
        h = _vp()
        rc = self.L.itx_create(device, C.byref(h))
        if rc != ITX_OK:
            raise ItxError(f"itx_create failed ({rc}): {self.L.itx_last_error(None).decode()}")
        self.h = h
        _live_contexts.add(self)
        if stream is not None:
            self.set_stream(stream)

    def close(self):
        if getattr(self, "h", None):
            self.L.itx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int, what: str = ""):
        if rc != ITX_OK:
            raise ItxError(f"{what} failed ({rc}): {self.L.itx_last_error(self.h).decode()}")

    def set_stream(self, stream: int):
        self.check(self.L.itx_set_stream(self.h, _vp(stream)), "itx_set_stream")

    def sync(self):
        self.check(self.L.itx_synchronize(self.h), "itx_synchronize")

    @property
    def cus(self) -> int:
        return self.L.itx_device_cus(self.h)

    # ---- contractions (device addresses) ----
    def contract_dense(self, A: int, dimsA, labA, B: int, dimsB, labB, Cp: int, dimsC, labC,
                       alpha=1.0, beta=0.0, prec="f64"):
        rc = self.L.itx_contract_dense(
            self.h, _vp(A), len(dimsA), _arr(C.c_int64, dimsA), _arr(C.c_int32, labA),
            _vp(B), len(dimsB), _arr(C.c_int64, dimsB), _arr(C.c_int32, labB),
            _vp(Cp), len(dimsC), _arr(C.c_int64, dimsC), _arr(C.c_int32, labC),
            float(alpha), float(beta), PREC[prec])
        self.check(rc, "itx_contract_dense")

    def contract_qn(self, A: int, labA, B: int, labB, Cp: int, labC, offs, blkA, blkB,
                    alpha=1.0, prec="f64", noalloc=False):
        """itx_contract_qn, or with noalloc itx_contract_qn_noalloc (no
        device allocation: the plan goes into the library's static scratch
        array)."""
        offs = np.ascontiguousarray(offs, dtype=np.int64).ravel()
        blkA = np.ascontiguousarray(blkA, dtype=np.int64).ravel()
        blkB = np.ascontiguousarray(blkB, dtype=np.int64).ravel()
        nt = len(offs) // 3
        name = "itx_contract_qn_noalloc" if noalloc else "itx_contract_qn"
        rc = getattr(self.L, name)(
            self.h, _vp(A), len(labA), _arr(C.c_int32, labA), _vp(B), len(labB), _arr(C.c_int32, labB),
            _vp(Cp), len(labC), _arr(C.c_int32, labC), nt,
            offs.ctypes.data_as(_i64p), blkA.ctypes.data_as(_i64p), blkB.ctypes.data_as(_i64p),
            float(alpha), PREC[prec])
        self.check(rc, name)

    def split_bf16(self, x: int, y: int, n: int):
        """y (n 32-bit words) = the BF16 hi / lo terms of the n doubles at x."""
        self.check(self.L.itx_split_bf16(self.h, _vp(x), _vp(y), n), "itx_split_bf16")

    def plan_dense(self, dimsA, labA, dimsB, labB, dimsC, labC, prec="f64") -> "Plan":
        h = _vp()
        rc = self.L.itx_plan_dense(self.h, len(dimsA), _arr(C.c_int64, dimsA), _arr(C.c_int32, labA),
                                   len(dimsB), _arr(C.c_int64, dimsB), _arr(C.c_int32, labB),
                                   len(dimsC), _arr(C.c_int64, dimsC), _arr(C.c_int32, labC),
                                   PREC[prec], C.byref(h))
        self.check(rc, "itx_plan_dense")
        return Plan(self, h)

    def plan_qn(self, labA, labB, labC, offs, blkA, blkB, prec="f64") -> "Plan":
        offs = np.ascontiguousarray(offs, dtype=np.int64).ravel()
        blkA = np.ascontiguousarray(blkA, dtype=np.int64).ravel()
        blkB = np.ascontiguousarray(blkB, dtype=np.int64).ravel()
        h = _vp()
        rc = self.L.itx_plan_qn(self.h, len(labA), _arr(C.c_int32, labA), len(labB), _arr(C.c_int32, labB),
                                len(labC), _arr(C.c_int32, labC), len(offs) // 3,
                                offs.ctypes.data_as(_i64p), blkA.ctypes.data_as(_i64p),
                                blkB.ctypes.data_as(_i64p), PREC[prec], C.byref(h))
        self.check(rc, "itx_plan_qn")
        return Plan(self, h)

    def chain(self, plans: list["Plan"]) -> "Chain":
        arr = (_vp * len(plans))(*[p.h for p in plans])
        h = _vp()
        self.check(self.L.itx_chain_create(self.h, len(plans), arr, C.byref(h)), "itx_chain_create")
        return Chain(self, h, plans)

    # ---- tile autotuner ----
    def set_autotune(self, on: bool = True):
        self.check(self.L.itx_set_autotune(self.h, 1 if on else 0), "itx_set_autotune")

    def tune_save(self, path: str):
        self.check(self.L.itx_tune_save(self.h, path.encode()), "itx_tune_save")

    def tune_load(self, path: str):
        self.check(self.L.itx_tune_load(self.h, path.encode()), "itx_tune_load")

    def tune_size(self) -> int:
        return self.L.itx_tune_size(self.h)

    # ---- vector kernels ----
    def dot(self, x: int, y: int, n: int) -> float:
        r = C.c_double()
        self.check(self.L.itx_dot(self.h, _vp(x), _vp(y), n, C.byref(r)), "itx_dot")
        return r.value

    def nrm2(self, x: int, n: int) -> float:
        r = C.c_double()
        self.check(self.L.itx_nrm2(self.h, _vp(x), n, C.byref(r)), "itx_nrm2")
        return r.value

    def multi_dot(self, V: list[int], y: int, n: int) -> np.ndarray:
        out = np.zeros(len(V))
        ptrs = (_vp * len(V))(*[_vp(v) for v in V])
        self.check(self.L.itx_multi_dot(self.h, ptrs, len(V), _vp(y), n, out.ctypes.data_as(_dp)),
                   "itx_multi_dot")
        return out

    def axpy(self, a: float, x: int, y: int, n: int):
        self.check(self.L.itx_axpy(self.h, float(a), _vp(x), _vp(y), n), "itx_axpy")

    def scal(self, a: float, x: int, n: int):
        self.check(self.L.itx_scal(self.h, float(a), _vp(x), n), "itx_scal")

    def lincomb(self, V: list[int], c, y: int, n: int):
        ptrs = (_vp * len(V))(*[_vp(v) for v in V])
        cc = np.ascontiguousarray(c, dtype=np.float64)
        self.check(self.L.itx_lincomb(self.h, ptrs, cc.ctypes.data_as(_dp), len(V), _vp(y), n), "itx_lincomb")

    def permute(self, X: int, Y: int, dimsY, perm, alpha=1.0, beta=0.0):
        self.check(self.L.itx_permute(self.h, _vp(X), _vp(Y), len(dimsY), _arr(C.c_int64, dimsY),
                                      _arr(C.c_int32, perm), float(alpha), float(beta)), "itx_permute")

    def syev(self, n: int, A: int, w: int):
        self.check(self.L.itx_syev(self.h, n, _vp(A), _vp(w)), "itx_syev")

    def syev_batched(self, ns, a_offs, A: int, w_offs, w: int):
        ns = np.ascontiguousarray(ns, dtype=np.int32)
        ao = np.ascontiguousarray(a_offs, dtype=np.int64)
        wo = np.ascontiguousarray(w_offs, dtype=np.int64)
        self.check(self.L.itx_syev_batched(self.h, len(ns), ns.ctypes.data_as(_i32p), ao.ctypes.data_as(_i64p),
                                           _vp(A), wo.ctypes.data_as(_i64p), _vp(w)), "itx_syev_batched")

    def syev_batched_host(self, ns, a_offs, A: int, w_offs, w: int) -> np.ndarray:
        """itx_syev_batched_host: as syev_batched, returning the eigenvalues
        the call also hands to the host."""
        ns = np.ascontiguousarray(ns, dtype=np.int32)
        ao = np.ascontiguousarray(a_offs, dtype=np.int64)
        wo = np.ascontiguousarray(w_offs, dtype=np.int64)
        wh = np.zeros(int(max([o + n for o, n in zip(wo, ns)] + [1])))
        self.check(self.L.itx_syev_batched_host(self.h, len(ns), ns.ctypes.data_as(_i32p), ao.ctypes.data_as(_i64p),
                                                _vp(A), wo.ctypes.data_as(_i64p), _vp(w), _vp(wh.ctypes.data)),
                   "itx_syev_batched_host")
        return wh

    def gesvd(self, m: int, n: int, A: int, U: int, S: int, V: int):
        self.check(self.L.itx_gesvd(self.h, m, n, _vp(A), _vp(U), _vp(S), _vp(V)), "itx_gesvd")

    def gesvd_batched(self, ms, ns, a_offs, A: int, u_offs, U: int, s_offs, S: int, v_offs, V: int):
        arr = [np.ascontiguousarray(x, dtype=t) for x, t in
               ((ms, np.int32), (ns, np.int32), (a_offs, np.int64), (u_offs, np.int64), (s_offs, np.int64),
                (v_offs, np.int64))]
        self.check(self.L.itx_gesvd_batched(
            self.h, len(arr[0]), arr[0].ctypes.data_as(_i32p), arr[1].ctypes.data_as(_i32p),
            arr[2].ctypes.data_as(_i64p), _vp(A), arr[3].ctypes.data_as(_i64p), _vp(U),
            arr[4].ctypes.data_as(_i64p), _vp(S), arr[5].ctypes.data_as(_i64p), _vp(V)), "itx_gesvd_batched")

    def block_copy(self, X: int, Y: int, r: int, desc, alpha=1.0, beta=0.0):
        desc = np.ascontiguousarray(desc, dtype=np.int64).ravel()
        nt = len(desc) // (2 + 3 * r)
        self.check(self.L.itx_block_copy(self.h, _vp(X), _vp(Y), nt, r, desc.ctypes.data_as(_i64p),
                                         float(alpha), float(beta)), "itx_block_copy")


class Plan:
    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h

    @property
    def flops(self) -> float:
        return self.ctx.L.itx_plan_flops(self.h)

    @property
    def info(self) -> dict:
        """tile index, arithmetic form, block-pair tasks and work units."""
        t, f, nt, nu = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self.ctx.check(self.ctx.L.itx_plan_info(self.h, C.byref(t), C.byref(f), C.byref(nt), C.byref(nu)),
                       "itx_plan_info")
        return {"tile": t.value, "form": {v: k for k, v in PREC.items()}[f.value], "ntasks": nt.value,
                "nunits": nu.value}

    def set_split(self, on: bool = True):
        """BF16x3 plan reading pre-split operand words (itx_split_bf16)."""
        self.ctx.check(self.ctx.L.itx_plan_set_split(self.ctx.h, self.h, int(bool(on))), "itx_plan_set_split")

    def execute(self, A: int, B: int, Cp: int, alpha=1.0, beta=0.0):
        self.ctx.check(self.ctx.L.itx_plan_execute(self.ctx.h, self.h, _vp(A), _vp(B), _vp(Cp),
                                                   float(alpha), float(beta)), "itx_plan_execute")

    def close(self):
        # a destroyed context has released its plans already
        if self.h and self.ctx.h:
            self.ctx.L.itx_plan_destroy(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Chain:
    """Plan executions replayed as one HIP graph (itx_chain_*)."""

    def __init__(self, ctx: Context, h, plans):
        self.ctx, self.h, self.plans = ctx, h, list(plans)

    def launch(self, A: list[int], B: list[int], Cp: list[int], alpha=None, beta=None):
        n = len(self.plans)
        pa = (_vp * n)(*[_vp(x) for x in A])
        pb = (_vp * n)(*[_vp(x) for x in B])
        pc = (_vp * n)(*[_vp(x) for x in Cp])
        al = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
        be = None if beta is None else np.ascontiguousarray(beta, dtype=np.float64)
        self.ctx.check(self.ctx.L.itx_chain_launch(
            self.ctx.h, self.h, pa, pb, pc,
            None if al is None else al.ctypes.data_as(_dp),
            None if be is None else be.ctypes.data_as(_dp)), "itx_chain_launch")

    def close(self):
        # a destroyed context has released its chains already
        if self.h and self.ctx.h:
            self.ctx.L.itx_chain_destroy(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
