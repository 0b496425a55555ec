"""Build recipe of the native parts of itensor_amd (run in-tree).

Targets (all written under itensor_amd/lib/):
  libitx.so            HIP kernels + the C ABI declared in include/itx.h
                       (hipcc, --offload-arch=gfx950)
  libitensor_host.so   the ITensor-compatible C++ host layer (itensor_amd/host),
                       linked against libitx.so
  _itensor_host*.so    pybind11 module exposing the host layer to Python

Usage:  python -m itensor_amd.build   (or __graft_entry__.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIBDIR = os.path.join(PKG, "lib")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["itx_runtime.hip", "itx_contract.hip", "itx_vec.hip", "itx_eig.hip", "itx_svd.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for all device code (they have
# produced wrong results on this hardware); -O3 for the kernels.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-sign-compare", "-Wno-unused-variable"]


def _run(cmd: list[str]) -> None:
    print("[build]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = (r.stdout or "") + (r.stderr or "")
    # hipcc's host pass does not know the packed-fp32 feature flag: harmless
    lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
    if r.returncode != 0:
        sys.stderr.write("\n".join(lines[-80:]) + "\n")
        raise RuntimeError(f"build step failed ({r.returncode}): {cmd[0]}")


def _jobs() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def _newer(target: str, sources: list[str]) -> bool:
    if not os.path.exists(target):
        return False
    t = os.path.getmtime(target)
    return all(os.path.getmtime(s) <= t for s in sources)


def build_libitx(force: bool = True) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libitx.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    deps = srcs + [os.path.join(CSRC, "itx_internal.h"), os.path.join(INCLUDE, "itx.h")]
    if not force and _newer(out, deps):
        return out
    # one object per source, compiled side by side, then one link
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in srcs]
    cflags = [f for f in HIP_FLAGS if f != "-shared"]
    hdrs = deps[len(srcs):]
    todo = [(s, o) for s, o in zip(srcs, objs) if force or not _newer(o, [s, *hdrs])]
    with ThreadPoolExecutor(max_workers=min(len(todo), _jobs()) or 1) as ex:
        for f in [ex.submit(_run, [HIPCC, *cflags, "-c", s, "-o", o]) for s, o in todo]:
            f.result()
    tmp = out + ".tmp"
    # rocSOLVER (with rocBLAS) diagonalises the density-matrix blocks above
    # 256 rows
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp, "-lrocsolver", "-lrocblas"])
    os.replace(tmp, out)
    return out


def _host_sources() -> list[str]:
    if not os.path.isdir(HOST):
        return []
    return sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cc"))


def _host_headers() -> list[str]:
    if not os.path.isdir(HOST):
        return []
    return sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".h"))


def build_host(force: bool = True) -> str | None:
    srcs = [s for s in _host_sources() if not s.endswith("pybind.cc")]
    if not srcs:
        return None
    out = os.path.join(LIBDIR, "libitensor_host.so")
    deps = srcs + _host_headers() + [os.path.join(INCLUDE, "itx.h"), os.path.join(LIBDIR, "libitx.so")]
    if not force and _newer(out, deps):
        return out
    tmp = out + ".tmp"
    _run([CXX, *CXX_FLAGS, "-shared", f"-I{INCLUDE}", f"-I{PKG}", *srcs, "-o", tmp,
          f"-L{LIBDIR}", "-litx", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_pybind(force: bool = True) -> str | None:
    src = os.path.join(HOST, "pybind.cc")
    if not os.path.exists(src):
        return None
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(LIBDIR, "_itensor_host" + suffix)
    deps = [src] + _host_headers() + [os.path.join(LIBDIR, "libitensor_host.so")]
    if not force and _newer(out, deps):
        return out
    py_inc = sysconfig.get_paths()["include"]
    tmp = out + ".tmp"
    _run([CXX, *CXX_FLAGS, "-shared", f"-I{INCLUDE}", f"-I{PKG}", f"-I{pybind11.get_include()}",
          f"-I{py_inc}", src, "-o", tmp, f"-L{LIBDIR}", "-litensor_host", "-litx",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_all(force: bool = True) -> None:
    build_libitx(force)
    build_host(force)
    build_pybind(force)


if __name__ == "__main__":
    build_all(force="--lazy" not in sys.argv)
