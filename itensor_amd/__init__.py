"""MI355X-native DMRG H*psi engine with an ITensor-compatible host layer.

itensor_amd.itx   ctypes view of the C ABI (include/itx.h, libitx.so)
itensor_amd.host  pybind11 view of the C++ host layer (libitensor_host.so)

torch (when installed) is imported before any native library of this
package: torch ships its own libamdhip64.so.7 and the dynamic loader binds
every later user of that soname to the copy loaded first, so importing
torch first keeps a single HIP runtime in the process.
"""
try:  # noqa: SIM105
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional for the C++ layer
    pass
