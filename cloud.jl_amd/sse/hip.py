"""ctypes binding of libsse_hip.so (the HIP residual library).

The library is built in-tree by ``__graft_entry__.build()`` into
``cloud.jl_amd/csrc/libsse_hip.so``. There is no CPU fallback: if the library
is missing, loading raises.
"""

from __future__ import annotations

import ctypes as C
import os

from .abi import EXPORTED_SYMBOLS, SSE_ABI_VERSION, SseDesc, SseGeometryOps

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "csrc", "libsse_hip.so"))

_lib = None


class HipLibraryMissing(RuntimeError):
    pass


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libsse_hip.so and declare its signatures (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise HipLibraryMissing(
            f"{path} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7);
    # loading it first makes the library and torch share one runtime instance.
    import torch  # noqa: F401
    lib = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    lib.sse_create.argtypes = [C.POINTER(SseDesc), C.POINTER(vp)]
    lib.sse_create.restype = C.c_int
    lib.sse_residual.argtypes = [vp, vp, vp, C.c_double, C.c_int32]
    lib.sse_residual.restype = C.c_int
    lib.sse_rk_steps.argtypes = [vp, vp, C.c_double, C.c_double, C.c_int32, C.c_int32]
    lib.sse_rk_steps.restype = C.c_int
    lib.sse_ec_flux_batch.argtypes = [C.c_int32, C.c_double, C.c_int32, vp, vp, vp, vp, C.c_int32]
    lib.sse_ec_flux_batch.restype = C.c_int
    lib.sse_erk_steps.argtypes = [vp, vp, C.c_double, C.c_double, C.c_int32, C.c_int32, dp, dp, C.c_int32]
    lib.sse_erk_steps.restype = C.c_int
    lib.sse_profile_residual.argtypes = [vp, vp, vp, C.c_int32, dp]
    lib.sse_profile_residual.restype = C.c_int
    lib.sse_set_stream.argtypes = [vp, vp]
    lib.sse_set_stream.restype = C.c_int
    lib.sse_set_diffusion.argtypes = [vp, C.c_double]
    lib.sse_set_diffusion.restype = C.c_int
    lib.sse_synchronize.argtypes = [vp]
    lib.sse_synchronize.restype = C.c_int
    lib.sse_last_error.argtypes = [vp]
    lib.sse_last_error.restype = C.c_char_p
    lib.sse_destroy.argtypes = [vp]
    lib.sse_destroy.restype = None
    lib.sse_describe.argtypes = [vp, C.c_char_p, C.c_int32]
    lib.sse_describe.restype = C.c_int
    lib.sse_batch_create.argtypes = [C.POINTER(C.POINTER(SseDesc)), C.c_int32, C.POINTER(vp)]
    lib.sse_batch_create.restype = C.c_int
    lib.sse_batch_rk_steps.argtypes = [vp, C.POINTER(vp), dp, C.c_int32, C.c_int32, C.c_int32]
    lib.sse_batch_rk_steps.restype = C.c_int
    lib.sse_batch_rk_tangent.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), dp, C.c_int32, C.c_int32]
    lib.sse_batch_rk_tangent.restype = C.c_int
    lib.sse_batch_describe.argtypes = [vp, C.c_char_p, C.c_int32]
    lib.sse_batch_describe.restype = C.c_int
    lib.sse_batch_last_error.argtypes = [vp]
    lib.sse_batch_last_error.restype = C.c_char_p
    lib.sse_batch_destroy.argtypes = [vp]
    lib.sse_batch_destroy.restype = None
    lib.sse_functional.argtypes = [vp, vp, vp, C.c_int32, C.c_int32, dp]
    lib.sse_functional.restype = C.c_int
    lib.sse_functional_noalloc.argtypes = [vp, vp, vp, C.c_int32, vp]
    lib.sse_functional_noalloc.restype = C.c_int
    lib.sse_residual_vjp.argtypes = [vp, vp, vp, vp, dp, C.c_int32]
    lib.sse_residual_vjp.restype = C.c_int
    lib.sse_rk_adjoint.argtypes = [vp, vp, C.c_double, C.c_int32, C.c_int32, C.c_int32, dp, vp, dp, C.c_int32]
    lib.sse_rk_adjoint.restype = C.c_int
    lib.sse_geometric_factors.argtypes = [C.POINTER(SseGeometryOps), C.c_int32, vp, vp, vp, vp, vp, vp, vp, vp,
                                          C.c_int32]
    lib.sse_geometric_factors.restype = C.c_int
    lib.sse_apply_reference_mapping.argtypes = [C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, vp, C.c_int32]
    lib.sse_apply_reference_mapping.restype = C.c_int
    lib.sse_project_l2.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, C.c_int32]
    lib.sse_project_l2.restype = C.c_int
    lib.sse_l2_error.argtypes = [vp, vp, vp, C.c_int32, C.c_int32, dp]
    lib.sse_l2_error.restype = C.c_int
    lib.sse_abi_version.argtypes = []
    lib.sse_abi_version.restype = C.c_int32
    for name in EXPORTED_SYMBOLS:
        getattr(lib, name)
    if lib.sse_abi_version() != SSE_ABI_VERSION:
        raise RuntimeError("libsse_hip.so ABI version mismatch")
    _lib = lib
    return lib


class SseError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sse error {code}: {msg}")
        self.code = code


def check(lib, handle, code: int):
    if code != 0:
        msg = lib.sse_last_error(handle)
        raise SseError(code, msg.decode() if msg else "")
