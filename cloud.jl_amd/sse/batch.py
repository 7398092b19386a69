"""Batched independent problems on one GPU (libsse_hip's ``sse_batch_*``).

The reference solves one problem per ``Solver``; a study over many problems
loops over ``solve`` calls. ``BatchedSolver`` instead groups problems that share
a discretisation (element type, degree, operators, law, form) into one set of
kernel launches per phase, and replays the whole CarpenterKennedy2N54 step of
all groups as one HIP graph. Scalar StandardForm problems on one mesh that
differ only in the diffusion coefficient (a parameter sweep; build them with
``Solver.with_law``) form an ensemble that shares one copy of the geometry.
"""

from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import abi
from .solvers import Solver


class BatchError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sse batch error {code}: {msg}")
        self.code = code


class BatchedSolver:
    """Many ``Solver``s advanced together; states stay the callers' arrays."""

    def __init__(self, solvers: Sequence[Solver]):
        from .hip import load_library
        self.lib = load_library()
        self.solvers = list(solvers)
        n = len(self.solvers)
        ptrs = (C.POINTER(abi.SseDesc) * n)(*[C.pointer(s.arrays.desc) for s in self.solvers])
        h = C.c_void_p()
        code = self.lib.sse_batch_create(ptrs, n, C.byref(h))
        if code:
            raise BatchError(code, self.lib.sse_batch_last_error(None).decode())
        self.handle = h

    def describe(self) -> str:
        buf = C.create_string_buffer(4096)
        self.lib.sse_batch_describe(self.handle, buf, 4096)
        return buf.value.decode()

    def rk_steps(self, states, dts, nsteps: int, use_graph: bool = True):
        """Advances state i (a CUDA tensor of shape (N_e, N_c, N_p), or a Fortran
        NumPy array (N_p, N_c, N_e)) by ``nsteps`` steps of ``dts[i]``, in place."""
        n = len(self.solvers)
        if len(states) != n or len(dts) != n:
            raise ValueError("one state and one dt per problem")
        ptrs, host = self._pointers(states, "state")
        dt = np.ascontiguousarray(dts, dtype=np.float64)
        code = self.lib.sse_batch_rk_steps(self.handle, ptrs, dt.ctypes.data_as(C.POINTER(C.c_double)),
                                           int(nsteps), abi.SSE_HOST if host else abi.SSE_DEVICE,
                                           1 if use_graph else 0)
        if code:
            raise BatchError(code, self.lib.sse_batch_last_error(self.handle).decode())
        return states

    def _pointers(self, states, what):
        n = len(self.solvers)
        host = isinstance(states[0], np.ndarray)
        ptrs = (C.c_void_p * n)()
        for i, (x, s) in enumerate(zip(states, self.solvers)):
            N_p, N_c, N_e = s.size()
            if host:
                if not (x.flags.f_contiguous and x.shape == (N_p, N_c, N_e) and x.dtype == np.float64):
                    raise ValueError(f"{what} {i} must be a Fortran float64 array {(N_p, N_c, N_e)}")
                ptrs[i] = x.ctypes.data
            else:
                if not (x.is_cuda and x.is_contiguous() and tuple(x.shape) == (N_e, N_c, N_p)):
                    raise ValueError(f"{what} {i} must be a contiguous CUDA tensor {(N_e, N_c, N_p)}")
                ptrs[i] = x.data_ptr()
        return ptrs, host

    def rk_tangent(self, states, tangents, dts, nsteps: int):
        """Advances each state u_i and its derivative v_i = du_i/db_i with respect
        to the problem's diffusion coefficient together (``sse_batch_rk_tangent``),
        in place. Every problem must belong to an ensemble (advection-diffusion
        problems sharing one mesh, see ``Solver.with_law``)."""
        n = len(self.solvers)
        if len(states) != n or len(tangents) != n or len(dts) != n:
            raise ValueError("one state, one tangent and one dt per problem")
        pu, host = self._pointers(states, "state")
        pv, host_v = self._pointers(tangents, "tangent")
        if host != host_v:
            raise ValueError("states and tangents must both be host arrays or both CUDA tensors")
        dt = np.ascontiguousarray(dts, dtype=np.float64)
        code = self.lib.sse_batch_rk_tangent(self.handle, pu, pv, dt.ctypes.data_as(C.POINTER(C.c_double)),
                                             int(nsteps), abi.SSE_HOST if host else abi.SSE_DEVICE)
        if code:
            raise BatchError(code, self.lib.sse_batch_last_error(self.handle).decode())
        return states, tangents

    def close(self):
        if self.handle is not None and self.handle.value:
            self.lib.sse_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
