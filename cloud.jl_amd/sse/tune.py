"""Launch-shape autotuner for one discretisation on its own data.

libsse_hip compiles, per (element type, degree), a few kernel configurations:
the kernel family of the time derivative (line-structured, table-driven pairs,
runtime-shaped generic) and, for the tensor kernels, launch-shape variants
(threads per element, elements per workgroup, pairs per thread of the
projection and pair kernels). A configuration is chosen when a handle is
created, from ``SSE_KERNELS`` (``tensor`` or ``generic``; unset = the fastest
family the shape has) and ``SSE_TENSOR_VARIANT``.

``autotune(solver, u)`` creates one handle per configuration on the solver's
arrays (``Solver.with_law`` shares them), times the residual with HIP events on
the handle's stream (``sse_profile_residual``), checks that every configuration
gives the same residual, and returns the timings with the fastest first.
``select(best)`` sets the environment so that handles created afterwards use it.

The reference has no counterpart: its ``Serial()``/``Threaded()`` parallelism
needs no launch shapes.
"""

from __future__ import annotations

import ctypes as C
import os
from contextlib import contextmanager
from dataclasses import dataclass
from typing import List, Optional

from .solvers import Solver

# launch-shape variants compiled for some (element, degree) shapes; a variant that
# a shape lacks falls back to variant 0, and duplicates are dropped by description
MAX_VARIANT = 5
FAMILIES = ("struct", "tensor", "generic")


@dataclass
class Timing:
    family: str
    variant: Optional[int]
    describe: str
    ms: List[float]          # per kernel slot (nodal values, time derivative)

    @property
    def total_ms(self) -> float:
        return sum(self.ms)


@contextmanager
def _selection(family: str, variant: Optional[int]):
    keys = ("SSE_KERNELS", "SSE_TENSOR_VARIANT")
    saved = {k: os.environ.get(k) for k in keys}
    try:
        if family == "struct":
            os.environ.pop("SSE_KERNELS", None)
        else:
            os.environ["SSE_KERNELS"] = family
        if variant is None:
            os.environ.pop("SSE_TENSOR_VARIANT", None)
        else:
            os.environ["SSE_TENSOR_VARIANT"] = str(variant)
        yield
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _profile(solver: Solver, u, dudt, nrep: int) -> List[float]:
    from . import abi
    from .hip import check
    lib, h = solver.lib, solver.handle
    ms = (C.c_double * abi.SSE_NUM_KERNELS)()
    check(lib, h, lib.sse_profile_residual(h, u.data_ptr(), dudt.data_ptr(), int(nrep), ms))
    return list(ms)


def autotune(solver: Solver, u, nrep: int = 20, check_tol: float = 1e-12) -> List[Timing]:
    """Times every compiled configuration of `solver`'s discretisation on the device
    state `u` (a CUDA tensor in the solver's layout); fastest first. Raises if two
    configurations disagree on the residual by more than `check_tol` (relative)."""
    import torch
    results: List[Timing] = []
    seen = set()
    ref = None
    for family in FAMILIES:
        variants = [None] if family == "generic" else list(range(MAX_VARIANT))
        for v in variants:
            with _selection(family, v):
                s = solver.with_law(solver.conservation_law)
                try:
                    desc = s.describe()
                    if desc in seen:
                        continue
                    seen.add(desc)
                    dudt = torch.empty_like(u)
                    _profile(s, u, dudt, 2)  # warm-up: kernel attributes, caches
                    ms = _profile(s, u, dudt, nrep)
                    torch.cuda.synchronize()
                    if ref is None:
                        ref = dudt.clone()
                    else:
                        err = float((dudt - ref).abs().max() / ref.abs().max())
                        if err > check_tol:
                            raise RuntimeError(f"{desc}: residual differs from the first configuration by {err:.3e}")
                    results.append(Timing(family, v, desc, ms))
                finally:
                    s.close()
    results.sort(key=lambda t: t.total_ms)
    return results


def select(best: Timing) -> None:
    """Makes handles created from now on use `best`'s configuration."""
    if best.family == "struct":
        os.environ.pop("SSE_KERNELS", None)
    else:
        os.environ["SSE_KERNELS"] = best.family
    if best.variant is None:
        os.environ.pop("SSE_TENSOR_VARIANT", None)
    else:
        os.environ["SSE_TENSOR_VARIANT"] = str(best.variant)
