"""Results directories of a run, laid out as the reference's ``File`` module lays
them out (``src/File/save.jl``, ``src/File/load.jl``):

    <results_path>/project.json      the problem (``project.jld2``)
    <results_path>/time_steps.npy    steps written so far (``time_steps.jld2``)
    <results_path>/sol_<n>.npz       u (N_p, N_c, N_e), t, and du = R(u) (``sol_<n>.jld2``)
    <results_path>/screen.txt        one line per written step

The containers are NumPy's (``.npz``/``.npy``, read with ``allow_pickle=False``) and
JSON instead of JLD2: no HDF5 library is available to this package. Array shapes,
memory order (column-major ``u``), the file naming and the step bookkeeping are
the reference's, so a converter to JLD2 is a per-file copy.
"""

from __future__ import annotations

import json
import os
import shutil
from typing import List, Optional, Tuple, Union

import numpy as np


def new_path(results_path: str, overwrite: bool = False, clear: bool = False) -> str:
    """(``save.jl:1-20``): the directory itself if new (or if `overwrite`, emptied
    first when `clear`), else the first free ``<path>_<n>/``."""
    if not os.path.isdir(results_path):
        path = results_path
    elif overwrite:
        if clear:
            shutil.rmtree(results_path, ignore_errors=True)
        path = results_path
    else:
        suffix = 1
        while True:
            path = results_path.rstrip("/") + f"_{suffix}/"
            if not os.path.isdir(path):
                break
            suffix += 1
    os.makedirs(path, exist_ok=True)
    return path if path.endswith("/") else path + "/"


def _describe(obj) -> dict:
    """A JSON record of a law, form, initial data or discretisation: class name and
    plain fields (numbers, strings, tuples of them)."""
    out = {"type": type(obj).__name__}
    for k, v in getattr(obj, "__dict__", {}).items():
        if isinstance(v, (int, float, str, bool)) or v is None:
            out[k] = v
        elif isinstance(v, (tuple, list)) and all(isinstance(x, (int, float)) for x in v):
            out[k] = list(v)
        elif hasattr(v, "__dict__") and not isinstance(v, np.ndarray):
            out[k] = {"type": type(v).__name__}
    return out


def save_project(conservation_law, spatial_discretization, initial_data, form,
                 tspan: Tuple[float, float], results_path: str, overwrite: bool = False,
                 clear: bool = False) -> str:
    """(``save.jl:22-38``): writes the problem record and an empty step list;
    returns the directory used."""
    path = new_path(results_path, overwrite, clear)
    ref = spatial_discretization.reference_approximation
    record = {
        "conservation_law": _describe(conservation_law),
        "spatial_discretization": {"N_e": int(spatial_discretization.N_e), "N_p": int(ref.N_p),
                                   "N_q": int(ref.N_q), "N_f": int(ref.N_f), "d": int(ref.d),
                                   "element_type": str(ref.element_type)},
        "initial_data": _describe(initial_data),
        "form": _describe(form),
        "tspan": [float(tspan[0]), float(tspan[1])],
        "N_c": int(conservation_law.N_c),
    }
    with open(os.path.join(path, "project.json"), "w") as f:
        json.dump(record, f, indent=1)
    np.save(os.path.join(path, "time_steps.npy"), np.zeros(0, dtype=np.int64))
    return path


def load_project(results_path: str) -> dict:
    """(``load.jl:12-19``): the problem record written by `save_project`."""
    with open(os.path.join(results_path, "project.json")) as f:
        return json.load(f)


def load_time_steps(results_path: str) -> np.ndarray:
    """(``load.jl:29-31``)."""
    return np.load(os.path.join(results_path, "time_steps.npy"), allow_pickle=False)


def save_solution(u: np.ndarray, t: float, results_path: str, time_step: Union[int, str] = 0,
                  du: Optional[np.ndarray] = None) -> None:
    """(``save.jl:40-72``): ``sol_<time_step>.npz`` with u (and du when given); an
    integer step is appended to the step list (a step already written is kept)."""
    name = os.path.join(results_path, f"sol_{time_step}.npz")
    if os.path.isfile(name) and isinstance(time_step, int):
        return
    arrays = {"u": np.asfortranarray(u, dtype=np.float64), "t": np.float64(t)}
    if du is not None:
        arrays["du"] = np.asfortranarray(du, dtype=np.float64)
    np.savez(name, **arrays)
    with open(os.path.join(results_path, "screen.txt"), "a") as f:
        f.write(f"writing time step {time_step}  t = {t}\n")
    if isinstance(time_step, int):
        steps = load_time_steps(results_path)
        np.save(os.path.join(results_path, "time_steps.npy"), np.append(steps, np.int64(time_step)))


def load_solution(results_path: str, time_step: Union[int, str] = 0, load_du: bool = False):
    """(``load.jl:1-10``): (u, t), or (u, du, t) with `load_du`."""
    with np.load(os.path.join(results_path, f"sol_{time_step}.npz"), allow_pickle=False) as d:
        u = np.asfortranarray(d["u"])
        t = float(d["t"])
        if load_du:
            return u, np.asfortranarray(d["du"]), t
        return u, t


def load_snapshots(results_path: str, time_steps: List[int]):
    """(``load.jl:33-50``): the snapshot matrix (N_p N_c N_e, N_t), columns vec(u) in
    column-major order, and the mean step between snapshots."""
    cols, times = [], []
    for s in time_steps:
        u, t = load_solution(results_path, s)
        cols.append(u.reshape(-1, order="F"))
        times.append(t)
    X = np.stack(cols, axis=1)
    t_s = (times[-1] - times[0]) / (len(times) - 1.0) if len(times) > 1 else 0.0
    return X, t_s


def load_snapshots_with_derivatives(results_path: str, time_steps: List[int]):
    """(``load.jl:51-69``): (U, dU, t_s)."""
    U, dU, times = [], [], []
    for s in time_steps:
        u, du, t = load_solution(results_path, s, load_du=True)
        U.append(u.reshape(-1, order="F"))
        dU.append(du.reshape(-1, order="F"))
        times.append(t)
    t_s = (times[-1] - times[0]) / (len(times) - 1.0) if len(times) > 1 else 0.0
    return np.stack(U, axis=1), np.stack(dU, axis=1), t_s


class SaveCallback:
    """(``save.jl:74-86``): write the solution and its time derivative every
    `interval` steps and at both ends of tspan. Pass to ``solve(..., callback=)``."""

    def __init__(self, results_path: str, interval: int = 1, restart_step: int = 0):
        self.results_path = results_path
        self.interval = int(interval)
        self.restart_step = int(restart_step)

    def __call__(self, solver, u_dev, t: float, step: int) -> None:
        import torch

        from .solvers import semi_discrete_residual
        from .time_integration import to_host
        du = torch.empty_like(u_dev)
        semi_discrete_residual(du, u_dev, solver, t)
        save_solution(to_host(u_dev), t, self.results_path, step + self.restart_step, to_host(du))
