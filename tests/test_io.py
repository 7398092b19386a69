"""Results directories (the reference's src/File/save.jl and load.jl layout)."""

import os

import numpy as np

from sse import io


def test_new_path_suffixes_and_clear(tmp_path):
    base = str(tmp_path / "run")
    p0 = io.new_path(base)
    assert os.path.isdir(p0)
    p1 = io.new_path(base)
    assert p1.rstrip("/").endswith("run_1") and os.path.isdir(p1)
    open(os.path.join(p0, "junk"), "w").close()
    p2 = io.new_path(base, overwrite=True, clear=True)
    assert p2.rstrip("/") == base and os.listdir(p2) == []


def test_solution_round_trip_steps_and_snapshots(tmp_path):
    from cases import vortex_2d_config
    law, sd, ic, form, ode = vortex_2d_config(M=2, p=2)
    path = io.save_project(law, sd, ic, form, (0.0, 1.0), str(tmp_path / "res"))
    rec = io.load_project(path)
    assert rec["tspan"] == [0.0, 1.0] and rec["spatial_discretization"]["N_e"] == sd.N_e
    assert rec["conservation_law"]["type"] == "EulerEquations"
    rng = np.random.default_rng(3)
    us, dus = [], []
    for step, t in ((0, 0.0), (5, 0.5), (10, 1.0)):
        u = np.asfortranarray(rng.standard_normal(ode.u0.shape))
        du = np.asfortranarray(rng.standard_normal(ode.u0.shape))
        io.save_solution(u, t, path, step, du)
        us.append(u)
        dus.append(du)
    io.save_solution(us[0] * 2, 0.0, path, 0)  # an existing step is kept
    assert list(io.load_time_steps(path)) == [0, 5, 10]
    u, t = io.load_solution(path, 5)
    assert t == 0.5 and u.flags.f_contiguous and np.array_equal(u, us[1])
    u, du, t = io.load_solution(path, 10, load_du=True)
    assert np.array_equal(du, dus[2])
    X, t_s = io.load_snapshots(path, [0, 5, 10])
    assert X.shape == (us[0].size, 3) and t_s == 0.5
    assert np.array_equal(X[:, 1], us[1].reshape(-1, order="F"))
    U, dU, t_s = io.load_snapshots_with_derivatives(path, [0, 10])
    assert np.array_equal(dU[:, 1], dus[2].reshape(-1, order="F")) and t_s == 1.0
    assert "writing time step 10" in open(os.path.join(path, "screen.txt")).read()
