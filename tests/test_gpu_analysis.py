"""Error and conservation analyses on the GPU (sse_l2_error, sse_functional over the
written time steps of a results directory), run the way the reference's
test/test_driver.jl:1-110 runs them: solve with the save callback, then
``analyze(ErrorAnalysis, u_final, exact, t)``, ``analyze(PrimaryConservationAnalysis,
time_steps)`` and ``analyze(EnergyConservationAnalysis, time_steps)``, checked
against the reference's golden values and tolerances (test/runtests.jl:12-61).
"""

import numpy as np
import pytest

from cases import (advection_2d_modal_tensor_tri, advection_diffusion_1d_modal_multi,
                   density_wave_3d_config, l2_error)

pytestmark = pytest.mark.gpu

TOL = 1.0e-10  # test/runtests.jl:12


def _driver(make, tmp_path, T=1.0, dt=1.0 / 100.0):
    """test_driver.jl: solve with save_callback(results_path, (0, T), floor(1/(dt 50))),
    then the three analyses; returns (l2, max |dE/dt| mass, max |dE/dt| energy)."""
    from sse import io
    from sse.analysis import (EnergyConservationAnalysis, ErrorAnalysis,
                              PrimaryConservationAnalysis, analyze)
    from sse.time_integration import CarpenterKennedy2N54, solve
    law, sd, exact, form, ode = make()
    ode.tspan = (0.0, T)
    path = io.save_project(law, sd, exact, form, ode.tspan, str(tmp_path / "res"))
    interval = int(np.floor(1.0 / (dt * 50)))
    sol = solve(ode, CarpenterKennedy2N54(), dt=dt, callback=io.SaveCallback(path, interval=interval))
    solver = ode.p
    err = analyze(ErrorAnalysis(solver, path), sol.u, exact, 1.0)
    steps = io.load_time_steps(path)
    cons = analyze(PrimaryConservationAnalysis(solver, path), steps)
    energy = analyze(EnergyConservationAnalysis(solver, path), steps)
    assert cons.E.shape == (len(steps), law.N_c) and cons.dEdt.shape == cons.E.shape
    return err, float(np.abs(cons.dEdt[:, 0]).max()), float(np.abs(energy.dEdt[:, 0]).max()), sd, sol, exact


def test_reference_driver_advection_diffusion_1d(tmp_path):
    # test/runtests.jl:13-36
    l2, cons, _, sd, sol, exact = _driver(advection_diffusion_1d_modal_multi, tmp_path)
    np.testing.assert_allclose(l2, [6.988216111882884e-6], rtol=0, atol=TOL)
    assert cons == pytest.approx(0.0, abs=TOL)


def test_reference_driver_advection_2d_energy_conservative_tri(tmp_path):
    # test/runtests.jl:38-61
    l2, cons, energy, sd, sol, exact = _driver(advection_2d_modal_tensor_tri, tmp_path)
    np.testing.assert_allclose(l2, [0.2660013939427627], rtol=0, atol=TOL)
    assert cons == pytest.approx(0.0, abs=TOL)
    assert energy == pytest.approx(0.0, abs=TOL)


@pytest.mark.parametrize("normalize", [False, True])
def test_l2_error_matches_the_quadrature_sum(normalize):
    """sse_l2_error on a five-variable curved-tet state (device and host memory)
    against the volume-quadrature sum of error.jl:58-91 written out in NumPy."""
    import torch
    from sse.analysis import ErrorAnalysis, analyze
    from sse.time_integration import to_device
    law, sd, ic, form, ode = density_wave_3d_config(M=2, p=4)
    rng = np.random.default_rng(7)
    u = np.asfortranarray(ode.u0 + 1e-3 * rng.standard_normal(ode.u0.shape))
    want = l2_error(sd, u, ic, 0.0)
    if normalize:
        ref = sd.reference_approximation
        vol = float((ref.W[:, None] * sd.geometric_factors.J_q).sum())
        want = want / np.sqrt(vol)
    ea = ErrorAnalysis(ode.p)
    got_host = analyze(ea, u, ic, 0.0, normalize=normalize)
    got_dev = analyze(ea, to_device(u), ic, 0.0, normalize=normalize)
    torch.cuda.synchronize()
    assert got_host.shape == (5,)
    np.testing.assert_allclose(got_host, want, rtol=1e-12, atol=0)
    np.testing.assert_array_equal(got_dev, got_host)


def test_conservation_analyze_initial_and_final(tmp_path):
    # conservation.jl:221-234: (initial, final, difference) between two written steps
    from sse import io
    from sse.analysis import PrimaryConservationAnalysis, analyze
    from sse.time_integration import CarpenterKennedy2N54, solve
    law, sd, ic, form, ode = density_wave_3d_config(M=2, p=3, ic_strength=0.2)
    ode.tspan = (0.0, 0.004)
    path = io.save_project(law, sd, ic, form, ode.tspan, str(tmp_path / "res"))
    sol = solve(ode, CarpenterKennedy2N54(), dt=1e-3, callback=io.SaveCallback(path, interval=2))
    io.save_solution(sol.u, 0.004, path, "final")
    a = PrimaryConservationAnalysis(ode.p, path)
    initial, final, diff = analyze(a, 0, "final")
    assert initial.shape == (5,)
    assert np.abs(diff).max() <= 1e-12 * np.abs(initial).max()  # mass is conserved on a periodic mesh
    res = analyze(a, list(io.load_time_steps(path)), normalize=True)
    np.testing.assert_allclose(res.E[0], 1.0, rtol=1e-15)
    np.testing.assert_allclose(res.t, [0.0, 0.002, 0.004], rtol=0, atol=1e-15)
