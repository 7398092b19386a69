"""HIP residual (libsse_hip.so, through its C-ABI) against the CPU oracle.

Bar: max-norm error relative to max |dudt| <= 1e-12 for one residual; <= 1e-10
after 100 CK2N54 steps; the reference's own golden L2 errors reproduced on the
GPU at the reference's tolerance (atol 1e-10, test/runtests.jl:12).
"""

import math

import numpy as np
import pytest

from cases import (advection_2d_modal_tensor_tri, advection_3d_modal_tensor_tet,
                   advection_diffusion_1d_modal_multi, advection_diffusion_3d_config,
                   density_wave_3d_config, euler_1d_gauss, euler_vortex_2d_modal, l2_error,
                   vortex_2d_config)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RES_TOL = 1e-12
RK_TOL = 1e-10


def _gpu():
    import torch
    assert torch.cuda.is_available()
    return torch


def _gpu_residual(ode, u=None):
    torch = _gpu()
    from sse import semi_discrete_residual
    from sse.time_integration import to_device, to_host
    u = ode.u0 if u is None else u
    ud = to_device(u)
    dd = torch.empty_like(ud)
    semi_discrete_residual(dd, ud, ode.p)
    torch.cuda.synchronize()
    return to_host(dd)


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _perturbed(ode, amp=1e-3, seed=20261202):
    """Initial state plus a small seeded perturbation (exercises every mode)."""
    rng = np.random.default_rng(seed)
    return np.asfortranarray(ode.u0 + amp * rng.standard_normal(ode.u0.shape) * np.abs(ode.u0).max())


CASES = {
    "euler1d_gauss_p5": lambda: euler_1d_gauss()[4],
    "vortex2d_modal_p3_warped": lambda: euler_vortex_2d_modal()[4],
    "vortex2d_tri_p4_8x8": lambda: vortex_2d_config(M=8, p=4)[4],
    "vortex2d_tri_p2": lambda: vortex_2d_config(M=4, p=2)[4],
    "vortex2d_tri_p5": lambda: vortex_2d_config(M=3, p=5)[4],
    "vortex2d_tri_p6": lambda: vortex_2d_config(M=3, p=6)[4],
    "wave3d_tet_p2_curved": lambda: density_wave_3d_config(M=2, p=2, ic_strength=0.2)[4],
    "wave3d_tet_p4_curved": lambda: density_wave_3d_config(M=2, p=4)[4],
    "wave3d_tet_p3_curved": lambda: density_wave_3d_config(M=3, p=3, ic_strength=0.2)[4],
    "wave3d_tet_p5_curved": lambda: density_wave_3d_config(M=2, p=5, ic_strength=0.2)[4],
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_residual_matches_oracle(name):
    ode = CASES[name]()
    u = _perturbed(ode)
    got = _gpu_residual(ode, u)
    ref = O.residual(ode.p.arrays, u)
    assert _rel(got, ref) <= RES_TOL, (name, _rel(got, ref))
    # Tri/Tet ModalTensor Euler runs the line-structured compile-time-shaped kernels
    expected = "generic" if name.startswith("euler1d") else "struct"
    assert ode.p.describe().startswith(expected), ode.p.describe()


def test_residual_host_memory_path_equals_device_path():
    # equal up to the order of K2's LDS atomic sums (the partner shares of the volume
    # pairs and the facet-correction row and column sums)
    from sse import semi_discrete_residual
    ode = vortex_2d_config(M=4, p=4)[4]
    dev = _gpu_residual(ode)
    host = np.zeros_like(ode.u0, order="F")
    semi_discrete_residual(host, ode.u0, ode.p)
    assert _rel(host, dev) <= 1e-14


def test_repeated_residuals_agree_to_rounding():
    # K2 accumulates the partner shares of the volume pairs and the facet-correction
    # sums with LDS atomic adds, so their order may change between calls; everything
    # else is fixed
    ode = density_wave_3d_config(M=2, p=4)[4]
    first = _gpu_residual(ode)
    for _ in range(3):
        assert _rel(_gpu_residual(ode), first) <= 1e-14


def test_rk_100_steps_config1_matches_oracle():
    from sse.time_integration import rk_steps, to_device, to_host
    law, sd, ic, form, ode = vortex_2d_config(M=8, p=4)
    dt = 0.01
    ud = to_device(ode.u0)
    rk_steps(ode.p, ud, dt, 100)
    got = to_host(ud)
    ref = O.rk_steps(ode.p.arrays, ode.u0, dt, 100)
    assert _rel(got, ref) <= RK_TOL


def test_gpu_reproduces_reference_golden_euler_1d():
    # test/runtests.jl:89-96, whole trajectory on the GPU
    from sse.time_integration import CarpenterKennedy2N54, solve
    law, sd, exact, form, ode = euler_1d_gauss()
    sol = solve(ode, CarpenterKennedy2N54(), dt=2.0 / 1000)
    err = l2_error(sd, sol.u, exact, 2.0)
    gold = [3.5808560177567635e-5, 5.2129828619609155e-5, 0.00012637647535378534]
    np.testing.assert_allclose(err, gold, rtol=0, atol=1e-10)


def test_gpu_reproduces_reference_golden_vortex_modal_tri():
    # test/runtests.jl:111-121, whole trajectory on the GPU
    from sse.time_integration import CarpenterKennedy2N54, solve
    law, sd, exact, form, ode, T = euler_vortex_2d_modal()
    sol = solve(ode, CarpenterKennedy2N54(), dt=T / 1000)
    err = l2_error(sd, sol.u, exact, T)
    gold = [0.015568197027072704, 0.040539693811761104, 0.04060141777050208, 0.043960971468832745]
    np.testing.assert_allclose(err, gold, rtol=0, atol=1e-10)


def test_nonphysical_state_is_reported():
    from sse import semi_discrete_residual
    from sse.hip import SseError
    ode = vortex_2d_config(M=2, p=2)[4]
    u = np.array(ode.u0, order="F", copy=True)
    u[:, 3, 0] = -10.0  # negative energy -> negative pressure in element 0
    out = np.zeros_like(u, order="F")
    with pytest.raises(SseError) as ei:
        semi_discrete_residual(out, u, ode.p)
    assert ei.value.code == 4
    # the handle stays usable
    semi_discrete_residual(out, ode.u0, ode.p)


@pytest.mark.parametrize("name", ["vortex2d_tri_p4_8x8", "wave3d_tet_p4_curved", "wave3d_tet_p3_curved"])
def test_kernel_families_agree(name, monkeypatch):
    """The three kernel families (line-structured, table-driven tensor, and the
    runtime-shaped generic kernels; chosen with SSE_KERNELS at handle creation)
    agree with each other and with the oracle."""
    u = _perturbed(CASES[name]())
    ref = None
    got = {}
    for fam in ("struct", "tensor", "generic"):
        if fam == "struct":
            monkeypatch.delenv("SSE_KERNELS", raising=False)
        else:
            monkeypatch.setenv("SSE_KERNELS", fam)
        ode = CASES[name]()
        got[fam] = _gpu_residual(ode, u)
        assert ode.p.describe().startswith(fam), ode.p.describe()
        if ref is None:
            ref = O.residual(ode.p.arrays, u)
        assert _rel(got[fam], ref) <= RES_TOL, (fam, _rel(got[fam], ref))
    assert _rel(got["struct"], got["tensor"]) <= RES_TOL
    assert _rel(got["struct"], got["generic"]) <= RES_TOL


# ----------------------------------------------------------------------------- full benchmark size
# 3D Euler, curved tets, ModalTensor p=4, 16^3 hexes x 6 tets (24,576 elements):
# one residual and 20 CK2N54 steps against the oracle, and the size-independent
# invariants (conservation, entropy conservation with the EC interface flux,
# free-stream preservation) on the GPU residual itself.
@pytest.fixture(scope="module")
def full_lf():
    return density_wave_3d_config(M=16, p=4, factor=0.05, ic_strength=0.5)


def test_full_size_residual_and_20_rk_steps_match_oracle(full_lf):
    from sse.time_integration import rk_steps, to_device, to_host
    law, sd, ic, form, ode = full_lf
    O.set_threads(min(16, O.lib().sse_oracle_max_threads()))
    got = _gpu_residual(ode)
    ref = O.residual(ode.p.arrays, ode.u0)
    assert _rel(got, ref) <= RES_TOL, _rel(got, ref)
    dt = 0.1 * (2.0 / 16) / (4 + 1) ** 2
    ud = to_device(ode.u0)
    rk_steps(ode.p, ud, dt, 20)
    got = to_host(ud)
    ref = O.rk_steps(ode.p.arrays, ode.u0, dt, 20)
    assert _rel(got, ref) <= RK_TOL, _rel(got, ref)


def _entropy_rate(sd, law, u, dudt):
    # src/Analysis/conservation.jl:169-189: with P = M_k^{-1} V' WJ_k the rate
    # (P w_q)' M_k dudt reduces to w_q' WJ_k V dudt
    from sse.conservation_laws import conservative_to_entropy
    ref = sd.reference_approximation
    V = ref.Vdense()
    WJ = ref.W[:, None] * sd.geometric_factors.J_q
    wq = conservative_to_entropy(law, np.einsum("qp,pck->qkc", V, u))
    terms = WJ[:, :, None] * wq * np.einsum("qp,pck->qkc", V, dudt)
    return terms.sum(), np.abs(terms).sum()


def test_full_size_conservation_and_free_stream(full_lf):
    law, sd, ic, form, ode = full_lf
    ref = sd.reference_approximation
    V = ref.Vdense()
    WJ = ref.W[:, None] * sd.geometric_factors.J_q
    dudt = _gpu_residual(ode)
    rate = np.einsum("qk,qp,pck->c", WJ, V, dudt)
    scale = np.einsum("qk,qp,pck->c", WJ, np.abs(V), np.abs(dudt))
    assert np.all(np.abs(rate) <= 1e-13 * scale), rate / scale
    rho, v, p = 1.2, np.array([0.3, -0.2, 0.5]), 0.9
    ucons = np.array([rho, *(rho * v), p / 0.4 + 0.5 * rho * v @ v])
    coef0 = np.linalg.lstsq(V, np.ones(ref.N_q), rcond=None)[0]
    u = np.asfortranarray(coef0[:, None, None] * ucons[None, :, None] * np.ones((1, 1, sd.N_e)))
    assert np.abs(_gpu_residual(ode, u)).max() < 1e-11


def test_full_size_entropy_conservation_with_ec_interface_flux():
    from sse.conservation_laws import EntropyConservativeNumericalFlux
    from sse.solvers import FluxDifferencingForm
    form = FluxDifferencingForm(inviscid_numerical_flux=EntropyConservativeNumericalFlux())
    law, sd, ic, form, ode = density_wave_3d_config(M=16, p=4, factor=0.05, form=form, ic_strength=0.5)
    u = _perturbed(ode)
    rate, scale = _entropy_rate(sd, law, u, _gpu_residual(ode, u))
    assert abs(rate) <= 1e-12 * scale, (rate, scale)



# ----------------------------------------------------------------------------- StandardForm
def _euler_tri_standard():
    from sse.solvers import SkewSymmetricMapping, StandardForm
    return vortex_2d_config(M=4, p=3, form=StandardForm(mapping_form=SkewSymmetricMapping()))[4]


def _advdiff_standard_mapping():
    from sse.conservation_laws import BR1, LaxFriedrichsNumericalFlux, LinearAdvectionDiffusionEquation
    from sse.solvers import PhysicalOperator, StandardForm, StandardMapping, semidiscretize
    law, sd, ex, form, ode = advection_diffusion_3d_config(M=2, p=3, b=2e-2, sigma=0.2)
    form = StandardForm(mapping_form=StandardMapping(), inviscid_numerical_flux=LaxFriedrichsNumericalFlux(),
                        viscous_numerical_flux=BR1())
    return semidiscretize(law, sd, ex.initial_data, form, (0.0, 1.0), PhysicalOperator())


STD_CASES = {
    "advdiff1d_modal_multi_br1": lambda: advection_diffusion_1d_modal_multi()[4],
    "adv2d_tri_skew_reference_operator": lambda: advection_2d_modal_tensor_tri()[4],
    "adv3d_tet_skew_central": lambda: advection_3d_modal_tensor_tet()[4],
    "advdiff3d_tet_p5_br1": lambda: advection_diffusion_3d_config(M=2, p=5)[4],
    "advdiff3d_tet_p3_standard_mapping": _advdiff_standard_mapping,
    "euler2d_tri_standard_skew": _euler_tri_standard,
    # every other degree the compile-time-shaped scalar kernels are built for
    "adv2d_tri_p2": lambda: advection_2d_modal_tensor_tri(M=3, p=2)[4],
    "adv2d_tri_p3": lambda: advection_2d_modal_tensor_tri(M=3, p=3)[4],
    "adv2d_tri_p5": lambda: advection_2d_modal_tensor_tri(M=2, p=5)[4],
    "adv2d_tri_p6": lambda: advection_2d_modal_tensor_tri(M=2, p=6)[4],
    "advdiff3d_tet_p2_br1": lambda: advection_diffusion_3d_config(M=2, p=2)[4],
    "advdiff3d_tet_p3_br1": lambda: advection_diffusion_3d_config(M=2, p=3)[4],
    "advdiff3d_tet_p4_br1": lambda: advection_diffusion_3d_config(M=2, p=4)[4],
}


@pytest.mark.parametrize("name", sorted(STD_CASES))
def test_standard_form_residual_matches_oracle(name):
    ode = STD_CASES[name]()
    u = _perturbed(ode, amp=1e-3 if name.startswith("euler") else 1e-2)  # Euler states stay physical
    got = _gpu_residual(ode, u)
    ref = O.residual(ode.p.arrays, u)
    assert _rel(got, ref) <= RES_TOL, (name, _rel(got, ref))
    assert ode.p.describe().startswith("standard"), ode.p.describe()


def test_standard_form_br1_rk_steps_match_oracle():
    from sse.time_integration import rk_steps, to_device, to_host
    ode = advection_diffusion_3d_config(M=2, p=4)[4]
    dt = 1e-3
    ud = to_device(ode.u0)
    rk_steps(ode.p, ud, dt, 50)
    got = to_host(ud)
    ref = O.rk_steps(ode.p.arrays, ode.u0, dt, 50)
    assert _rel(got, ref) <= RK_TOL, _rel(got, ref)


def _advdiff_ensemble(p, bs, M=2):
    from sse.conservation_laws import LinearAdvectionDiffusionEquation
    law, sd, ex, form, ode = advection_diffusion_3d_config(M=M, p=p)
    solvers = [ode.p.with_law(LinearAdvectionDiffusionEquation(law.a, b)) for b in bs]
    u0s = [np.asfortranarray(ode.u0 * (1.0 + 0.25 * i)) for i in range(len(bs))]
    return solvers, u0s


@pytest.mark.parametrize("p,chunk", [(3, None), (5, None), (4, "2")])
def test_diffusivity_ensemble_matches_oracle(p, chunk, monkeypatch):
    # a sweep of diffusion coefficients on one mesh runs as one ensemble group
    # (one geometry copy, members side by side); chunk 2 splits it over launches
    from sse.batch import BatchedSolver
    from sse.time_integration import to_device, to_host
    if chunk:
        monkeypatch.setenv("SSE_ENSEMBLE_CHUNK", chunk)
    bs = (2e-3, 5e-3, 1e-2, 4e-3, 7e-3)
    solvers, u0s = _advdiff_ensemble(p, bs)
    batch = BatchedSolver(solvers)
    assert batch.describe().count("one mesh") == 1, batch.describe()
    dt, n = 1e-3, 10
    states = [to_device(u) for u in u0s]
    batch.rk_steps(states, [dt] * len(bs), n)
    for s, u0, x in zip(solvers, u0s, states):
        ref = O.rk_steps(s.arrays, u0, dt, n)
        assert _rel(to_host(x), ref) <= RK_TOL, (s.conservation_law.b, _rel(to_host(x), ref))


@pytest.mark.parametrize("chunk", [None, "2"])
def test_diffusivity_tangent_matches_oracle_forward_mode(chunk, monkeypatch):
    # (u, du/db) pairs advanced together against forward-mode duals through the
    # oracle's RK; the energy gradient dJ/db = u' M du/db against the oracle's too
    import torch
    from oracle import functionals as F
    from sse.analysis import EnergyConservationAnalysis
    from sse.batch import BatchedSolver
    from sse.time_integration import to_device, to_host
    if chunk:
        monkeypatch.setenv("SSE_ENSEMBLE_CHUNK", chunk)
    bs = (3e-3, 1e-2, 6e-3)
    solvers, u0s = _advdiff_ensemble(4, bs)
    batch = BatchedSolver(solvers)
    dt, n = 1e-3, 10
    states = [to_device(u) for u in u0s]
    tangents = [torch.zeros_like(x) for x in states]
    batch.rk_tangent(states, tangents, [dt] * len(bs), n)
    for s, u0, x, v in zip(solvers, u0s, states, tangents):
        uN, duN = O.rk_jvp(s.arrays, u0, np.zeros_like(u0), dt, n, db=1.0)
        assert _rel(to_host(x), uN) <= RK_TOL
        assert _rel(to_host(v), duN) <= RK_TOL, _rel(to_host(v), duN)
        g = F.energy_rate(s.spatial_discretization, uN, duN)[0]
        got = EnergyConservationAnalysis(s).evaluate_conservation_residual(x, v)[0]
        assert abs(got - g) <= 1e-10 * abs(g), (got, g)


def test_gpu_reproduces_reference_golden_advection_diffusion_1d():
    # test/runtests.jl:13-36, whole trajectory on the GPU
    from sse.time_integration import CarpenterKennedy2N54, solve
    law, sd, exact, form, ode = advection_diffusion_1d_modal_multi()
    sol = solve(ode, CarpenterKennedy2N54(), dt=1.0 / 100)
    err = l2_error(sd, sol.u, exact, 1.0)
    np.testing.assert_allclose(err, [6.988216111882884e-6], rtol=0, atol=1e-10)


def test_gpu_reproduces_reference_golden_advection_2d_tri():
    # test/runtests.jl:38-61, whole trajectory on the GPU
    from sse.time_integration import CarpenterKennedy2N54, solve
    law, sd, exact, form, ode = advection_2d_modal_tensor_tri()
    sol = solve(ode, CarpenterKennedy2N54(), dt=1.0 / 100)
    err = l2_error(sd, sol.u, exact, 1.0)
    np.testing.assert_allclose(err, [0.2660013939427627], rtol=0, atol=1e-10)


def test_gpu_reproduces_reference_golden_euler_3d_nodal_tensor_hex(tmp_path):
    # test/runtests.jl:131-144 with the flow of test/euler_3d.jl: NodalTensor(4) Hex
    # (generic kernels: sparse S, diagonal E, EC interface flux), DP8 at dt = T / 250
    # (sse_erk_steps) with save_callback every 5 steps, then the three analyses
    from sse import io
    from sse.analysis import (EntropyConservationAnalysis, ErrorAnalysis, PrimaryConservationAnalysis,
                              analyze)
    from sse.time_integration import DP8, solve
    from sse.workloads import euler_3d_hex
    law, sd, exact, form, ode, T, N_t = euler_3d_hex()
    assert ode.p.describe().startswith("generic"), ode.p.describe()
    path = io.save_project(law, sd, exact, form, ode.tspan, str(tmp_path / "res"))
    sol = solve(ode, DP8(), dt=T / N_t, callback=io.SaveCallback(path, interval=N_t // 50))
    err = analyze(ErrorAnalysis(ode.p, path), sol.u, exact, T)
    gold = [0.18342164491797003, 0.1834216449179776, 0.18342164491796725, 0.18342164491796784,
            0.2751324673769553]
    np.testing.assert_allclose(err, gold, rtol=0, atol=1e-10)
    steps = io.load_time_steps(path)
    assert len(steps) == 51
    cons = analyze(PrimaryConservationAnalysis(ode.p, path), steps)
    np.testing.assert_allclose(cons.E[-1] - cons.E[0], 0.0, atol=1e-10)
    ent = analyze(EntropyConservationAnalysis(ode.p, path), steps)
    assert np.abs(ent.dEdt[:, 0]).max() < 1e-10


@pytest.mark.parametrize("d", [1, 2, 3])
def test_device_ec_flux_matches_oracle_flux(d):
    # the kernels' two-point flux (hoisted logarithms, one refined reciprocal instead of a
    # divide) against the oracle's Ranocha flux (the reference's formulas), contracted with
    # a random vector, pair by pair: distinct states, near-equal ones (the series branch of
    # the logarithmic mean) and equal ones
    import ctypes as C
    from sse import abi
    from sse.hip import check, load_library
    lib = load_library()
    rng = np.random.default_rng(31 + d)
    n, gamma = 3000, 1.4
    nc = d + 2

    def states(m):
        rho = rng.uniform(0.3, 3.0, m)
        v = rng.uniform(-1.5, 1.5, (m, d))
        p = rng.uniform(0.3, 3.0, m)
        E = p / (gamma - 1.0) + 0.5 * rho * (v ** 2).sum(1)
        return np.column_stack([rho, rho[:, None] * v, E])

    uL = states(n)
    uR = states(n)
    uR[1000:2000] = uL[1000:2000] * (1.0 + 1e-5 * rng.standard_normal((1000, nc)))
    uR[2000:2200] = uL[2000:2200]
    a = rng.standard_normal((n, d))
    c = np.zeros((n, nc))
    dp = lambda x: np.ascontiguousarray(x).ctypes.data
    uL, uR, a = (np.ascontiguousarray(x) for x in (uL, uR, a))
    check(lib, None, lib.sse_ec_flux_batch(d, gamma, n, dp(uL), dp(uR), dp(a), c.ctypes.data, abi.SSE_HOST))
    worst = 0.0
    for t in range(n):
        F = O.two_point(1, d, gamma, uL[t], uR[t])   # N_c x d
        ref = F @ a[t]
        err = np.abs(c[t] - ref).max() / np.abs(ref).max()
        worst = max(worst, err)
    assert worst <= 2e-14, worst


def test_non_antisymmetric_s_leaves_the_pair_once_kernels():
    # the line-structured kernels give a pair's partner -S_m[i, j]; a descriptor whose S
    # is not antisymmetric runs on kernels that, like the reference's flux_difference!
    # (src/Solvers/flux_differencing_form.jl:37-75), read only the entries with i < j
    ode = density_wave_3d_config(M=2, p=3, ic_strength=0.2)[4]
    S = ode.p.arrays.S                      # S[m, j, i] = S_m[i, j]
    d, n = S.shape[0], S.shape[1]
    low = np.tril(np.ones((n, n), dtype=bool), -1)     # i > j, stored at S[m, j, i]
    for m in range(d):
        Sm = S[m].T.copy()                  # S_m[i, j]
        Sm[low] *= 1.0 + 1e-3
        S[m] = Sm.T
    assert not ode.p.describe().startswith("struct"), ode.p.describe()
    got = _gpu_residual(ode)
    ref = O.residual(ode.p.arrays, ode.u0)
    assert _rel(got, ref) <= RES_TOL, _rel(got, ref)


def _modal_multi():
    from sse.reference_approximation import ModalMulti
    return ModalMulti


def _euler_3d_hex(**kw):
    from sse.workloads import euler_3d_hex
    return euler_3d_hex(**kw)


def _deterministic(make):
    from sse.solvers import MI355X, Solver
    law, sd, ic, form, ode = make()[:5]
    solver = Solver(law, sd, form, parallelism=MI355X(deterministic=True))
    assert solver.describe().endswith("deterministic"), solver.describe()
    return law, sd, ode, solver


DET_CASES = {
    "tet_p4": lambda: density_wave_3d_config(M=2, p=4),
    "tet_p3_even_lines": lambda: density_wave_3d_config(M=2, p=3, ic_strength=0.2),
    "tri_p4": lambda: vortex_2d_config(M=4, p=4),
    "tet_p5": lambda: density_wave_3d_config(M=1, p=5, ic_strength=0.2),
    # the general kernel (owner lists in place of its LDS atomics)
    "tet_p3_multi": lambda: density_wave_3d_config(M=2, p=3, ic_strength=0.2, approx=_modal_multi()),
    "euler1d_gauss_p5": euler_1d_gauss,
    "hex_p3": lambda: _euler_3d_hex(M=2, p=3),
}


@pytest.mark.parametrize("name", sorted(DET_CASES))
def test_deterministic_residuals_are_bitwise_equal_and_match_oracle(name):
    # the deterministic mode of the line-structured kernels: repeated residuals are equal
    # bit for bit (the default mode differs in the last bits), and equal the oracle
    torch = _gpu()
    from sse import semi_discrete_residual
    from sse.time_integration import to_device, to_host
    law, sd, ode, solver = _deterministic(DET_CASES[name])
    u = to_device(_perturbed(ode, amp=1e-3))
    outs = []
    for _ in range(4):
        d = torch.empty_like(u)
        semi_discrete_residual(d, u, solver)
        outs.append(d)
    torch.cuda.synchronize()
    for d in outs[1:]:
        assert torch.equal(d, outs[0])
    ref = O.residual(solver.arrays, to_host(u))
    assert _rel(to_host(outs[0]), ref) <= RES_TOL, _rel(to_host(outs[0]), ref)


@pytest.mark.parametrize("name", ["tet_p4", "tet_p3_even_lines", "tri_p4", "tet_p3_multi"])
def test_deterministic_vjp_is_bitwise_repeatable_and_passes_the_dot_product_test(name):
    from sse.adjoint import residual_vjp
    law, sd, ode, solver = _deterministic(DET_CASES[name])
    u = _perturbed(ode, amp=1e-3)
    rng = np.random.default_rng(41)
    ybar = np.asfortranarray(rng.standard_normal(u.shape))
    du = np.asfortranarray(rng.standard_normal(u.shape))
    ub1, lam1 = residual_vjp(solver, u, ybar)
    ub2, lam2 = residual_vjp(solver, u, ybar)
    assert np.array_equal(ub1, ub2) and lam1 == lam2
    dlam = 0.7
    r, dr = O.residual_jvp(solver.arrays, u, du, dlam)
    lhs = float(np.sum(ybar * dr))
    rhs = float(np.sum(ub1 * du)) + lam1 * dlam
    assert abs(lhs - rhs) <= 1e-11 * float(np.sum(np.abs(ybar * dr))), (lhs, rhs)


def test_deterministic_rk_adjoint_is_bitwise_repeatable():
    # with every sum ordered, the checkpointed adjoint's recomputed stage states equal the
    # forward sweep's, and two gradients agree bit for bit
    from sse.adjoint import rk_adjoint
    law, sd, ode, solver = _deterministic(DET_CASES["tet_p3_even_lines"])
    u0 = _perturbed(ode, amp=1e-3)
    J1, g1, l1 = rk_adjoint(solver, u0, 2e-3, 8, checkpoint_every=3)
    J2, g2, l2 = rk_adjoint(solver, u0, 2e-3, 8, checkpoint_every=3)
    assert J1 == J2 and l1 == l2 and np.array_equal(g1, g2)


@pytest.mark.parametrize("nsteps,every", [(8, 3), (9, 2), (4, 2)])
def test_rk_adjoint_recompute_beside_reverse_sweep_equals_one_stream(monkeypatch, nsteps, every):
    # segments recomputed on a second stream while the reverse sweep runs (two buffer sets,
    # reused every other segment) give the gradient of one stream bit for bit; odd and
    # even segment counts, and a count of two
    from sse.adjoint import rk_adjoint
    law, sd, ode, solver = _deterministic(DET_CASES["tet_p3_even_lines"])
    u0 = _perturbed(ode, amp=1e-3)
    J1, g1, l1 = rk_adjoint(solver, u0, 2e-3, nsteps, checkpoint_every=every)
    monkeypatch.setenv("SSE_ADJ_OVERLAP", "0")
    J0, g0, l0 = rk_adjoint(solver, u0, 2e-3, nsteps, checkpoint_every=every)
    assert J1 == J0 and l1 == l0 and np.array_equal(g1, g0)


@pytest.mark.parametrize("make", [lambda: density_wave_3d_config(M=2, p=3, ic_strength=0.2),
                                  lambda: density_wave_3d_config(M=2, p=4, ic_strength=0.2),
                                  lambda: density_wave_3d_config(M=1, p=5, ic_strength=0.2),
                                  lambda: vortex_2d_config(M=4, p=4)],
                         ids=["tet_p3", "tet_p4", "tet_p5", "tri_p4"])
def test_kept_mass_matrices_match_the_transforms_and_the_oracle(monkeypatch, make):
    # each element's kept weight-adjusted inverse V' diag(W/J_k) V (the default) against the
    # sum-factorised V, W/J, V' (SSE_MASS_MATRICES=0): residual and VJP agree to round-off
    # (the default mode's LDS atomics add ~1e-14 of their own), and the residual matches
    # the oracle
    from sse import semi_discrete_residual
    from sse.adjoint import residual_vjp
    from sse.solvers import Solver
    from sse.time_integration import to_device, to_host
    torch = _gpu()
    law, sd, ic, form, ode = make()[:5]
    u = _perturbed(ode)
    rng = np.random.default_rng(5)
    ybar = np.asfortranarray(rng.standard_normal(u.shape))
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("SSE_MASS_MATRICES", flag)
        s = Solver(law, sd, form)
        try:
            ud = to_device(u)
            dd = torch.empty_like(ud)
            semi_discrete_residual(dd, ud, s)
            torch.cuda.synchronize()
            r = to_host(dd)
            ub, lam = residual_vjp(s, u, ybar)
            out[flag] = (r, ub, lam)
        finally:
            s.close()
    ref = O.residual(ode.p.arrays, u)
    assert _rel(out["1"][0], ref) <= RES_TOL
    assert _rel(out["1"][0], out["0"][0]) <= RES_TOL
    assert _rel(out["1"][1], out["0"][1]) <= 1e-11
    assert abs(out["1"][2] - out["0"][2]) <= 1e-11 * max(abs(out["0"][2]), 1e-300)


def test_rk_steps_graph_replay_matches_eager_and_oracle(monkeypatch):
    # sse_rk_steps replays one captured step; a new state pointer or a new dt recaptures
    from sse.time_integration import rk_steps, to_device, to_host
    ode = vortex_2d_config(M=8, p=4)[4]
    dt = 1e-2
    ua, ub = to_device(ode.u0), to_device(ode.u0)
    rk_steps(ode.p, ua, dt, 12)
    rk_steps(ode.p, ub, dt, 12)         # another state: recaptured
    rk_steps(ode.p, ub, 0.5 * dt, 6)    # another step size: recaptured
    rk_steps(ode.p, ua, 0.5 * dt, 6)
    monkeypatch.setenv("SSE_RK_GRAPH", "0")
    ue = to_device(ode.u0)
    rk_steps(ode.p, ue, dt, 12)
    rk_steps(ode.p, ue, 0.5 * dt, 6)
    ref = O.rk_steps(ode.p.arrays, O.rk_steps(ode.p.arrays, ode.u0, dt, 12), 0.5 * dt, 6)
    for x in (ua, ub):
        assert _rel(to_host(x), to_host(ue)) <= 1e-13
        assert _rel(to_host(x), ref) <= RK_TOL


def test_erk_steps_match_oracle_and_host_path():
    # DP8 through sse_erk_steps on a ModalTensor Tet (the line-structured kernels) and a
    # Tri against the oracle's explicit RK, device and host memory
    from sse.time_integration import DP8, erk_steps, to_device, to_host
    alg = DP8()
    for ode, dt in ((density_wave_3d_config(M=2, p=3, ic_strength=0.2)[4], 2e-3),
                    (vortex_2d_config(M=4, p=4)[4], 1e-2)):
        ref = O.erk_steps(ode.p.arrays, ode.u0, dt, 6, alg.A, alg.b)
        ud = to_device(ode.u0)
        erk_steps(ode.p, ud, dt, 6, alg)
        assert _rel(to_host(ud), ref) <= RK_TOL, _rel(to_host(ud), ref)
        import ctypes as C
        from sse import abi
        from sse.hip import check
        uh = np.array(ode.u0, order="F", copy=True)
        dp = C.POINTER(C.c_double)
        lib, h = ode.p.lib, ode.p.handle
        check(lib, h, lib.sse_erk_steps(h, uh.ctypes.data, 0.0, dt, 6, 12, alg.A.ctypes.data_as(dp),
                                        alg.b.ctypes.data_as(dp), abi.SSE_HOST))
        assert _rel(uh, ref) <= RK_TOL


# ----------------------------------------------------------------------------- batched problems
def test_batched_problems_match_individual_solves():
    """Problems of uneven size and degree advanced together (grouped launches,
    HIP-graph RK step replayed across two calls) equal per-problem solves."""
    from sse.batch import BatchedSolver
    from sse.time_integration import rk_steps, to_device, to_host
    from sse.workloads import batched_vortex_problems
    probs = batched_vortex_problems(n=7, seed=5, M_range=(2, 4), p_range=(2, 4))
    batch = BatchedSolver([ode.p for ode, dt, key in probs])
    assert batch.describe().startswith("batch groups=")
    states = [to_device(ode.u0) for ode, dt, key in probs]
    dts = [dt for ode, dt, key in probs]
    batch.rk_steps(states, dts, 12)            # graph built and replayed
    batch.rk_steps(states, dts, 8)             # replayed again in a second call
    eager = [to_device(ode.u0) for ode, dt, key in probs]
    batch.rk_steps(eager, dts, 20, use_graph=False)
    for (ode, dt, key), got, got_eager in zip(probs, states, eager):
        ud = to_device(ode.u0)
        rk_steps(ode.p, ud, dt, 20)
        ref = to_host(ud)
        assert _rel(to_host(got), ref) <= 1e-13, key
        assert _rel(to_host(got_eager), ref) <= 1e-13, key


# ----------------------------------------------------------------------------- functionals, LinearResidual
FUNC_CASES = {
    "wave3d_tet_p3_curved": lambda: density_wave_3d_config(M=2, p=3, ic_strength=0.2),
    "euler1d_gauss_nodal": lambda: euler_1d_gauss(),
    "advdiff3d_tet_p4": lambda: advection_diffusion_3d_config(M=2, p=4),
}


@pytest.mark.parametrize("name", sorted(FUNC_CASES))
def test_functionals_match_oracle(name):
    from oracle import functionals as F
    from sse.analysis import (EnergyConservationAnalysis, EntropyConservationAnalysis,
                              PrimaryConservationAnalysis)
    from sse.time_integration import to_device
    law, sd, ic, form, ode = FUNC_CASES[name]()
    u = ode.u0
    dudt = O.residual(ode.p.arrays, u)
    ud, dd = to_device(u), to_device(dudt)
    expect = {
        PrimaryConservationAnalysis: (F.mass(sd, u), F.mass_rate(sd, u, dudt)),
        EnergyConservationAnalysis: (F.energy(sd, u), F.energy_rate(sd, u, dudt)),
        EntropyConservationAnalysis: (F.entropy(sd, law, u), F.entropy_rate(sd, law, u, dudt)),
    }
    for cls, (val, rate) in expect.items():
        a = cls(ode.p)
        for got in (a.evaluate_conservation(ud), a.evaluate_conservation(u)):
            np.testing.assert_allclose(got, val, rtol=1e-12, atol=1e-12 * np.abs(val).max(), err_msg=cls.__name__)
        scale = np.abs(F.energy(sd, np.abs(dudt))).max() + np.abs(val).max()
        for got in (a.evaluate_conservation_residual(ud, dd), a.evaluate_conservation_residual(u, dudt)):
            np.testing.assert_allclose(got, rate, rtol=1e-10, atol=1e-12 * scale, err_msg=cls.__name__)


def test_linear_residual_is_linear_and_skew_adjoint():
    # central skew-symmetric advection: M R is skew-symmetric, <x, M R y> = -<y, M R x>
    from oracle import functionals as F
    from sse.linear import LinearResidual
    law, sd, ex, form, ode = advection_2d_modal_tensor_tri()
    L = LinearResidual(ode.p)
    rng = np.random.default_rng(11)
    x, y = rng.standard_normal(L.shape[0]), rng.standard_normal(L.shape[0])
    Rx, Ry, Rxy = L @ x, L @ y, L @ (2.0 * x - 3.0 * y)
    np.testing.assert_allclose(Rxy, 2.0 * Rx - 3.0 * Ry, rtol=0, atol=1e-12 * np.abs(Rxy).max())
    shape = ode.p.size()

    def minner(a, b):
        A, B = a.reshape(shape, order="F"), b.reshape(shape, order="F")
        return sum(A[:, 0, k] @ F.mass_matrix(sd, k) @ B[:, 0, k] for k in range(sd.N_e))

    lhs, rhs = minner(x, Ry), -minner(y, Rx)
    assert abs(lhs - rhs) < 1e-11 * (abs(lhs) + 1.0), (lhs, rhs)


def test_linear_residual_adjoint_dot_product():
    # <v, R w>_M = <R* v, w>_M with R* the velocity-reversed residual, for BR1
    # advection-diffusion with LF on curved tets (the operator of the adjoint in b)
    from oracle import functionals as F
    from sse.linear import LinearResidual
    law, sd, ex, form, ode = advection_diffusion_3d_config(M=2, p=3)
    L = LinearResidual(ode.p)
    La = L.adjoint()
    rng = np.random.default_rng(5)
    v, w = rng.standard_normal(L.shape[0]), rng.standard_normal(L.shape[0])
    shape = ode.p.size()

    def minner(a, b):
        return F.energy_rate(sd, np.asfortranarray(a.reshape(shape, order="F")),
                             np.asfortranarray(b.reshape(shape, order="F")))[0]

    lhs, rhs = minner(v, L @ w), minner(La @ v, w)
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (lhs, rhs)


# ----------------------------------------------------------------------------- adjoint
def _ec_interface(mk):
    from sse.conservation_laws import EntropyConservativeNumericalFlux
    from sse.solvers import FluxDifferencingForm
    return mk(FluxDifferencingForm(inviscid_numerical_flux=EntropyConservativeNumericalFlux()))


VJP_CASES = {
    "wave3d_tet_p3_lf": lambda: density_wave_3d_config(M=2, p=3, ic_strength=0.2)[4],
    "wave3d_tet_p4_lf": lambda: density_wave_3d_config(M=2, p=4, ic_strength=0.2)[4],
    "vortex2d_tri_p2_lf": lambda: vortex_2d_config(M=3, p=2)[4],
    "wave3d_tet_p3_ec": lambda: _ec_interface(lambda f: density_wave_3d_config(M=2, p=3, ic_strength=0.2, form=f))[4],
    "vortex2d_tri_p4_lf": lambda: vortex_2d_config(M=3, p=4)[4],
    "vortex2d_tri_p6_lf": lambda: vortex_2d_config(M=2, p=6)[4],
    "wave3d_tet_p5_lf": lambda: density_wave_3d_config(M=1, p=5, ic_strength=0.2)[4],
    "wave3d_tet_p2_lf": lambda: density_wave_3d_config(M=2, p=2, ic_strength=0.2)[4],
    "euler1d_gauss_nodal_ec": lambda: euler_1d_gauss()[4],
}


@pytest.mark.parametrize("name", sorted(VJP_CASES))
def test_residual_vjp_dot_product_against_oracle_jvp(name):
    """<ybar, R'(u) du + dR/dλ dλ> (oracle, forward-mode dual numbers) equals
    <VJP(ybar), du> + lambda_bar dλ (GPU)."""
    from sse.adjoint import residual_vjp
    ode = VJP_CASES[name]()
    u = _perturbed(ode, amp=1e-3)
    rng = np.random.default_rng(17)
    du = np.asfortranarray(rng.standard_normal(u.shape))
    ybar = np.asfortranarray(rng.standard_normal(u.shape))
    dlam = 0.7
    r, dr = O.residual_jvp(ode.p.arrays, u, du, dlam)
    ubar, lam_bar = residual_vjp(ode.p, u, ybar)
    lhs = float(np.sum(ybar * dr))
    rhs = float(np.sum(ubar * du)) + lam_bar * dlam
    scale = float(np.sum(np.abs(ybar * dr)))
    assert abs(lhs - rhs) <= 1e-11 * scale, (name, lhs, rhs, scale)
    # the device-memory path gives the same pull-back
    from sse.time_integration import to_device, to_host
    ub_dev, lam_dev = residual_vjp(ode.p, to_device(u), to_device(ybar))
    assert _rel(to_host(ub_dev), ubar) <= 1e-14 and abs(lam_dev - lam_bar) <= 1e-12 * (1 + abs(lam_bar))


@pytest.mark.parametrize("make", [lambda: density_wave_3d_config(M=2, p=4)[4],
                                  lambda: vortex_2d_config(M=4, p=4)[4],
                                  lambda: advection_diffusion_3d_config(M=2, p=3)[4]],
                         ids=["tet_p4_euler", "tri_p4_euler", "tet_p3_br1"])
def test_autotune_times_every_configuration_and_selects_one(make, monkeypatch):
    # every compiled configuration gives the same residual (checked inside autotune)
    # and the selection is what a new handle then runs
    from sse import tune
    from sse.time_integration import to_device
    monkeypatch.delenv("SSE_KERNELS", raising=False)
    monkeypatch.delenv("SSE_TENSOR_VARIANT", raising=False)
    ode = make()
    u = to_device(_perturbed(ode, amp=1e-3))
    timings = tune.autotune(ode.p, u, nrep=3)
    assert len(timings) >= 2 and all(t.total_ms > 0 for t in timings)
    assert len({t.describe for t in timings}) == len(timings)
    tune.select(timings[0])
    fresh = ode.p.with_law(ode.p.conservation_law)
    assert fresh.describe() == timings[0].describe
    monkeypatch.delenv("SSE_KERNELS", raising=False)
    monkeypatch.delenv("SSE_TENSOR_VARIANT", raising=False)


def test_solve_with_save_callback_writes_the_reference_layout(tmp_path):
    # src/File/save.jl:74-86: every `interval` steps and both ends of tspan, with du = R(u)
    from sse import io
    from sse.time_integration import CarpenterKennedy2N54, solve
    law, sd, ic, form, ode = vortex_2d_config(M=3, p=3)
    ode.tspan = (0.0, 0.0105)
    path = io.save_project(law, sd, ic, form, ode.tspan, str(tmp_path / "res"))
    sol = solve(ode, CarpenterKennedy2N54(), dt=1e-3, callback=io.SaveCallback(path, interval=4))
    assert list(io.load_time_steps(path)) == [0, 4, 8, 11]
    u, du, t = io.load_solution(path, 11, load_du=True)
    assert t == 0.0105 and np.array_equal(u, sol.u)
    assert _rel(du, O.residual(ode.p.arrays, u)) <= RES_TOL
    u0, t0 = io.load_solution(path, 0)
    assert t0 == 0.0 and np.array_equal(u0, ode.u0)
    ref = O.rk_steps(ode.p.arrays, ode.u0, 1e-3, 8)
    assert _rel(io.load_solution(path, 8)[0], ref) <= RK_TOL


def test_reference_scaling_benchmark_problem_and_timings(tmp_path):
    # src/Analysis/benchmark.jl:116-183: the residual of its problem (ChanWarping-curved
    # Tri p=4, isentropic vortex) against the oracle, then the timing report
    from sse.benchmark import scaling_test_euler_2d, scaling_test_euler_2d_problem
    ode = scaling_test_euler_2d_problem(4, 4)
    got = _gpu_residual(ode)
    ref = O.residual(ode.p.arrays, ode.u0)
    assert _rel(got, ref) <= RES_TOL, _rel(got, ref)
    for _ in range(2):
        r = scaling_test_euler_2d(4, 4, path=str(tmp_path), samples=5, verbose=False)
        assert 0 < r["min_ms"] <= r["median_ms"] and r["N_e"] == 32
    assert np.load(tmp_path / "minimum.npy").shape == (2,)
    assert np.load(tmp_path / "median.npy").shape == (2,)


@pytest.mark.parametrize("name", ["wave3d_tet_p3_lf", "wave3d_tet_p4_lf", "vortex2d_tri_p4_lf"])
def test_tensor_vjp_equals_general_vjp(name, monkeypatch):
    # the compile-time-shaped VJP kernels against the general ones, entry by entry
    from sse.adjoint import residual_vjp
    ode = VJP_CASES[name]()
    u = _perturbed(ode, amp=1e-3)
    ybar = np.asfortranarray(np.random.default_rng(5).standard_normal(u.shape))
    fast = ode.p.with_law(ode.p.conservation_law)
    ub_fast, lam_fast = residual_vjp(fast, u, ybar)
    monkeypatch.setenv("SSE_VJP", "generic")
    general = ode.p.with_law(ode.p.conservation_law)
    ub_gen, lam_gen = residual_vjp(general, u, ybar)
    assert _rel(ub_fast, ub_gen) <= 1e-12, _rel(ub_fast, ub_gen)
    assert abs(lam_fast - lam_gen) <= 1e-12 * abs(lam_gen), (lam_fast, lam_gen)


@pytest.mark.parametrize("every", [3, 4, 8, 20])
def test_rk_adjoint_matches_oracle_forward_mode(every):
    """dJ along (du0, dλ) from the discrete adjoint (GPU, checkpoints every `every` of 8
    steps: uneven, even, one segment, one segment longer than the run; the last segment
    is kept from the forward sweep) equals the oracle's forward-mode derivative through
    the same 8 CK2N54 steps."""
    from oracle import functionals as F
    from sse.adjoint import rk_adjoint
    law, sd, ic, form, ode = density_wave_3d_config(M=2, p=3, ic_strength=0.2)
    u0 = _perturbed(ode, amp=1e-3)
    dt, nsteps = 2e-3, 8
    J, g, glam = rk_adjoint(ode.p, u0, dt, nsteps, checkpoint_every=every)
    uN = O.rk_steps(ode.p.arrays, u0, dt, nsteps)
    assert abs(J - F.entropy(sd, law, uN)[0]) <= 1e-12 * abs(J)
    rng = np.random.default_rng(23)
    du = np.asfortranarray(rng.standard_normal(u0.shape))
    dlam = 0.5
    uN2, duN = O.rk_jvp(ode.p.arrays, u0, du, dt, nsteps, dlam)
    # dJ = <dJ/du_N, du_N>, dJ/du_N = V'(WJ w(V u_N)) per element
    ref = sd.reference_approximation
    V = ref.Vdense()
    WJ = ref.W[:, None] * sd.geometric_factors.J_q
    from sse.conservation_laws import conservative_to_entropy
    wq = conservative_to_entropy(law, np.einsum("qp,pck->qkc", V, uN2))
    dJ_fwd = float(np.einsum("qk,qkc,qp,pck->", WJ, wq, V, duN))
    dJ_adj = float(np.sum(g * du)) + glam * dlam
    assert abs(dJ_fwd - dJ_adj) <= 1e-10 * (abs(dJ_fwd) + 1e-3), (dJ_fwd, dJ_adj)


def test_set_diffusion_equals_a_new_handle_and_recaptures_the_rk_step():
    # sse_set_diffusion on a handle gives the residual and the (graph-captured) RK steps of
    # a handle built with that coefficient
    import torch
    from sse import semi_discrete_residual
    from sse.hip import check
    from sse.time_integration import rk_steps, to_device, to_host
    solvers, u0s = _advdiff_ensemble(3, (4e-3, 7e-3))
    s0, s1 = solvers
    u = to_device(u0s[0])
    x0 = u.clone()
    rk_steps(s0, x0, 1e-3, 2)  # captures the step with b = 4e-3
    check(s0.lib, s0.handle, s0.lib.sse_set_diffusion(s0.handle, 7e-3))
    d0, d1 = torch.empty_like(u), torch.empty_like(u)
    semi_discrete_residual(d0, u, s0)
    semi_discrete_residual(d1, u, s1)
    torch.cuda.synchronize()
    assert _rel(to_host(d0), to_host(d1)) <= 1e-14
    x0, x1 = u.clone(), u.clone()
    rk_steps(s0, x0, 1e-3, 3)
    rk_steps(s1, x1, 1e-3, 3)
    assert _rel(to_host(x0), to_host(x1)) <= 1e-14


def test_energy_adjoint_in_b_with_shared_helpers_matches_oracle():
    # helpers built for b = 4e-3 serve a member with b = 7e-3
    from oracle import functionals as F
    from sse.adjoint import energy_adjoint_b, energy_adjoint_helpers
    from sse.time_integration import to_device
    solvers, u0s = _advdiff_ensemble(3, (4e-3, 7e-3))
    helpers = energy_adjoint_helpers(solvers[0])
    try:
        s, u0 = solvers[1], u0s[1]
        dt, n = 1e-3, 4
        J, dJdb, g = energy_adjoint_b(s, to_device(u0), dt, n, helpers)
        uN, duN = O.rk_jvp(s.arrays, u0, np.zeros_like(u0), dt, n, db=1.0)
        sd = s.spatial_discretization
        assert abs(J - 0.5 * F.energy_rate(sd, uN, uN)[0]) <= 1e-12 * abs(J)
        want = F.energy_rate(sd, uN, duN)[0]
        assert abs(dJdb - want) <= 1e-10 * abs(want), (dJdb, want)
    finally:
        for h_ in helpers:
            h_.close()


@pytest.mark.parametrize("keep_bytes", [16 << 30, 0], ids=["kept", "recomputed"])
def test_energy_adjoint_in_b_matches_oracle_forward_mode(keep_bytes):
    # the discrete adjoint of the BR1 solve (run on the forward kernels with the
    # velocity reversed) against the oracle's forward-mode duals: dJ/db for
    # J = 1/2 u_N' M u_N, and the gradient state through a dot-product test; with the
    # stages' viscous terms kept from the forward sweep, and recomputed from step states
    from oracle import functionals as F
    from sse.adjoint import energy_adjoint_b
    from sse.time_integration import to_device, to_host
    solvers, u0s = _advdiff_ensemble(3, (4e-3,))
    s, u0 = solvers[0], u0s[0]
    dt, n = 1e-3, 6
    J, dJdb, g = energy_adjoint_b(s, to_device(u0), dt, n, keep_bytes=keep_bytes)
    uN, duN = O.rk_jvp(s.arrays, u0, np.zeros_like(u0), dt, n, db=1.0)
    sd = s.spatial_discretization
    assert abs(J - 0.5 * F.energy_rate(sd, uN, uN)[0]) <= 1e-12 * abs(J)
    want = F.energy_rate(sd, uN, duN)[0]
    assert abs(dJdb - want) <= 1e-10 * abs(want), (dJdb, want)
    rng = np.random.default_rng(11)
    du0 = np.asfortranarray(rng.standard_normal(u0.shape))
    uN2, dN = O.rk_jvp(s.arrays, u0, du0, dt, n, db=0.0)
    lhs = F.energy_rate(sd, to_host(g), du0)[0]   # <g, du0>_M = dJ/du0 . du0
    rhs = F.energy_rate(sd, uN2, dN)[0]           # u_N' M (dU_N/du0 du0)
    assert abs(lhs - rhs) <= 1e-10 * abs(rhs), (lhs, rhs)
