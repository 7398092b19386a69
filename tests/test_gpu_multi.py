"""The multidimensional dense-operator schemes (ModalMulti, NodalMulti on
triangles and tetrahedra, src/SpatialDiscretizations/multidimensional.jl:1-75)
through the HIP library against the CPU oracle.

These schemes have dense S_m and C = R'B, so the library runs its general
(runtime-sized, LDS-staged) kernels, with the dense Hadamard sum and facet
correction (src/Solvers/flux_differencing_form.jl:1-35, 91-124). Bar as for the
other parity tests: 1e-12 per residual, 1e-10 after RK steps, 1e-11 for the VJP
dot-product identity.
"""

import numpy as np
import pytest

from cases import advection_2d_modal_tensor_tri, density_wave_3d_config, euler_vortex_2d_modal
from oracle import oracle as O
from sse.reference_approximation import ModalMulti, NodalMulti

pytestmark = pytest.mark.gpu

RES_TOL = 1e-12
RK_TOL = 1e-10


def _ec(mk):
    from sse.conservation_laws import EntropyConservativeNumericalFlux
    from sse.solvers import FluxDifferencingForm
    return mk(FluxDifferencingForm(inviscid_numerical_flux=EntropyConservativeNumericalFlux()))


CASES = {
    "vortex2d_modal_multi_p4_lf": lambda: euler_vortex_2d_modal(M=3, p=4, approx=ModalMulti)[4],
    "vortex2d_nodal_multi_p4_lf": lambda: euler_vortex_2d_modal(M=3, p=4, approx=NodalMulti)[4],
    "vortex2d_modal_multi_p2_ec": lambda: _ec(lambda f: euler_vortex_2d_modal(M=3, p=2, approx=ModalMulti,
                                                                             form=f))[4],
    "wave3d_modal_multi_p3_ec": lambda: _ec(lambda f: density_wave_3d_config(M=2, p=3, approx=ModalMulti,
                                                                             ic_strength=0.2, form=f))[4],
    "wave3d_modal_multi_p3_lf": lambda: density_wave_3d_config(M=2, p=3, approx=ModalMulti, ic_strength=0.2)[4],
    "wave3d_nodal_multi_p3_lf": lambda: density_wave_3d_config(M=2, p=3, approx=NodalMulti, ic_strength=0.2)[4],
    "wave3d_modal_multi_p4_lf": lambda: density_wave_3d_config(M=2, p=4, approx=ModalMulti, ic_strength=0.2)[4],
    "adv2d_nodal_multi_p4_skew_central": lambda: advection_2d_modal_tensor_tri(M=3, p=4, approx=NodalMulti)[4],
    "adv2d_modal_multi_p3_skew_central": lambda: advection_2d_modal_tensor_tri(M=3, p=3, approx=ModalMulti)[4],
}


def _perturbed(ode, amp=1e-3, seed=20261202):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(ode.u0 + amp * rng.standard_normal(ode.u0.shape) * np.abs(ode.u0).max())


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _gpu_residual(ode, u):
    import torch
    from sse import semi_discrete_residual
    from sse.time_integration import to_device, to_host
    assert torch.cuda.is_available()
    ud = to_device(u)
    dd = torch.empty_like(ud)
    semi_discrete_residual(dd, ud, ode.p)
    torch.cuda.synchronize()
    return to_host(dd)


@pytest.mark.parametrize("name", sorted(CASES))
def test_multi_residual_matches_oracle(name):
    ode = CASES[name]()
    u = _perturbed(ode)
    got = _gpu_residual(ode, u)
    ref = O.residual(ode.p.arrays, u)
    assert _rel(got, ref) <= RES_TOL, (name, _rel(got, ref))
    kind = "standard" if name.startswith("adv") else "generic"
    assert ode.p.describe().startswith(kind), ode.p.describe()


@pytest.mark.parametrize("name", ["wave3d_modal_multi_p3_lf", "vortex2d_nodal_multi_p4_lf",
                                  "adv2d_nodal_multi_p4_skew_central"])
def test_multi_rk_steps_match_oracle(name):
    from sse.time_integration import rk_steps, to_device, to_host
    ode = CASES[name]()
    dt = 1e-3
    ud = to_device(ode.u0)
    rk_steps(ode.p, ud, dt, 20)
    got = to_host(ud)
    ref = O.rk_steps(ode.p.arrays, ode.u0, dt, 20)
    assert _rel(got, ref) <= RK_TOL, _rel(got, ref)


@pytest.mark.parametrize("name", ["wave3d_modal_multi_p3_ec", "vortex2d_nodal_multi_p4_lf"])
def test_multi_residual_vjp_dot_product_against_oracle_jvp(name):
    """<ybar, R'(u) du + dR/dλ dλ> (oracle, forward-mode dual numbers) equals
    <VJP(ybar), du> + lambda_bar dλ (GPU)."""
    from sse.adjoint import residual_vjp
    ode = CASES[name]()
    u = _perturbed(ode)
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


def test_multi_repeated_residuals_agree_to_rounding():
    # the general kernels add each flux to both of its nodes with LDS atomic adds,
    # so the order of those sums may change between calls; nothing else does
    ode = CASES["wave3d_modal_multi_p3_lf"]()
    u = _perturbed(ode)
    first = _gpu_residual(ode, u)
    for _ in range(3):
        assert _rel(_gpu_residual(ode, u), first) <= 1e-14


MFMA_CASES = ["wave3d_modal_multi_p3_ec", "wave3d_modal_multi_p3_lf", "wave3d_modal_multi_p4_lf",
              "vortex2d_modal_multi_p4_lf", "vortex2d_modal_multi_p2_ec"]


@pytest.mark.parametrize("name", MFMA_CASES)
def test_matrix_core_nodal_values_match_oracle(name, monkeypatch):
    # the nodal values on the matrix cores (16 elements per workgroup, MFMA tiles for V,
    # V' and R), switched on for these small meshes, whose last workgroup is partly empty
    monkeypatch.setenv("SSE_NODAL_MFMA", "1")
    ode = CASES[name]()
    assert "nodal_values=mfma" in ode.p.describe(), ode.p.describe()
    u = _perturbed(ode)
    got = _gpu_residual(ode, u)
    ref = O.residual(ode.p.arrays, u)
    assert _rel(got, ref) <= RES_TOL, (name, _rel(got, ref))


def test_matrix_core_nodal_values_are_chosen_on_a_large_mesh_and_match_oracle():
    # 10^3 hexes x 6 tets = 6,000 elements: the matrix-core kernel by default; 20 RK steps
    ode = density_wave_3d_config(M=10, p=3, approx=ModalMulti, ic_strength=0.2)[4]
    assert "nodal_values=mfma" in ode.p.describe(), ode.p.describe()
    u = _perturbed(ode)
    got = _gpu_residual(ode, u)
    ref = O.residual(ode.p.arrays, u)
    assert _rel(got, ref) <= RES_TOL, _rel(got, ref)
    from sse.time_integration import rk_steps, to_device, to_host
    dt = 1e-3
    ud = to_device(ode.u0)
    rk_steps(ode.p, ud, dt, 20)
    ref = O.rk_steps(ode.p.arrays, ode.u0, dt, 20)
    assert _rel(to_host(ud), ref) <= RK_TOL, _rel(to_host(ud), ref)
