"""Pins of the oracle that no reference-held vector covers.

The reference's golden values (test_oracle_golden.py) pin the 1D dense path and
the 2D ModalTensor triangle path end to end. The curved-tetrahedron path of the
benchmark and the oracle's pointwise physics are pinned here by what the
mathematics fixes, each check chosen so that a plausible slip (a dropped term, a
sign, a transposed operand, a wrong face index) fails it:

* EC two-point flux: consistency with the physical flux, symmetry, Tadmor's
  entropy-conservation condition (src/ConservationLaws/euler_navierstokes.jl:171-225);
* logmean / inv_logmean against 40-digit arithmetic on both sides of the series
  branch (src/ConservationLaws/ConservationLaws.jl:132-156);
* entropy variables: gradient of the entropy (finite differences) and the
  round trip (src/ConservationLaws/euler_navierstokes.jl:93-131);
* CK2N54: stability polynomial exp(z) through z^4 and exactness on t^3;
* reference operators: SBP property, orthonormal V, antisymmetric S, sum
  factorisation equal to the dense V, R exact on degree-p data;
* whole residual on curved tets: conservation, entropy conservation with the
  EC interface flux (src/Analysis/conservation.jl:169-189) and entropy
  dissipation with Lax-Friedrichs, free-stream preservation, high-order
  convergence to the exactly advected density wave, and the dense
  flux-differencing algorithm equal to the sparse one.
"""

import math

import mpmath
import numpy as np
import pytest

from cases import (advection_2d_modal_tensor_tri, advection_3d_modal_tensor_tet,
                   advection_diffusion_3d_config, density_wave_3d_config, energy_rate, l2_error,
                   vortex_2d_config)
from oracle import oracle as O
from sse.conservation_laws import (EntropyConservativeNumericalFlux, EulerEquations,
                                   LaxFriedrichsNumericalFlux, conservative_to_entropy,
                                   entropy)
from sse.reference_approximation import (ModalTensor, ReferenceApproximationTet,
                                         ReferenceApproximationTri, check_sbp_property,
                                         flux_differencing_operators)
from sse.simplex import vandermonde_tet, vandermonde_tri
from sse.solvers import FluxDifferencingForm

GAMMA = 1.4
RNG = np.random.default_rng(20261202)


def random_state(d, n=1, rng=RNG):
    rho = rng.uniform(0.5, 2.0, n)
    v = rng.uniform(-1.0, 1.0, (n, d))
    p = rng.uniform(0.5, 2.0, n)
    E = p / (GAMMA - 1) + 0.5 * rho * (v ** 2).sum(1)
    return np.concatenate([rho[:, None], rho[:, None] * v, E[:, None]], 1)


def physical_flux(u, d):
    rho, m, E = u[0], u[1:d + 1], u[d + 1]
    v = m / rho
    p = (GAMMA - 1) * (E - 0.5 * rho * (v ** 2).sum())
    F = np.zeros((d + 2, d))
    for j in range(d):
        F[0, j] = m[j]
        F[1:d + 1, j] = m * v[j]
        F[1 + j, j] += p
        F[d + 1, j] = (E + p) * v[j]
    return F


# ----------------------------------------------------------------------------- pointwise physics
@pytest.mark.parametrize("d", [1, 2, 3])
def test_ec_flux_consistency_and_symmetry(d):
    for u, w in zip(random_state(d, 20), random_state(d, 20)):
        np.testing.assert_allclose(O.two_point(1, d, GAMMA, u, u), physical_flux(u, d),
                                   rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(O.two_point(1, d, GAMMA, u, w), O.two_point(1, d, GAMMA, w, u),
                                   rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_ec_flux_tadmor_condition(d):
    # (w_R - w_L) . F(u_L, u_R) n = psi_R - psi_L with the potential psi = rho v . n
    law = EulerEquations(d, GAMMA)
    for uL, uR in zip(random_state(d, 50), random_state(d, 50)):
        n = RNG.normal(size=d)
        F = O.two_point(1, d, GAMMA, uL, uR) @ n
        wL, wR = conservative_to_entropy(law, uL), conservative_to_entropy(law, uR)
        psi = lambda u: u[1:d + 1] @ n  # noqa: E731
        lhs = (wR - wL) @ F
        assert abs(lhs - (psi(uR) - psi(uL))) < 1e-13 * (1 + np.abs(wR - wL) @ np.abs(F))


def test_ec_flux_tadmor_fails_for_planted_error():
    # the Tadmor check above is sharp: an arithmetic-mean density flux (a
    # plausible slip for the logarithmic mean) violates it
    d = 2
    law = EulerEquations(d, GAMMA)
    uL, uR = random_state(d, 2)
    n = np.array([1.0, 0.3])
    F = O.two_point(1, d, GAMMA, uL, uR)
    vL, vR = uL[1:3] / uL[0], uR[1:3] / uR[0]
    F_bad = F.copy()
    F_bad[0] = 0.5 * (uL[0] + uR[0]) * 0.5 * (vL + vR)
    wL, wR = conservative_to_entropy(law, uL), conservative_to_entropy(law, uR)
    assert abs((wR - wL) @ (F_bad @ n) - (uR[1:3] - uL[1:3]) @ n) > 1e-6


def test_logmean_against_high_precision():
    mpmath.mp.dps = 40
    xs = [1.0, 0.37, 2.5, 1e-3]
    # f = (x - y) / (x + y); series branch for f^2 < 1e-4, both sides tested
    for x in xs:
        for f in [0.0, 1e-9, 1e-5, 3e-3, 9.99e-3, 1.001e-2, 0.05, 0.3, 0.9]:
            y = x * (1 - f) / (1 + f)
            if f == 0.0:
                exact = mpmath.mpf(x)
            else:
                X, Y = mpmath.mpf(x), mpmath.mpf(y)
                exact = (X - Y) / (mpmath.log(X) - mpmath.log(Y))
            # the series branch is accurate to a few ulps; the log branch
            # (x - y) / (log x - log y) loses the cancellation of the logs
            tol = 4e-16
            if f * f >= 1e-4:
                tol *= 1 + (abs(math.log(x)) + abs(math.log(y))) / abs(math.log(x) - math.log(y))
            assert abs(O.logmean(x, y) - float(exact)) <= tol * float(exact), (x, f)
            assert abs(O.inv_logmean(x, y) - float(1 / exact)) <= tol * float(1 / exact), (x, f)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_entropy_variables_are_the_entropy_gradient(d):
    law = EulerEquations(d, GAMMA)
    for u in random_state(d, 10):
        w = O.c2e(d, GAMMA, u)
        np.testing.assert_allclose(w, conservative_to_entropy(law, u), rtol=1e-13, atol=1e-13)
        h = 1e-5
        grad = np.array([(entropy(law, u + h * e) - entropy(law, u - h * e)) / (2 * h)
                         for e in np.eye(d + 2)])
        np.testing.assert_allclose(w, grad, rtol=1e-7, atol=1e-8)
        np.testing.assert_allclose(O.e2c(d, GAMMA, w), u, rtol=1e-13, atol=1e-13)


# ----------------------------------------------------------------------------- time stepping
def _ck_step(f, y, t, dt):
    A, B, C = O.ck2n54()
    k = 0 * y
    for s in range(5):
        k = A[s] * k + dt * f(t + C[s] * dt, y)
        y = y + B[s] * k
    return y


def test_ck2n54_stability_polynomial_is_exp_through_z4():
    # one step of y' = z y from y = 1 is R(z); a degree-5 polynomial fitted
    # through 6 points recovers its coefficients exactly
    zs = np.linspace(-1.0, 1.0, 6)
    R = np.array([_ck_step(lambda t, y, z=z: z * y, 1.0, 0.0, 1.0) for z in zs])
    coef = np.polynomial.polynomial.polyfit(zs, R, 5)
    np.testing.assert_allclose(coef[:5], [1, 1, 1 / 2, 1 / 6, 1 / 24], atol=1e-12)
    assert abs(coef[5] - 1 / 120) > 1e-4  # 5 stages, order 4: R differs from exp at z^5


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_ck2n54_integrates_t_cubed_exactly(k):
    t0, dt = 0.3, 0.7
    y = _ck_step(lambda t, y: t ** k, 0.0, t0, dt)
    exact = ((t0 + dt) ** (k + 1) - t0 ** (k + 1)) / (k + 1)
    assert abs(y - exact) < 1e-14


# ----------------------------------------------------------------------------- reference operators
@pytest.mark.parametrize("elem,p", [("tri", 2), ("tri", 4), ("tri", 6), ("tet", 3), ("tet", 4)])
def test_reference_operators(elem, p):
    ref = (ReferenceApproximationTri if elem == "tri" else ReferenceApproximationTet)(ModalTensor(p))
    assert max(check_sbp_property(ref)) < 1e-12
    V = ref.Vdense()
    np.testing.assert_allclose(V.T @ (ref.W[:, None] * V), np.eye(ref.N_p), atol=1e-12)
    S, S_sparse, C, C_kind = flux_differencing_operators(ref)
    np.testing.assert_allclose(S, -np.transpose(S, (0, 2, 1)), atol=1e-14)
    np.testing.assert_allclose(C, ref.R.T * ref.B[None, :], atol=1e-14)
    # sum-factorised V equals the dense Vandermonde on random coefficients
    x = RNG.normal(size=ref.N_p)
    np.testing.assert_allclose(ref.warped.dense() @ x, V @ x, atol=1e-12)
    # facet extrapolation is exact on degree-p polynomials: R maps their values at
    # the volume nodes to their values at the facet nodes
    vdm = vandermonde_tri if elem == "tri" else vandermonde_tet
    re = ref.reference_element
    np.testing.assert_allclose(ref.R @ vdm(p, *re.rstq), vdm(p, *re.rstf), atol=1e-12)


def test_sbp_check_catches_planted_transpose():
    ref = ReferenceApproximationTri(ModalTensor(3))
    S, *_ = flux_differencing_operators(ref)
    # S_1^T in place of S_1 flips the skew part: W D + D^T W no longer matches
    assert np.abs(S[0] + S[0].T).max() < 1e-14
    assert np.abs(S[0].T - S[0]).max() > 1e-3


# ----------------------------------------------------------------------------- whole residual, curved tets
@pytest.fixture(scope="module")
def tet_ec():
    form = FluxDifferencingForm(inviscid_numerical_flux=EntropyConservativeNumericalFlux())
    return density_wave_3d_config(M=2, p=3, factor=0.05, form=form, ic_strength=0.2)


@pytest.fixture(scope="module")
def tet_lf():
    return density_wave_3d_config(M=2, p=3, factor=0.05, ic_strength=0.2)


def _entropy_rate(sd, law, u, dudt):
    # src/Analysis/conservation.jl:169-189 with P = M_k^{-1} V' WJ_k:
    # (P w_q)' M_k dudt = w_q' WJ_k V dudt
    ref = sd.reference_approximation
    V = ref.Vdense()
    WJ = ref.W[:, None] * sd.geometric_factors.J_q
    uq = np.einsum("qp,pck->qkc", V, u)
    wq = conservative_to_entropy(law, uq)
    Vd = np.einsum("qp,pck->qkc", V, dudt)
    terms = WJ[:, :, None] * wq * Vd
    return terms.sum(), np.abs(terms).sum()


def _perturbed(u, scale=0.02, rng=None):
    rng = rng or np.random.default_rng(7)
    v = u.copy()
    v[:, 0, :] *= 1 + scale * rng.uniform(-1, 1, v[:, 0, :].shape) * (np.arange(v.shape[0]) > 0)[:, None]
    return v


def test_tet_entropy_conservation_with_ec_interface_flux(tet_ec):
    law, sd, ic, form, ode = tet_ec
    u = _perturbed(ode.u0)
    dudt = O.residual(ode.p.arrays, u)
    rate, scale = _entropy_rate(sd, law, u, dudt)
    assert abs(rate) < 1e-12 * scale, (rate, scale)


def test_tet_entropy_dissipation_with_lax_friedrichs(tet_lf, tet_ec):
    law, sd, ic, form, ode = tet_lf
    u = _perturbed(ode.u0)
    rate, scale = _entropy_rate(sd, law, u, O.residual(ode.p.arrays, u))
    assert rate < -1e-6 * scale


def test_tet_conservation(tet_lf):
    law, sd, ic, form, ode = tet_lf
    ref = sd.reference_approximation
    u = _perturbed(ode.u0)
    dudt = O.residual(ode.p.arrays, u)
    WJ = ref.W[:, None] * sd.geometric_factors.J_q
    rate = np.einsum("qk,qp,pck->c", WJ, ref.Vdense(), dudt)
    scale = np.einsum("qk,qp,pck->c", WJ, np.abs(ref.Vdense()), np.abs(dudt))
    assert np.all(np.abs(rate) < 1e-13 * scale), rate / scale


def test_tet_free_stream_preservation(tet_lf):
    law, sd, ic, form, ode = tet_lf
    ref = sd.reference_approximation
    # constant state rho = 1.2, v = (0.3, -0.2, 0.5), p = 0.9 as modal coefficients
    rho, v, p = 1.2, np.array([0.3, -0.2, 0.5]), 0.9
    ucons = np.array([rho, *(rho * v), p / (GAMMA - 1) + 0.5 * rho * v @ v])
    coef0 = np.linalg.lstsq(ref.Vdense(), np.ones(ref.N_q), rcond=None)[0]
    u = np.asfortranarray(coef0[:, None, None] * ucons[None, :, None] * np.ones((1, 1, sd.N_e)))
    dudt = O.residual(ode.p.arrays, u)
    assert np.abs(dudt).max() < 1e-12


def test_tet_density_wave_converges_to_exact_advection():
    # the density wave rho = 1 + a sin(pi (x+y+z)), v = (1,1,1), p = 1 is advected
    # unchanged: u(x, t) = u0(x - t (1,1,1)). CK2N54 to T = 0.2 on two meshes;
    # the L2 density error must be small and fall at high order (a wrong sign or
    # scale anywhere in the residual gives an O(1) error that does not converge)
    T, a, p = 0.2, 0.2, 3
    errs = []
    for M in (4, 6):
        law, sd, ic, form, ode = density_wave_3d_config(M=M, p=p, factor=0.05, ic_strength=a)
        n = 4 * M
        uT = O.rk_steps(ode.p.arrays, ode.u0, T / n, n)
        errs.append(l2_error(sd, uT, lambda x, t: ic(tuple(xm - t for xm in x)), T)[0])
    order = math.log(errs[0] / errs[1]) / math.log(6 / 4)
    assert errs[1] < 0.02 and order > 2.5, (errs, order)


def test_dense_flux_differencing_equals_sparse(tet_lf):
    law, sd, ic, form, ode = tet_lf
    arrays = ode.p.arrays
    u = _perturbed(ode.u0)
    sparse = O.residual(arrays, u)
    saved = (arrays.desc.S_sparse, arrays.desc.C_kind)
    try:
        arrays.desc.S_sparse = 0
        arrays.desc.C_kind = 1
        dense = O.residual(arrays, u)
    finally:
        arrays.desc.S_sparse, arrays.desc.C_kind = saved
    np.testing.assert_allclose(dense, sparse, rtol=0, atol=1e-12 * np.abs(sparse).max())


def test_tri_entropy_conservation_with_ec_interface_flux():
    form = FluxDifferencingForm(inviscid_numerical_flux=EntropyConservativeNumericalFlux())
    law, sd, ic, form, ode = vortex_2d_config(M=4, p=4, form=form)
    u = _perturbed(ode.u0)
    rate, scale = _entropy_rate(sd, law, u, O.residual(ode.p.arrays, u))
    assert abs(rate) < 1e-12 * scale


# ----------------------------------------------------------------------------- StandardForm
@pytest.mark.parametrize("case", ["tri", "tet"])
def test_standard_form_skew_central_conserves_energy(case):
    # the skew-symmetric split form with the central flux conserves the discrete
    # energy on curved meshes (the reference's "energy ~ 0" checks,
    # test/runtests.jl:58-60, test/advection_3d.jl)
    law, sd, ex, form, ode = (advection_2d_modal_tensor_tri() if case == "tri"
                              else advection_3d_modal_tensor_tet())
    u = np.asfortranarray(ode.u0 + 0.01 * np.random.default_rng(3).standard_normal(ode.u0.shape))
    rate, scale = energy_rate(sd, u, O.residual(ode.p.arrays, u))
    assert abs(rate) < 1e-12 * scale, (rate, scale)


def test_standard_form_br1_dissipates_energy_and_conserves_mass():
    law, sd, ex, form, ode = advection_diffusion_3d_config(M=2, p=3, b=5e-2, sigma=0.25)
    ref = sd.reference_approximation
    u = ode.u0
    dudt = O.residual(ode.p.arrays, u)
    rate, scale = energy_rate(sd, u, dudt)
    assert rate < -1e-6 * scale
    WJ = ref.W[:, None] * sd.geometric_factors.J_q
    mass = np.einsum("qk,qp,pk->", WJ, ref.Vdense(), dudt[:, 0, :])
    mscale = np.einsum("qk,qp,pk->", WJ, np.abs(ref.Vdense()), np.abs(dudt[:, 0, :]))
    assert abs(mass) < 1e-13 * mscale


def test_standard_form_br1_converges_to_exact_gaussian():
    # advection-diffusion of a Gaussian on curved tets against the closed form of
    # src/ConservationLaws/linear_advection_diffusion.jl:147-165 (pulse away from the
    # periodic boundary over the short horizon)
    T = 0.05
    errs = []
    for M in (2, 3):
        law, sd, ex, form, ode = advection_diffusion_3d_config(M=M, p=4, b=1e-2, sigma=0.15)
        n = 10 * M
        uT = O.rk_steps(ode.p.arrays, ode.u0, T / n, n)
        errs.append(l2_error(sd, uT, ex, T)[0])
    assert errs[1] < 0.7 * errs[0] and errs[1] < 0.02, errs


def _advdiff_with_b(b, M=2, p=3):
    from sse.conservation_laws import LinearAdvectionDiffusionEquation
    law, sd, ex, form, ode = advection_diffusion_3d_config(M=M, p=p, b=5e-3, sigma=0.25)
    return ode, ode.p.with_law(LinearAdvectionDiffusionEquation(law.a, b))


def test_br1_residual_derivative_in_b_is_the_viscous_part():
    # the BR1 residual is affine in b (viscous flux -b q, interface flux -b {q}.n,
    # q independent of b): dR/db = R(u; 1) - R(u; 0) whatever b; a tangent that
    # missed the interface term or the volume term fails this
    ode, s7 = _advdiff_with_b(7e-3)
    _, s0 = _advdiff_with_b(0.0)
    _, s1 = _advdiff_with_b(1.0)
    u = ode.u0
    r, dr = O.residual_jvp(s7.arrays, u, np.zeros_like(u), 0.0, 1.0)
    np.testing.assert_allclose(r, O.residual(s7.arrays, u), rtol=0, atol=1e-14 * np.abs(r).max())
    want = O.residual(s1.arrays, u) - O.residual(s0.arrays, u)
    assert np.abs(dr - want).max() <= 1e-12 * np.abs(want).max()
    # and R(u; b) = R(u; 0) + b dR/db
    np.testing.assert_allclose(r, O.residual(s0.arrays, u) + 7e-3 * dr, rtol=0, atol=1e-12 * np.abs(r).max())


def test_br1_rk_tangent_in_b_matches_central_differences():
    # d u_N / d b through 10 CK2N54 steps (forward-mode duals) against a central
    # difference in b; u_N is a polynomial in b, the difference error is O(h^2)
    b, h, dt, n = 5e-3, 1e-6, 2e-3, 10
    ode, sb = _advdiff_with_b(b)
    _, sp = _advdiff_with_b(b + h)
    _, sm = _advdiff_with_b(b - h)
    u0 = ode.u0
    uN, duN = O.rk_jvp(sb.arrays, u0, np.zeros_like(u0), dt, n, db=1.0)
    fd = (O.rk_steps(sp.arrays, u0, dt, n) - O.rk_steps(sm.arrays, u0, dt, n)) / (2 * h)
    np.testing.assert_allclose(uN, O.rk_steps(sb.arrays, u0, dt, n), rtol=0, atol=1e-14 * np.abs(uN).max())
    assert np.abs(duN).max() > 0
    assert np.abs(duN - fd).max() <= 1e-6 * np.abs(duN).max(), np.abs(duN - fd).max() / np.abs(duN).max()


# ----------------------------------------------------------------------------- functionals
def test_functionals_closed_forms():
    from oracle import functionals as F
    law, sd, ic, form, ode = density_wave_3d_config(M=2, p=3, factor=0.05, ic_strength=0.2)
    ref = sd.reference_approximation
    # a constant state: mass = state x volume (8), entropy = S(state) x volume
    rho, v, p = 1.3, np.array([0.2, -0.1, 0.4]), 0.8
    uc = np.array([rho, *(rho * v), p / (GAMMA - 1) + 0.5 * rho * v @ v])
    coef0 = np.linalg.lstsq(ref.Vdense(), np.ones(ref.N_q), rcond=None)[0]
    u = np.asfortranarray(coef0[:, None, None] * uc[None, :, None] * np.ones((1, 1, sd.N_e)))
    np.testing.assert_allclose(F.mass(sd, u), 8.0 * uc, rtol=1e-12)
    S = -rho * math.log(p / rho ** GAMMA) / (GAMMA - 1)
    np.testing.assert_allclose(F.entropy(sd, law, u), [8.0 * S], rtol=1e-12)
    # entropy rate: the literal (P w)' M dudt equals w_q' WJ V dudt
    dudt = O.residual(ode.p.arrays, ode.u0)
    rate, _ = _entropy_rate(sd, law, ode.u0, dudt)
    np.testing.assert_allclose(F.entropy_rate(sd, law, ode.u0, dudt), [rate], rtol=1e-9, atol=1e-12)


def test_energy_functional_on_affine_mesh_is_half_l2_norm():
    # with constant J the weight-adjusted mass matrix is V' W J V exactly, so the
    # energy is half the quadrature L2 norm of the polynomial
    from oracle import functionals as F
    law, sd, ic, form, ode = vortex_2d_config(M=2, p=3)
    ref = sd.reference_approximation
    uq = np.einsum("qp,pck->qck", ref.Vdense(), ode.u0)
    WJ = ref.W[:, None] * sd.geometric_factors.J_q
    np.testing.assert_allclose(F.energy(sd, ode.u0), 0.5 * np.einsum("qk,qck->c", WJ, uq ** 2), rtol=1e-12)


def test_reference_scaling_benchmark_problem():
    # src/Analysis/benchmark.jl:120-159: sizes, end time L / Ma, and a mass-conserving
    # residual on the ChanWarping-curved mesh
    from sse.benchmark import scaling_test_euler_2d_problem
    ode = scaling_test_euler_2d_problem(3, 3)
    assert ode.p.size() == (10, 4, 18) and ode.tspan == (0.0, 2.5)
    sd = ode.p.spatial_discretization
    ref = sd.reference_approximation
    dudt = O.residual(ode.p.arrays, ode.u0)
    WJ = ref.W[:, None] * sd.geometric_factors.J_q
    mass = np.einsum("qk,qp,pck->c", WJ, ref.Vdense(), dudt)
    scale = np.einsum("qk,qp,pck->c", WJ, np.abs(ref.Vdense()), np.abs(dudt))
    assert np.all(np.abs(mass) <= 1e-13 * scale), (mass, scale)


def test_br1_weak_form_transpose_is_the_velocity_reversed_weak_form():
    """(M A(a, b))' = M A(-a, b) for the skew-symmetric StandardForm with LF and BR1 on
    curved tets: advection is skew in M, LF and BR1 symmetric. This is what the
    velocity-reversal adjoint of the advection-diffusion solve rests on; a planted
    error (the LF coefficient changed on one side) must break it."""
    from sse.conservation_laws import LinearAdvectionDiffusionEquation
    from sse.workloads import advection_diffusion_3d_config
    law, sd, ex, form, ode = advection_diffusion_3d_config(M=1, p=2)
    s = ode.p
    shape = ode.u0.shape
    ref = sd.reference_approximation
    V, W, Jq = ref.V, ref.W, sd.geometric_factors.J_q
    N_p, _, N_e = shape
    n = N_p * N_e
    M = np.zeros((n, n))
    for k in range(N_e):
        M[k * N_p:(k + 1) * N_p, k * N_p:(k + 1) * N_p] = np.linalg.inv(V.T @ ((W / Jq[:, k])[:, None] * V))

    def weak(arrays):
        A = np.empty((n, n))
        for i in range(n):
            e = np.zeros(n)
            e[i] = 1.0
            A[:, i] = O.residual(arrays, np.asfortranarray(e.reshape(shape, order="F"))).reshape(-1, order="F")
        return M @ A

    B = weak(s.arrays)
    rev = s.with_law(LinearAdvectionDiffusionEquation(tuple(-x for x in law.a), law.b))
    Brev = weak(rev.arrays)
    assert np.abs(B.T - Brev).max() <= 1e-12 * np.abs(B).max()
    planted = rev.arrays
    planted.desc.half_lambda = 0.5 * planted.desc.half_lambda
    assert np.abs(B.T - weak(planted)).max() > 1e-6 * np.abs(B).max()
