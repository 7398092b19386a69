"""Pins of the multidimensional (dense-operator) schemes ModalMulti and NodalMulti
on triangles and tetrahedra (src/SpatialDiscretizations/multidimensional.jl:1-75),
checked on the CPU by what the mathematics fixes:

* the degree-2p volume and facet quadratures integrate every monomial of total
  degree <= 2p exactly (closed-form simplex integrals), and a rule one degree
  short fails;
* the operators: SBP property (SpatialDiscretizations.jl:441-455), orthonormal
  modal basis, D_m exact on degree-p polynomials, R exact on degree-p data,
  antisymmetric S, dense flux differencing and facet correction (C = R'B);
* the whole residual on curved meshes through the oracle's dense algorithms
  (flux_differencing_form.jl:1-35, 91-124): conservation, entropy conservation
  with the EC interface flux and dissipation with Lax-Friedrichs
  (src/Analysis/conservation.jl:169-189), free-stream preservation, and energy
  conservation of the skew-symmetric StandardForm with the central flux.

The reference takes its default Tri/Tet volume rules from StartUpDG's tables
(not available here); these schemes use collapsed rules of the same degree,
so the reference's notebook outputs are not reproduced value for value.
"""

import math

import numpy as np
import pytest

from cases import advection_2d_modal_tensor_tri, density_wave_3d_config, energy_rate, euler_vortex_2d_modal
from oracle import oracle as O
from sse.conservation_laws import EntropyConservativeNumericalFlux, conservative_to_entropy
from sse.reference_approximation import (GaussQuadrature, LGQuadrature, ModalMulti, NodalMulti,
                                         ReferenceApproximationTet, ReferenceApproximationTri,
                                         check_sbp_property, flux_differencing_operators, quadrature_tet,
                                         quadrature_tri)
from sse.simplex import basis_tet, basis_tri, vandermonde_tet, vandermonde_tri
from sse.solvers import FluxDifferencingForm

GAMMA = 1.4


def _simplex_monomial(powers):
    """Integral of prod r_i^k_i over the reference simplex with vertices -1 and e_i:
    r = 2x - 1 with x on the unit simplex, where the integral of prod x_i^a_i is
    prod a_i! / (sum a + d)! (Dirichlet's formula)."""
    d = len(powers)
    total = 0.0
    # expand (2x - 1)^k = sum_a C(k, a) 2^a x^a (-1)^(k - a)
    def rec(i, coef, alphas):
        nonlocal total
        if i == d:
            total += coef * math.prod(math.factorial(a) for a in alphas) / math.factorial(sum(alphas) + d)
            return
        k = powers[i]
        for a in range(k + 1):
            rec(i + 1, coef * math.comb(k, a) * 2.0 ** a * (-1.0) ** (k - a), alphas + [a])
    rec(0, 1.0, [])
    return total * 2.0 ** d


def _monomials(d, deg):
    if d == 2:
        return [(i, j) for i in range(deg + 1) for j in range(deg + 1 - i)]
    return [(i, j, k) for i in range(deg + 1) for j in range(deg + 1 - i) for k in range(deg + 1 - i - j)]


def _worst(rst, w, deg):
    errs = []
    for pw in _monomials(len(rst), deg):
        f = np.prod([x ** k for x, k in zip(rst, pw)], axis=0)
        errs.append(abs(w @ f - _simplex_monomial(pw)))
    return max(errs)


def test_closed_form_monomial_integrals():
    # area / volume of the reference triangle and tetrahedron, and a first moment
    assert _simplex_monomial((0, 0)) == pytest.approx(2.0, rel=1e-15)
    assert _simplex_monomial((0, 0, 0)) == pytest.approx(4.0 / 3.0, rel=1e-15)
    assert _simplex_monomial((1, 0)) == pytest.approx(-2.0 / 3.0, rel=1e-15)  # centroid r = -1/3


@pytest.mark.parametrize("elem,p", [("tri", 2), ("tri", 4), ("tet", 2), ("tet", 3)])
def test_multi_quadratures_are_exact_to_degree_2p(elem, p):
    ref = (ReferenceApproximationTri if elem == "tri" else ReferenceApproximationTet)(ModalMulti(p))
    re = ref.reference_element
    assert _worst(re.rstq, re.wq, 2 * p) < 1e-13
    # facet rule: each face is an affine image of the reference (d-1)-simplex, so
    # the facet weights of one face sum to its reference measure and integrate
    # monomials of the face parameters exactly; check through the face areas
    nf = ref.nodes_per_face
    areas = [ref.B[f * nf:(f + 1) * nf].sum() for f in range(len(re.face_vertices))]
    if elem == "tri":
        np.testing.assert_allclose(areas, [2.0, 2.0, 2.0], rtol=1e-14)  # the wf are in parameter length
    else:
        np.testing.assert_allclose(areas, [2.0] * 4, rtol=1e-14)
    # a rule one degree short is not exact at degree 2p
    short = ((LGQuadrature(p - 1), LGQuadrature(p - 1)) if elem == "tri"
             else (LGQuadrature(p - 1), LGQuadrature(p - 1), GaussQuadrature(p - 1, 1, 0)))
    *rst, w = quadrature_tri(short) if elem == "tri" else quadrature_tet(short)
    assert _worst(rst, w, 2 * p) > 1e-6


@pytest.mark.parametrize("approx", [ModalMulti, NodalMulti])
@pytest.mark.parametrize("elem,p", [("tri", 2), ("tri", 4), ("tet", 2), ("tet", 3)])
def test_multi_reference_operators(elem, p, approx):
    tri = elem == "tri"
    ref = (ReferenceApproximationTri if tri else ReferenceApproximationTet)(approx(p))
    re = ref.reference_element
    assert ref.reference_mapping is None and ref.warped is None
    assert max(check_sbp_property(ref)) < 1e-12
    if approx is ModalMulti:
        V = ref.V
        np.testing.assert_allclose(V.T @ (ref.W[:, None] * V), np.eye(ref.N_p), atol=1e-12)
    else:
        assert ref.V is None and ref.N_p == ref.N_q
    # D_m is the exact derivative, and R the exact extrapolation, of degree-p data
    # given by its values at the volume nodes
    VDM, *grads = (basis_tri if tri else basis_tet)(p, *re.rstq)
    for m in range(ref.d):
        np.testing.assert_allclose(ref.D[m] @ VDM, grads[m], atol=1e-11)
    vdm = vandermonde_tri if tri else vandermonde_tet
    np.testing.assert_allclose(ref.R @ VDM, vdm(p, *re.rstf), atol=1e-12)
    S, S_sparse, C, C_kind = flux_differencing_operators(ref)
    assert not S_sparse and C_kind == 1  # the dense algorithms
    np.testing.assert_allclose(S, -np.transpose(S, (0, 2, 1)), atol=1e-14)
    np.testing.assert_allclose(C, ref.R.T * ref.B[None, :], atol=1e-14)


def test_multi_sbp_check_catches_planted_swap():
    ref = ReferenceApproximationTri(ModalMulti(3))
    ref.D = (ref.D[1], ref.D[0])  # D_r and D_s exchanged
    assert max(check_sbp_property(ref)) > 1e-3


# ----------------------------------------------------------------------------- whole residual
def _make(elem, approx, form=None):
    if elem == "tri":
        return euler_vortex_2d_modal(M=3, p=4, approx=approx, form=form)[:5]
    return density_wave_3d_config(M=2, p=3, approx=approx, form=form, ic_strength=0.2)


def _perturbed(ode, amp=1e-3, seed=11):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(ode.u0 + amp * rng.standard_normal(ode.u0.shape) * np.abs(ode.u0).max())


def _entropy_rate(sd, law, u, dudt):
    ref = sd.reference_approximation
    V = ref.Vdense()
    WJ = ref.W[:, None] * sd.geometric_factors.J_q
    wq = conservative_to_entropy(law, np.einsum("qp,pck->qkc", V, u))
    terms = WJ[:, :, None] * wq * np.einsum("qp,pck->qkc", V, dudt)
    return terms.sum(), np.abs(terms).sum()


CASES = [(e, a) for e in ("tri", "tet") for a in (ModalMulti, NodalMulti)]
IDS = [f"{e}-{a.__name__}" for e, a in CASES]


@pytest.mark.parametrize("elem,approx", CASES, ids=IDS)
def test_multi_entropy_conservation_and_dissipation(elem, approx):
    form = FluxDifferencingForm(inviscid_numerical_flux=EntropyConservativeNumericalFlux())
    law, sd, ic, form, ode = _make(elem, approx, form)
    u = _perturbed(ode)
    rate, scale = _entropy_rate(sd, law, u, O.residual(ode.p.arrays, u))
    assert abs(rate) < 1e-12 * scale, (rate, scale)
    law, sd, ic, form, ode = _make(elem, approx)
    rate, scale = _entropy_rate(sd, law, u, O.residual(ode.p.arrays, u))
    assert rate < -1e-6 * scale


@pytest.mark.parametrize("elem,approx", CASES, ids=IDS)
def test_multi_conservation_and_free_stream(elem, approx):
    law, sd, ic, form, ode = _make(elem, approx)
    ref = sd.reference_approximation
    V = ref.Vdense()
    WJ = ref.W[:, None] * sd.geometric_factors.J_q
    u = _perturbed(ode)
    dudt = O.residual(ode.p.arrays, u)
    rate = np.einsum("qk,qp,pck->c", WJ, V, dudt)
    scale = np.einsum("qk,qp,pck->c", WJ, np.abs(V), np.abs(dudt))
    assert np.all(np.abs(rate) < 1e-13 * scale), rate / scale
    rho, p = 1.2, 0.9
    v = np.array([0.3, -0.2, 0.5])[:sd.reference_approximation.d]
    ucons = np.array([rho, *(rho * v), p / (GAMMA - 1) + 0.5 * rho * v @ v])
    coef0 = np.linalg.lstsq(V, np.ones(ref.N_q), rcond=None)[0]
    uf = np.asfortranarray(coef0[:, None, None] * ucons[None, :, None] * np.ones((1, 1, sd.N_e)))
    assert np.abs(O.residual(ode.p.arrays, uf)).max() < 1e-11


@pytest.mark.parametrize("approx", [ModalMulti, NodalMulti])
def test_multi_standard_form_skew_central_conserves_energy(approx):
    # the scheme of examples/advection_2d.ipynb (NodalMulti Tri, skew-symmetric split
    # form, central flux) on a curved mesh. The discrete energy is u' M_k dudt with
    # M_k = diag(W J_k) for the nodal scheme (DiagonalSolver) and the inverse of the
    # weight-adjusted V' diag(W / J_k) V for the modal one.
    law, sd, ex, form, ode = advection_2d_modal_tensor_tri(M=3, p=4, approx=approx)
    u = np.asfortranarray(ode.u0 + 0.01 * np.random.default_rng(3).standard_normal(ode.u0.shape))
    dudt = O.residual(ode.p.arrays, u)
    if approx is NodalMulti:
        ref, J = sd.reference_approximation, sd.geometric_factors.J_q
        terms = u[:, 0, :] * (ref.W[:, None] * J) * dudt[:, 0, :]
        rate, scale = terms.sum(), np.abs(terms).sum()
    else:
        rate, scale = energy_rate(sd, u, dudt)
    assert abs(rate) < 1e-12 * scale, (rate, scale)
    # the same rate with a planted error in the mass matrix (W alone) is not zero
    if approx is NodalMulti:
        wrong = (u[:, 0, :] * ref.W[:, None] * dudt[:, 0, :]).sum()
        assert abs(wrong) > 1e-6 * scale
