"""Set-up on the GPU (sse_geometric_factors, sse_apply_reference_mapping,
sse_project_l2) against the host set-up and against closed forms.

Bar: every array within 1e-12 of the host's, relative to its largest entry. The
host functions are pinned by the reference's golden errors (the vortex and 1D
cases run through them) and by the invariants of test_oracle_invariants.py; the
closed-form tests below pin the device path independently: on affine elements
both metric types equal the adjugate of the affine map, and the L2 projection
reproduces a polynomial of the solution degree.
"""

import numpy as np
import pytest

from cases import (advection_diffusion_3d_config, density_wave_3d_config, euler_1d_gauss,
                   vortex_2d_config)

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rel(a, b):
    scale = max(float(np.abs(b).max()), 1e-300)
    return float(np.abs(np.asarray(a) - np.asarray(b)).max()) / scale


def _compare_gf(g, h):
    for name in ("J_q", "Lambda_q", "J_f", "nJf", "nJq"):
        a, b = getattr(g, name), getattr(h, name)
        assert a.shape == b.shape, name
        err = _rel(a, b)
        assert err <= TOL, f"{name}: {err:.3e}"


def _sd(builder):
    return builder()[1]


SD_CASES = {
    "line_gauss_p5": lambda: _sd(euler_1d_gauss),
    "tri_p4_warped": lambda: _sd(lambda: vortex_2d_config(M=4, p=4)),
    "tri_p2": lambda: _sd(lambda: vortex_2d_config(M=3, p=2)),
    "tet_p3_curved": lambda: _sd(lambda: density_wave_3d_config(M=2, p=3)),
    "tet_p4_curved": lambda: _sd(lambda: density_wave_3d_config(M=2, p=4)),
    "tet_p5_curved": lambda: _sd(lambda: advection_diffusion_3d_config(M=2, p=5)),
}


@pytest.mark.parametrize("metrics", ["exact", "curl"])
@pytest.mark.parametrize("name", sorted(SD_CASES))
def test_geometric_factors_match_host(name, metrics):
    from sse.geometry import (ChanWilcoxMetrics, ExactMetrics, geometric_factors_curl,
                              geometric_factors_exact, geometric_factors_hip, project_jacobian)
    from sse.reference_approximation import LINE
    sd = SD_CASES[name]()
    ref, mesh = sd.reference_approximation, sd.mesh
    if metrics == "curl":
        if ref.element_type == LINE:
            pytest.skip("curl metrics are defined for triangles and tetrahedra")
        host = geometric_factors_curl(mesh, ref)
        dev = geometric_factors_hip(mesh, ref, ChanWilcoxMetrics())
    else:
        host = geometric_factors_exact(mesh, ref)
        host.J_q = project_jacobian(host.J_q, ref)
        dev = geometric_factors_hip(mesh, ref, ExactMetrics(), project_jacobian_flag=True)
    _compare_gf(dev, host)


def _affine_map(mesh, rd, k):
    """x = c + A r on element k, fitted to its mapping nodes (exact for affine maps)."""
    d = len(mesh.xyz)
    R = np.column_stack([np.ones_like(rd.rst[0])] + list(rd.rst[:d]))
    X = np.column_stack([mesh.xyz[m][:, k] for m in range(d)])
    coef, *_ = np.linalg.lstsq(R, X, rcond=None)
    A = coef[1:].T  # A[m, n] = d x_m / d r_n
    assert np.abs(R @ coef - X).max() < 1e-12
    return A


@pytest.mark.parametrize("metrics", ["exact", "curl"])
@pytest.mark.parametrize("dim", [2, 3])
def test_affine_elements_give_the_adjugate(dim, metrics):
    """Unwarped meshes: Lambda_q = J A^{-1}, J_q = det A, nJf = Lambda^T nrstJ and
    J_f = |nJf| at every node, for both metric types."""
    from sse.geometry import ChanWilcoxMetrics, ExactMetrics, geometric_factors_hip
    from sse.mesh import uniform_periodic_mesh
    from sse.reference_approximation import (ModalTensor, ReferenceApproximationTet,
                                             ReferenceApproximationTri)
    if dim == 2:
        ref = ReferenceApproximationTri(ModalTensor(4), mapping_degree=4)
        mesh = uniform_periodic_mesh(ref, ((0.0, 1.0), (0.0, 2.0)), (3, 2))
    else:
        ref = ReferenceApproximationTet(ModalTensor(3), mapping_degree=3)
        mesh = uniform_periodic_mesh(ref, ((0.0, 1.0), (0.0, 2.0), (0.0, 3.0)), (2, 2, 1))
    rd = ref.reference_element
    mt = ChanWilcoxMetrics() if metrics == "curl" else ExactMetrics()
    g = geometric_factors_hip(mesh, ref, mt, project_jacobian_flag=False)
    nrst = np.stack(rd.nrstJ, axis=0)  # d x N_f
    for k in range(mesh.N_e):
        A = _affine_map(mesh, rd, k)
        J = np.linalg.det(A)
        Lam = J * np.linalg.inv(A)  # Lam[l, m] = J d r_l / d x_m
        scale = np.abs(Lam).max()
        assert np.abs(g.J_q[:, k] - J).max() <= 1e-12 * abs(J)
        assert np.abs(g.Lambda_q[:, :, :, k] - Lam[None]).max() <= 1e-12 * scale
        nJf = Lam.T @ nrst
        assert np.abs(g.nJf[:, :, k] - nJf).max() <= 1e-12 * scale
        assert np.abs(g.J_f[:, k] - np.sqrt((nJf ** 2).sum(0))).max() <= 1e-12 * scale


def test_reference_mapping_matches_host():
    from sse.geometry import apply_reference_mapping
    sd = SD_CASES["tet_p5_curved"]()
    ref, gf = sd.reference_approximation, sd.geometric_factors
    host = apply_reference_mapping(gf, ref)
    dev = apply_reference_mapping(gf, ref, backend="hip")
    assert _rel(dev.Lambda_q, host.Lambda_q) <= TOL


def test_projection_matches_host_and_reproduces_polynomials():
    from sse.geometry import SpatialDiscretization
    from sse.solvers import project_function
    law, sd, ic, form, ode = advection_diffusion_3d_config(M=2, p=5)
    dev_sd = SpatialDiscretization(sd.mesh, sd.reference_approximation, sd.geometric_factors, "hip")
    u_host = project_function(ic, sd)
    u_dev = project_function(ic, dev_sd)
    assert _rel(u_dev, u_host) <= TOL
    # a polynomial of degree 5 on affine tets is reproduced exactly (5 components)
    from sse.mesh import uniform_periodic_mesh
    from sse.geometry import make_spatial_discretization
    ref = sd.reference_approximation
    mesh = uniform_periodic_mesh(ref, ((0.0, 1.0),) * 3, (2, 2, 2))
    flat = make_spatial_discretization(mesh, ref, backend="hip")

    def poly(x, t=0.0):
        x0, x1, x2 = x
        return np.stack([1.0 + x0 ** 5, x1 ** 3 * x2 ** 2 - x0, x0 * x1 * x2 * (x0 + x1),
                         x2 ** 5 - 2.0 * x1 ** 4, x0 ** 2 * x1 ** 2 * x2], axis=-1)

    u = project_function(poly, flat)
    uq = np.einsum("qp,pck->qck", ref.Vdense(), u)
    want = np.transpose(poly(mesh.xyzq), (0, 2, 1))
    assert np.abs(uq - want).max() <= 1e-11 * np.abs(want).max()


def test_projection_reports_a_non_positive_definite_mass_matrix():
    from sse.geometry import GeometricFactors, SpatialDiscretization
    from sse.hip import SseError
    from sse.solvers import project_function
    law, sd, ic, form, ode = advection_diffusion_3d_config(M=2, p=3)
    gf = sd.geometric_factors
    J_q = np.array(gf.J_q, order="F")
    J_q[:, 7] *= -1.0  # an inverted element: V' W J V is negative definite there
    bad = SpatialDiscretization(sd.mesh, sd.reference_approximation,
                                GeometricFactors(J_q, gf.Lambda_q, gf.J_f, gf.nJf, gf.nJq), "hip")
    with pytest.raises(SseError, match="element 7"):
        project_function(ic, bad)


@pytest.mark.parametrize("make", [lambda b: advection_diffusion_3d_config(M=2, p=4, backend=b),
                                  lambda b: density_wave_3d_config(M=2, p=3, backend=b)],
                         ids=["advdiff_tet_p4", "euler_tet_p3"])
def test_residual_of_a_gpu_set_up_problem_matches_the_host_set_up(make):
    """The whole set-up on the GPU (geometry, reference mapping, projection) gives the
    same initial state and residual as the host set-up."""
    import torch
    from sse import semi_discrete_residual
    from sse.time_integration import to_device, to_host
    host, dev = make("host")[4], make("hip")[4]
    assert _rel(dev.u0, host.u0) <= TOL
    out = []
    for ode in (host, dev):
        u = to_device(host.u0)
        du = torch.empty_like(u)
        semi_discrete_residual(du, u, ode.p)
        torch.cuda.synchronize()
        out.append(to_host(du))
    assert _rel(out[1], out[0]) <= TOL
