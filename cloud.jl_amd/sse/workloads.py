"""Discretisations of the reference's test scenarios and of the benchmark
workloads (host-side set-up only)."""

import math

import numpy as np

from .conservation_laws import (BR1, CentralNumericalFlux, EntropyConservativeNumericalFlux,
                                EulerEquations, EulerPeriodicTest, ExactSolution,
                                InitialDataCosine, InitialDataGaussian, InitialDataSine,
                                IsentropicVortex, LaxFriedrichsNumericalFlux,
                                LinearAdvectionDiffusionEquation, LinearAdvectionEquation,
                                ShuVortex)
from .geometry import ChanWilcoxMetrics, make_spatial_discretization
from .mesh import ChanWarping, UniformWarping, uniform_periodic_mesh, warp_mesh
from .reference_approximation import (LGQuadrature, ModalMulti, ModalTensor, NodalTensor,
                                      ReferenceApproximationHex, ReferenceApproximationLine,
                                      ReferenceApproximationTet, ReferenceApproximationTri)
from .solvers import (FluxDifferencingForm, PhysicalOperator, ReferenceOperator,
                      SkewSymmetricMapping, StandardForm, StandardMapping, semidiscretize)


def euler_1d_gauss():
    """test/euler_1d_gauss.jl:1-80 (golden L2 at test/runtests.jl:89-96)."""
    law = EulerEquations(1, 1.4)

    def exact(x, t=0.0):
        rho = 1.0 + 0.2 * np.sin(np.pi * x[0])
        return np.stack([rho, rho * 1.0, 1.0 / (1.4 - 1) + 0.5 * rho], -1)

    ref = ReferenceApproximationLine(NodalTensor(5), volume_quadrature_rule=LGQuadrature(5))
    mesh = uniform_periodic_mesh(ref, (0.0, 2.0), 4)
    sd = make_spatial_discretization(mesh, ref)
    form = FluxDifferencingForm(inviscid_numerical_flux=EntropyConservativeNumericalFlux())
    return law, sd, exact, form, semidiscretize(law, sd, exact, form, (0.0, 2.0))


def euler_3d_hex(M=2, p=4):
    """test/euler_3d.jl:1-90 (golden L2 at test/runtests.jl:131-144): the Euler periodic
    wave on [0, 2]^3, M^3 hexes curved by ChanWarping(1/16), NodalTensor(p) with
    Gauss-Lobatto volume and facet rules, conservative-curl metrics, the EC interface
    flux. Returns (law, sd, exact, form, ode, T, N_t); the reference integrates with
    DP8 at dt = T / N_t."""
    L = 2.0
    T = L
    law = EulerEquations(3, 1.4)
    exact = EulerPeriodicTest(3, 1.4)
    form = FluxDifferencingForm(inviscid_numerical_flux=EntropyConservativeNumericalFlux())
    ref = ReferenceApproximationHex(NodalTensor(p), mapping_degree=p)
    mesh = uniform_periodic_mesh(ref, ((0.0, L),) * 3, (M, M, M))
    mesh = warp_mesh(mesh, ref, ChanWarping(1.0 / 16.0, (L, L, L)))
    sd = make_spatial_discretization(mesh, ref, ChanWilcoxMetrics())
    ode = semidiscretize(law, sd, exact, form, (0.0, T))
    return law, sd, exact, form, ode, T, 25 * M * (p + 1)


def euler_vortex_2d_modal(M=4, form=None, p=3, approx=ModalTensor):
    """test/euler_vortex_2d_modal.jl:1-73 (golden L2 at test/runtests.jl:111-121).
    `approx=ModalMulti, p=4` is the scheme of examples/euler_vortex_2d.ipynb."""
    g, L, Ma = 1.4, 1.0, 0.4
    T = L / Ma
    strength = math.sqrt(2 / (g - 1) * (1 - 0.75 ** (g - 1)))
    law = EulerEquations(2, g)
    exact = IsentropicVortex(g, Ma, 0.0, 0.1, strength, 1.0, (L / 2, L / 2))
    ref = ReferenceApproximationTri(approx(p), mapping_degree=p)
    mesh = uniform_periodic_mesh(ref, ((0.0, L), (0.0, L)), (M, M))
    mesh = warp_mesh(mesh, ref, ChanWarping(1.0 / 16.0, (L, L)))
    sd = make_spatial_discretization(mesh, ref, ChanWilcoxMetrics())
    form = form or FluxDifferencingForm(inviscid_numerical_flux=LaxFriedrichsNumericalFlux())
    return law, sd, exact, form, semidiscretize(law, sd, exact, form, (0.0, T)), T


def vortex_2d_config(M=8, p=4, form=None):
    """2D Euler Shu vortex (strength 5, centre (5,5), free stream rho=u=v=p=1) on an
    affine M x M periodic grid of [0,10]^2, two triangles per square, ModalTensor(p)."""
    law = EulerEquations(2, 1.4)
    ic = ShuVortex(1.4, 5.0, (5.0, 5.0), (1.0, 1.0), (10.0, 10.0))
    ref = ReferenceApproximationTri(ModalTensor(p), mapping_degree=1)
    mesh = uniform_periodic_mesh(ref, ((0.0, 10.0), (0.0, 10.0)), (M, M))
    sd = make_spatial_discretization(mesh, ref)
    form = form or FluxDifferencingForm(inviscid_numerical_flux=LaxFriedrichsNumericalFlux())
    return law, sd, ic, form, semidiscretize(law, sd, ic, form, (0.0, 1.0))


def density_wave_3d_config(M=2, p=4, factor=0.05, form=None, ic_strength=0.5, backend="host",
                           approx=ModalTensor):
    """3D Euler on curved tets: periodic [0,2]^3, M^3 hexes x 6 tets, UniformWarping
    factor 0.05, ModalTensor(p) isoparametric, curl metrics,
    rho = 1 + 0.5 sin(pi (x+y+z)), u=v=w=1, p=1. `backend="hip"` runs the set-up
    (geometric factors, projection of the initial data) on the GPU. `approx=ModalMulti`
    with M=4, p=3 is the scheme of examples/euler_3d.ipynb (dense operators)."""
    law = EulerEquations(3, 1.4)
    ic = EulerPeriodicTest(3, 1.4, ic_strength, 2.0)
    ref = ReferenceApproximationTet(approx(p), mapping_degree=p)
    mesh = uniform_periodic_mesh(ref, ((0.0, 2.0),) * 3, (M, M, M))
    if factor:
        mesh = warp_mesh(mesh, ref, UniformWarping(factor, (2.0, 2.0, 2.0)))
    sd = make_spatial_discretization(mesh, ref, ChanWilcoxMetrics(), backend=backend)
    form = form or FluxDifferencingForm(inviscid_numerical_flux=LaxFriedrichsNumericalFlux())
    return law, sd, ic, form, semidiscretize(law, sd, ic, form, (0.0, 1.0))


def advection_diffusion_1d_modal_multi():
    """test/runtests.jl:13-36 (golden L2 6.988216111882884e-6): ModalMulti(4) on 4
    periodic elements of [0,1], a = 1, b = 5e-2, sine data, StandardForm with
    StandardMapping, LF + BR1, PhysicalOperator; dt = 1/100 to T = 1."""
    law = LinearAdvectionDiffusionEquation((1.0,), 5.0e-2)
    ic = InitialDataSine(1.0, (2 * math.pi,))
    ref = ReferenceApproximationLine(ModalMulti(4))
    mesh = uniform_periodic_mesh(ref, (0.0, 1.0), 4)
    sd = make_spatial_discretization(mesh, ref)
    form = StandardForm(mapping_form=StandardMapping(),
                        inviscid_numerical_flux=LaxFriedrichsNumericalFlux(),
                        viscous_numerical_flux=BR1())
    ode = semidiscretize(law, sd, ic, form, (0.0, 1.0), PhysicalOperator())
    return law, sd, ExactSolution(law, ic), form, ode


def advection_2d_modal_tensor_tri(M=2, p=4, approx=ModalTensor):
    """test/runtests.jl:38-61 (golden L2 0.2660013939427627): ModalTensor(4) Tri,
    mapping degree 4, a = (1,1), sine data, StandardForm skew-symmetric with
    LF(0) (central), ReferenceOperator; DelRey-warped (0.1) 2x2 periodic mesh of
    [0,1]^2 with exact metrics and projected Jacobian; dt = 1/100 to T = 1.
    `approx=NodalMulti, M=4` is the scheme of examples/advection_2d.ipynb."""
    law = LinearAdvectionEquation((1.0, 1.0))
    ic = InitialDataSine(1.0, (2 * math.pi, 2 * math.pi))
    ref = ReferenceApproximationTri(approx(p), mapping_degree=p)
    mesh = warp_mesh(uniform_periodic_mesh(ref, ((0.0, 1.0), (0.0, 1.0)), (M, M)), ref, 0.1)
    sd = make_spatial_discretization(mesh, ref)
    form = StandardForm(mapping_form=SkewSymmetricMapping(),
                        inviscid_numerical_flux=LaxFriedrichsNumericalFlux(0.0))
    ode = semidiscretize(law, sd, ic, form, (0.0, 1.0), ReferenceOperator())
    return law, sd, ExactSolution(law, ic), form, ode


def advection_3d_modal_tensor_tet(M=2, p=4):
    """test/advection_3d.jl (golden L2 0.1876141674772107 under a Jaskowiec-Sukumar
    degree-11 error quadrature): ModalTensor(4) Tet, a = (1,1,1), cosine data,
    StandardForm skew-symmetric with the central flux, ReferenceOperator,
    DelRey-warped (0.1) M^3 x 6 periodic mesh of [0,1]^3, curl metrics."""
    law = LinearAdvectionEquation((1.0, 1.0, 1.0))
    ic = InitialDataCosine(1.0, (2 * math.pi,) * 3)
    ref = ReferenceApproximationTet(ModalTensor(p), mapping_degree=4)
    mesh = warp_mesh(uniform_periodic_mesh(ref, ((0.0, 1.0),) * 3, (M, M, M)), ref, 0.1, 1.0)
    sd = make_spatial_discretization(mesh, ref, ChanWilcoxMetrics())
    form = StandardForm(mapping_form=SkewSymmetricMapping(),
                        inviscid_numerical_flux=CentralNumericalFlux())
    ode = semidiscretize(law, sd, ic, form, (0.0, 1.0), ReferenceOperator())
    return law, sd, ExactSolution(law, ic), form, ode


def advection_diffusion_3d_config(M=4, p=5, b=5.0e-3, sigma=0.1, factor=0.05, backend="host"):
    """3D linear advection-diffusion on curved tets: a = (1, 0.5, 0.25), diffusivity
    b, BR1, periodic [0,1]^3 with M^3 hexes x 6 tets, UniformWarping, ModalTensor(p)
    isoparametric with curl metrics, Gaussian pulse of width sigma at the centre;
    StandardForm (skew-symmetric, LF) with PhysicalOperator. `backend="hip"` runs the
    set-up (geometric factors, reference mapping, projection) on the GPU."""
    law = LinearAdvectionDiffusionEquation((1.0, 0.5, 0.25), b)
    ic = InitialDataGaussian(1.0, sigma, (0.5, 0.5, 0.5))
    ref = ReferenceApproximationTet(ModalTensor(p), mapping_degree=p)
    mesh = uniform_periodic_mesh(ref, ((0.0, 1.0),) * 3, (M, M, M))
    if factor:
        mesh = warp_mesh(mesh, ref, UniformWarping(factor, (1.0, 1.0, 1.0)))
    sd = make_spatial_discretization(mesh, ref, ChanWilcoxMetrics(), backend=backend)
    form = StandardForm(mapping_form=SkewSymmetricMapping(),
                        inviscid_numerical_flux=LaxFriedrichsNumericalFlux(),
                        viscous_numerical_flux=BR1())
    ode = semidiscretize(law, sd, ic, form, (0.0, 1.0), PhysicalOperator())
    return law, sd, ExactSolution(law, ic), form, ode


def batched_vortex_problems(n=256, seed=20261202, M_range=(4, 32), p_range=(2, 6), cfl=0.1,
                            progress=None):
    """n independent 2D Euler Shu-vortex problems of uneven size: periodic [0,10]^2
    with M x M squares split into triangles (M uniform in M_range), ModalTensor(p)
    with p uniform in p_range, vortex strength uniform in [1,5] and centre uniform
    in the domain (seeded), free stream (1,1), EC + LF flux differencing.
    Each problem's time step is dt = cfl * h / ((2p + 1) * 2.6), h = 10 / M, with
    2.6 >= |v| + c of the free stream. Problems with the same (M, p) share their
    mesh and geometry. Returns a list of (ode, dt, (M, p))."""
    rng = np.random.default_rng(seed)
    Ms = rng.integers(M_range[0], M_range[1] + 1, n)
    ps = rng.integers(p_range[0], p_range[1] + 1, n)
    strengths = rng.uniform(1.0, 5.0, n)
    centres = rng.uniform(0.0, 10.0, (n, 2))
    law = EulerEquations(2, 1.4)
    form = FluxDifferencingForm(inviscid_numerical_flux=LaxFriedrichsNumericalFlux())
    cache = {}
    out = []
    for i in range(n):
        M, p = int(Ms[i]), int(ps[i])
        if (M, p) not in cache:
            ref = ReferenceApproximationTri(ModalTensor(p), mapping_degree=1)
            mesh = uniform_periodic_mesh(ref, ((0.0, 10.0), (0.0, 10.0)), (M, M))
            cache[(M, p)] = make_spatial_discretization(mesh, ref)
        sd = cache[(M, p)]
        ic = ShuVortex(1.4, float(strengths[i]), (float(centres[i, 0]), float(centres[i, 1])), (1.0, 1.0),
                       (10.0, 10.0))
        ode = semidiscretize(law, sd, ic, form, (0.0, 1.0))
        out.append((ode, cfl * (10.0 / M) / ((2 * p + 1) * 2.6), (M, p)))
        if progress and (i + 1) % 32 == 0:
            progress(f"batched problems: {i + 1}/{n} set up")
    return out
