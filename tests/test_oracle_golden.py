"""The oracle against the reference's own golden values (test/runtests.jl).

Both scenarios are whole trajectories (mesh -> discretisation -> semidiscretize
-> 1000 CK2N54 steps -> L2 error), compared at the reference's tolerance
(atol 1e-10, test/runtests.jl:12)."""

import numpy as np

from cases import (advection_2d_modal_tensor_tri, advection_diffusion_1d_modal_multi,
                   euler_1d_gauss, euler_vortex_2d_modal, l2_error)
from oracle import oracle as O

TOL = 1.0e-10


def test_euler_1d_gauss_collocation_golden():
    # test/runtests.jl:89-96
    law, sd, exact, form, ode = euler_1d_gauss()
    uT = O.rk_steps(ode.p.arrays, ode.u0, 2.0 / 1000, 1000)
    err = l2_error(sd, uT, exact, 2.0)
    gold = [3.5808560177567635e-5, 5.2129828619609155e-5, 0.00012637647535378534]
    np.testing.assert_allclose(err, gold, rtol=0, atol=TOL)


def test_euler_vortex_2d_modal_tensor_tri_lf_golden():
    # test/runtests.jl:111-121: the ModalTensor Tri flux-differencing path
    law, sd, exact, form, ode, T = euler_vortex_2d_modal()
    uT = O.rk_steps(ode.p.arrays, ode.u0, T / 1000, 1000)
    err = l2_error(sd, uT, exact, T)
    gold = [0.015568197027072704, 0.040539693811761104, 0.04060141777050208, 0.043960971468832745]
    np.testing.assert_allclose(err, gold, rtol=0, atol=TOL)


def test_euler_vortex_2d_modal_conservation():
    # test/runtests.jl:120: conservation E[end] - E[1] ~ 0
    law, sd, exact, form, ode, T = euler_vortex_2d_modal()
    uT = O.rk_steps(ode.p.arrays, ode.u0, T / 1000, 100)
    ref = sd.reference_approximation
    WJ = ref.W[:, None] * sd.geometric_factors.J_q

    def mass(u):
        return np.einsum("qk,qp,pck->c", WJ, ref.V, u)

    np.testing.assert_allclose(mass(uT) - mass(ode.u0), 0.0, atol=TOL)


def test_advection_diffusion_1d_modal_multi_br1_golden():
    # test/runtests.jl:13-36: StandardForm, StandardMapping, LF + BR1, PhysicalOperator
    law, sd, exact, form, ode = advection_diffusion_1d_modal_multi()
    uT = O.rk_steps(ode.p.arrays, ode.u0, 1.0 / 100, 100)
    err = l2_error(sd, uT, exact, 1.0)
    np.testing.assert_allclose(err, [6.988216111882884e-6], rtol=0, atol=TOL)


def test_advection_2d_modal_tensor_tri_energy_conservative_golden():
    # test/runtests.jl:38-61: StandardForm, skew-symmetric, central, ReferenceOperator,
    # DelRey-warped isoparametric degree-4 mesh with exact metrics
    law, sd, exact, form, ode = advection_2d_modal_tensor_tri()
    uT = O.rk_steps(ode.p.arrays, ode.u0, 1.0 / 100, 100)
    err = l2_error(sd, uT, exact, 1.0)
    np.testing.assert_allclose(err, [0.2660013939427627], rtol=0, atol=TOL)
