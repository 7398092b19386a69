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


def test_dp8_tableau_order_conditions():
    # DP8's propagating weights reproduce exp(z) through z^8 on y' = z y (the linear
    # order conditions b' A^(k-1) 1 = 1/k!), and the method converges at 8th order
    # on a nonlinear ODE (y' = -y^2, y(0) = 1, exact 1/(1+t)); a changed digit in any
    # coefficient breaks one or the other
    import math
    from sse.time_integration import DP8
    alg = DP8()
    assert alg.stages == 12 and alg.A.shape == (12, 12)
    assert np.all(np.triu(alg.A) == 0.0)
    v = np.ones(12)
    for k in range(1, 9):
        assert abs(alg.b @ v - 1.0 / math.factorial(k)) < 1e-14 * math.factorial(8), k
        v = alg.A @ v
    np.testing.assert_allclose(alg.A.sum(axis=1), alg.c, rtol=0, atol=1e-14)

    def integrate(n):
        h, y = 1.0 / n, 1.0
        for _ in range(n):
            k = np.zeros(12)
            for i in range(12):
                yi = y + h * (alg.A[i, :i] @ k[:i])
                k[i] = -yi * yi
            y += h * (alg.b @ k)
        return abs(y - 0.5)

    e1, e2 = integrate(4), integrate(8)
    assert 7.5 < np.log2(e1 / e2) < 9.5, (e1, e2)


def test_euler_3d_nodal_tensor_hex_golden():
    # test/runtests.jl:131-144 (scenario test/euler_3d.jl): NodalTensor(4) Hex, curved
    # by ChanWarping(1/16), conservative-curl metrics, EC interface flux, DP8 at
    # dt = T / 250; the L2 errors at T, mass conservation E[end] - E[1] and the largest
    # entropy rate over the steps save_callback writes (every floor(250 / 50) = 5)
    from oracle import functionals as F
    from sse.time_integration import DP8
    from sse.workloads import euler_3d_hex
    law, sd, exact, form, ode, T, N_t = euler_3d_hex()
    alg = DP8()
    dt, every = T / N_t, N_t // 50
    u, rates = ode.u0, []
    for step in range(0, N_t + 1, every):
        if step:
            u = O.erk_steps(ode.p.arrays, u, dt, every, alg.A, alg.b)
        rates.append(F.entropy_rate(sd, law, u, O.residual(ode.p.arrays, u))[0])
    err = l2_error(sd, u, exact, T)
    gold = [0.18342164491797003, 0.1834216449179776, 0.18342164491796725, 0.18342164491796784,
            0.2751324673769553]
    np.testing.assert_allclose(err, gold, rtol=0, atol=TOL)
    np.testing.assert_allclose(F.mass(sd, u) - F.mass(sd, ode.u0), 0.0, atol=TOL)
    assert max(abs(r) for r in rates) < TOL
