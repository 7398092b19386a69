"""sse_functional_noalloc (the allocation-free variant of sse_functional, for callers
that cannot call hipMalloc) against sse_functional on the same handles and states:
the same kernels and fixed-order sums, so the values agree bit for bit. Two calls in a
row share the library's one workspace, the second on a smaller mesh; an oversized
handle is rejected with a clear error."""

import ctypes as C

import numpy as np
import pytest

from cases import density_wave_3d_config, vortex_2d_config

pytestmark = pytest.mark.gpu


def _states(ode, seed):
    rng = np.random.default_rng(seed)
    u = np.asfortranarray(ode.u0 * (1.0 + 1e-3 * rng.standard_normal(ode.u0.shape)))
    v = np.asfortranarray(rng.standard_normal(ode.u0.shape))
    return u, v


def _both(solver, u, v):
    """{(which, rate): (sse_functional values, sse_functional_noalloc values)}"""
    import torch
    from sse import abi
    from sse.hip import check
    from sse.time_integration import to_device
    lib, h = solver.lib, solver.handle
    N_c = solver.size()[1]
    ud, vd = to_device(u), to_device(v)
    out = {}
    for which in (abi.SSE_MASS, abi.SSE_ENERGY, abi.SSE_ENTROPY):
        n = 1 if which == abi.SSE_ENTROPY else N_c
        for rate in (False, True):
            want = np.zeros(n)
            check(lib, h, lib.sse_functional(h, ud.data_ptr(), vd.data_ptr() if rate else None, which,
                                             abi.SSE_DEVICE, want.ctypes.data_as(C.POINTER(C.c_double))))
            got_d = torch.full((n + 1,), -7.0, dtype=torch.float64, device="cuda")
            check(lib, h, lib.sse_functional_noalloc(h, ud.data_ptr(), vd.data_ptr() if rate else None, which,
                                                     got_d.data_ptr()))
            check(lib, h, lib.sse_synchronize(h))
            got = got_d.cpu().numpy()
            assert got[n] == -7.0, "wrote past the values it returns"
            out[(which, rate)] = (want, got[:n])
    return out


def _check(pairs):
    for key, (want, got) in pairs.items():
        assert np.all(np.isfinite(want)), key
        np.testing.assert_array_equal(got, want, err_msg=str(key))


def test_noalloc_functionals_match_sse_functional_twice_then_smaller():
    # first call: 48 curved tets, 5 variables; second: 32 triangles, 4 variables
    big = density_wave_3d_config(M=2, p=4)
    small = vortex_2d_config(M=4, p=4)
    assert big[1].N_e * 5 > small[1].N_e * 4
    for (law, sd, ic, form, ode), seed in ((big, 1), (small, 2)):
        u, v = _states(ode, seed)
        _check(_both(ode.p, u, v))
    # the larger mesh again after the smaller one (the workspace keeps the last call's contents)
    u, v = _states(big[4], 3)
    _check(_both(big[4].p, u, v))


def test_noalloc_rejects_a_mesh_beyond_its_workspace():
    import torch
    from sse import abi
    from sse.hip import SseError, check
    from sse.workloads import (EulerEquations, FluxDifferencingForm, LGQuadrature, NodalTensor,
                               ReferenceApproximationLine, make_spatial_discretization, semidiscretize,
                               uniform_periodic_mesh)
    N_e = abi.SSE_FUNCTIONAL_NOALLOC_MAX_DOUBLES // 3 + 1  # 1D Euler: N_e * 3 + 3 > the workspace
    law = EulerEquations(1, 1.4)
    ref = ReferenceApproximationLine(NodalTensor(1), volume_quadrature_rule=LGQuadrature(1))
    mesh = uniform_periodic_mesh(ref, (0.0, 1.0), N_e)
    sd = make_spatial_discretization(mesh, ref)

    def ic(x, t=0.0):
        rho = 1.0 + 0.1 * np.sin(2 * np.pi * x[0])
        return np.stack([rho, rho, 2.5 + 0.5 * rho], -1)

    ode = semidiscretize(law, sd, ic, FluxDifferencingForm(), (0.0, 1.0))
    solver = ode.p
    lib, h = solver.lib, solver.handle
    u = torch.zeros(solver.size()[::-1], dtype=torch.float64, device="cuda")
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    with pytest.raises(SseError, match="exceeds its workspace") as e:
        check(lib, h, lib.sse_functional_noalloc(h, u.data_ptr(), None, abi.SSE_MASS, out.data_ptr()))
    assert e.value.code == abi.SSE_ERR_INVALID_ARGUMENT
