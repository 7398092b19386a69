"""The three larger workloads at their configured sizes, HIP path against the oracle.

* Batched solves: 256 independent 2D Euler vortex problems (M x M squares split into
  triangles, M in 4..32, ModalTensor p in 2..6, seeded), advanced in grouped launches
  with the HIP-graph RK step; per p group the largest and the smallest problem against
  the oracle's CK2N54 (<= 1e-10 after 20 steps).
* Adjoint: 3D Euler p=3 on 12^3 hexes x 6 curved tets (10,368 elements), checkpoints
  every 20 steps over 40 steps; dJ along two seeded directions against the oracle's
  forward-mode derivative through the same steps (<= 1e-10), and against central
  differences (h = 1e-6) on four more (<= 1e-6).
* Advection-diffusion: BR1 on 24^3 hexes x 6 curved tets (82,944 elements), p=5: one
  residual against the oracle (<= 1e-12), and the 64-member diffusivity sweep for 2
  steps with its first and last members against the oracle's CK2N54 (<= 1e-10).

Tolerances follow the reference's own test bar (test/runtests.jl:12, atol 1e-10) and
the residual parity bar of 1e-12 (max-norm error relative to max |dudt|).
"""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_config3_256_batched_problems_against_oracle():
    from sse.batch import BatchedSolver
    from sse.time_integration import to_device, to_host
    from sse.workloads import batched_vortex_problems
    probs = batched_vortex_problems(n=256)  # seed 20261202, M in 4..32, p in 2..6
    assert len(probs) == 256
    batch = BatchedSolver([ode.p for ode, dt, key in probs])
    states = [to_device(ode.u0) for ode, dt, key in probs]
    dts = [dt for ode, dt, key in probs]
    batch.rk_steps(states, dts, 12)  # graph captured, replayed
    batch.rk_steps(states, dts, 8)   # and replayed again in a second call: 20 steps
    by_p = {}
    for i, (ode, dt, (M, p)) in enumerate(probs):
        by_p.setdefault(p, []).append((M, i))
    assert sorted(by_p) == [2, 3, 4, 5, 6]
    checked = []
    for p, members in sorted(by_p.items()):
        members.sort()
        for M, i in {members[0], members[-1]}:  # the smallest and the largest mesh of the group
            ode, dt, key = probs[i]
            ref = O.rk_steps(ode.p.arrays, ode.u0, dt, 20)
            got = to_host(states[i])
            assert _rel(got, ref) <= 1e-10, (key, i, _rel(got, ref))
            checked.append(key)
    assert max(M for M, p in checked) == 32, checked
    for x in states:
        assert bool(x.isfinite().all())


def test_config4_adjoint_12cube_p3_checkpointed_against_oracle_and_fd():
    from oracle import functionals as F
    from sse.adjoint import rk_adjoint
    from sse.analysis import EntropyConservationAnalysis
    from sse.conservation_laws import conservative_to_entropy
    from sse.time_integration import rk_steps, to_device
    from sse.workloads import density_wave_3d_config
    M, p, nsteps, every = 12, 3, 40, 20
    law, sd, ic, form, ode = density_wave_3d_config(M=M, p=p, factor=0.05, ic_strength=0.5)
    solver = ode.p
    assert solver.size() == (20, 5, 10368)
    rng = np.random.default_rng(20261202)
    u0 = np.asfortranarray(ode.u0 + 1e-3 * rng.standard_normal(ode.u0.shape))
    dt = 0.2 * (2.0 / M) / ((2 * p + 1) * 2.9)
    J, g, glam = rk_adjoint(solver, u0, dt, nsteps, checkpoint_every=every)
    ref = sd.reference_approximation
    V = ref.Vdense()
    WJ = ref.W[:, None] * sd.geometric_factors.J_q
    for t, dlam in enumerate((0.5, 0.0)):  # two seeded directions against forward mode
        du = np.asfortranarray(np.random.default_rng(100 + t).standard_normal(u0.shape))
        uN, duN = O.rk_jvp(solver.arrays, u0, du, dt, nsteps, dlam)
        if t == 0:
            JN = F.entropy(sd, law, uN)[0]
            assert abs(J - JN) <= 1e-12 * abs(JN), (J, JN)
        wq = conservative_to_entropy(law, np.einsum("qp,pck->qkc", V, uN))
        dJ_fwd = float(np.einsum("qk,qkc,qp,pck->", WJ, wq, V, duN))
        dJ_adj = float(np.sum(g * du)) + glam * dlam
        assert abs(dJ_fwd - dJ_adj) <= 1e-10 * (abs(dJ_fwd) + 1e-3), (t, dJ_fwd, dJ_adj)
    ent = EntropyConservationAnalysis(solver)

    def J_of(x):
        y = to_device(x)
        rk_steps(solver, y, dt, nsteps)
        return float(ent.evaluate_conservation(y)[0])

    h = 1e-6
    for t in range(4):  # central differences on four more seeded directions
        d = np.asfortranarray(np.random.default_rng(200 + t).standard_normal(u0.shape))
        fd = (J_of(u0 + h * d) - J_of(u0 - h * d)) / (2 * h)
        ad = float(np.sum(g * d))
        assert abs(fd - ad) <= 1e-6 * abs(ad), (t, fd, ad)


@pytest.fixture(scope="module")
def advdiff_24cube_p5():
    from sse.workloads import advection_diffusion_3d_config
    law, sd, ex, form, ode = advection_diffusion_3d_config(M=24, p=5, b=5e-3, sigma=0.1, backend="hip")
    assert ode.p.size() == (56, 1, 82944)
    yield law, ode
    ode.p.close()


def test_config5_br1_residual_82944_tets_against_oracle(advdiff_24cube_p5):
    import torch
    from sse import semi_discrete_residual
    from sse.time_integration import to_device, to_host
    law, ode = advdiff_24cube_p5
    u = to_device(ode.u0)
    dudt = torch.empty_like(u)
    semi_discrete_residual(dudt, u, ode.p)
    torch.cuda.synchronize()
    ref = O.residual(ode.p.arrays, ode.u0)
    assert _rel(to_host(dudt), ref) <= 1e-12, _rel(to_host(dudt), ref)


def test_config5_64_member_diffusivity_sweep_against_oracle(advdiff_24cube_p5):
    from sse.batch import BatchedSolver
    from sse.conservation_laws import LinearAdvectionDiffusionEquation
    from sse.time_integration import to_device, to_host
    law, ode = advdiff_24cube_p5
    bvals = np.random.default_rng(20261202).uniform(1e-3, 1e-2, 64)
    members = [ode.p.with_law(LinearAdvectionDiffusionEquation(law.a, float(b))) for b in bvals]
    batch = BatchedSolver(members)
    assert batch.describe().count("one mesh") == 1, batch.describe()[:300]
    dt, n = 1e-4, 2
    u = to_device(ode.u0)
    states = [u.clone() for _ in members]
    del u
    batch.rk_steps(states, [dt] * len(members), n)
    for i in (0, len(members) - 1):
        ref = O.rk_steps(members[i].arrays, ode.u0, dt, n)
        got = to_host(states[i])
        assert _rel(got, ref) <= 1e-10, (i, bvals[i], _rel(got, ref))
    batch.close()
