#!/usr/bin/env python3
"""Benchmark: DOF-residuals/s of the FP64 semi-discrete residual on one MI355X.

Workload (one "step" = one residual evaluation over the whole mesh): 3D
compressible Euler on curved tetrahedra -- periodic cube [0,2]^3, 16^3 hexes x 6
tets (24,576 elements), UniformWarping factor 0.05, ModalTensor p = 4
(N_p = 35, N_q = 125, N_f = 100, N_c = 5), isoparametric mapping with
conservative-curl metrics, density wave rho = 1 + 0.5 sin(pi (x+y+z)),
u = v = w = 1, p = 1, FluxDifferencingForm with the entropy-conservative
two-point flux and Lax-Friedrichs interface dissipation. State arrays are
resident in HBM for the timed region.

The same line carries the other two metrics of the headline, each measured on its
own workload after the residual: "batched" (256 uneven 2D Euler vortex problems,
200 RK steps, batched and graphed against a loop of per-problem GPU solves) and
"adjoint" (3D Euler p=3 on 10,368 curved tets, 200 RK steps, the whole entropy
gradient against one forward solve, and the reverse sweep alone).

Prints one JSON line (rank 0). Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "cloud.jl_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 (vector and matrix)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(M: int, p: int):
    from sse.workloads import density_wave_3d_config
    return density_wave_3d_config(M=M, p=p, factor=0.05, ic_strength=0.5)


def _solver(args, law, sd, form, ode):
    """The workload's solver; with --deterministic, one whose sums run in a fixed order."""
    if not getattr(args, "deterministic", False):
        return ode.p
    from sse.solvers import MI355X, Solver
    ode.p.close()
    return Solver(law, sd, form, parallelism=MI355X(deterministic=True))


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(arrays, u0, dof, target_s: float):
    """The CPU oracle (C++ restatement of the reference algorithm, OpenMP over elements
    like the reference's Threaded()) on the same workload: all the threads this job
    may use (the cores affine to the process, at most 16, the CPU share of one GPU
    on the GPU hosts, which set OMP_NUM_THREADS to 16) and one thread."""
    from oracle import oracle as O
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = max(1, min(16, visible, omp) if omp > 0 else min(16, visible))

    def rate(nthreads, budget):
        O.set_threads(nthreads)
        t0 = time.perf_counter()
        O.residual(arrays, u0)
        one = time.perf_counter() - t0
        nrep = max(1, int(budget / max(one, 1e-3)))
        t0 = time.perf_counter()
        O.residual_repeat(arrays, u0, nrep)
        el = time.perf_counter() - t0
        return dof * nrep / el, nrep, el

    v_all, n_all, el_all = rate(threads, target_s)
    v_one, n_one, el_one = rate(1, target_s)
    O.set_threads(threads)
    return {"value": v_all, "unit": "DOF-residuals/s", "cores": threads, "kind": "port",
            "value_1thread": v_one, "cpu_model": _cpu_model(), "host_cpus_visible": visible,
            "sample": f"{n_all} full residual evaluations of the same 24,576-element workload on {threads} "
                      f"OpenMP threads ({el_all:.1f} s), and {n_one} on one thread ({el_one:.1f} s); "
                      "C++ restatement of the reference algorithm (not the Julia reference)"}


def load_traffic(workload: str, kernel: str):
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(workload, {}).get(kernel)
    except (OSError, ValueError):
        return None


def bench_batched(args):
    """256 independent 2D Euler vortex problems of uneven size (M in 4..32, p in 2..6),
    200 CK2N54 steps each: one batched, graph-captured run against a loop of
    per-problem GPU solves. Returns the JSON line."""
    import torch

    from sse.batch import BatchedSolver
    from sse.time_integration import rk_steps, to_device
    from sse.workloads import batched_vortex_problems

    nsteps = args.batch_steps
    t0 = time.perf_counter()
    probs = batched_vortex_problems(n=args.batch_problems, progress=log)
    dof = sum(int(np.prod(ode.p.size())) for ode, _, _ in probs)
    log(f"prep: {time.perf_counter() - t0:.1f} s, {len(probs)} problems, {dof} DOF")
    dts = [dt for _, dt, _ in probs]
    # loop of plain per-problem solves (handles created and warmed up first)
    loop_states = [to_device(ode.u0) for ode, _, _ in probs]
    for (ode, dt, _), u in zip(probs, loop_states):
        rk_steps(ode.p, u, dt, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for (ode, dt, _), u in zip(probs, loop_states):  # one solve after the other, each to completion
        rk_steps(ode.p, u, dt, nsteps)
        ode.p.lib.sse_synchronize(ode.p.handle)
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t0
    log(f"loop of {len(probs)} solves: {t_loop:.3f} s")
    batch = BatchedSolver([ode.p for ode, _, _ in probs])
    log(batch.describe()[:400])
    states = [to_device(ode.u0) for ode, _, _ in probs]
    batch.rk_steps(states, dts, 2)  # warm-up: graph capture
    states = [to_device(ode.u0) for ode, _, _ in probs]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.rk_steps(states, dts, nsteps)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    log(f"batched: {t_batch:.3f} s")
    # both paths advanced the same problems from the same states (loop: 1 + nsteps steps)
    line = {
        "metric": "batched vs loop speedup (256 2D Euler vortex problems, 200 RK steps)",
        "value": t_loop / t_batch, "unit": "x", "n_gpus": 1, "steps": nsteps, "warmup": 1,
        "ms_per_step": 1e3 * t_batch / nsteps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "256 independent 2D Euler Shu-vortex problems, periodic [0,10]^2, "
                               "M x M squares x 2 tris with M in 4..32, ModalTensor p in 2..6, "
                               "FluxDifferencingForm EC + LF, CK2N54", "problems": len(probs), "DOF": dof,
                   "t_loop_s": t_loop, "t_batched_s": t_batch,
                   "solves_per_s_batched": len(probs) / t_batch, "solves_per_s_loop": len(probs) / t_loop,
                   "rk_steps_per_s_batched": nsteps / t_batch,
                   "dof_residuals_per_s_batched": dof * 5 * nsteps / t_batch},
    }
    return line


def bench_adjoint(args):
    """Discrete adjoint of 200 CK2N54 steps, 3D Euler p=3 on the curved-tet mesh at
    12^3 hexes x 6 tets (10,368 elements): gradient of the final entropy integral
    w.r.t. the initial state and the Lax-Friedrichs coefficient, checkpoints every
    20 steps. Reports adjoint/forward cost and the central-FD agreement on 32 seeded
    directions. Returns the JSON line; its value is the whole gradient's cost (the
    adjoint's own checkpointed forward sweep plus the reverse sweep with recomputation)
    over one forward solve, and the reverse sweep alone is reported beside it."""
    import torch

    from sse.adjoint import rk_adjoint
    from sse.time_integration import rk_steps, to_device, to_host

    M, p, nsteps, every = args.adj_mesh, 3, args.adj_steps, 20
    t0 = time.perf_counter()
    law, sd, ic, form, ode = build_workload(M, p)
    solver = _solver(args, law, sd, form, ode)
    rng = np.random.default_rng(20261202)
    u0 = np.asfortranarray(ode.u0 + 1e-3 * rng.standard_normal(ode.u0.shape))
    dt = 0.2 * (2.0 / M) / ((2 * p + 1) * 2.9)
    N_p, N_c, N_e = solver.size()
    log(f"prep: {time.perf_counter() - t0:.1f} s, N_e={N_e}, dt={dt:.3e}")
    ud0 = to_device(u0)
    # warm-up both paths
    u = ud0.clone()
    rk_steps(solver, u, dt, 2)
    rk_adjoint(solver, ud0, dt, 2, every)
    torch.cuda.synchronize()
    u = ud0.clone()
    t0 = time.perf_counter()
    rk_steps(solver, u, dt, nsteps)
    solver.lib.sse_synchronize(solver.handle)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    J, g, glam = rk_adjoint(solver, ud0, dt, nsteps, every)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    log(f"forward {t_fwd:.3f} s, adjoint (forward sweep + reverse with recompute) {t_adj:.3f} s")
    g = to_host(g)
    # central finite differences of J on seeded random directions, h = 1e-6
    from sse.analysis import EntropyConservationAnalysis
    ent = EntropyConservationAnalysis(solver)

    def J_of(x, lam_shift=0.0):
        hl = solver.arrays.desc.half_lambda
        if lam_shift:
            solver.close()
            solver.arrays.desc.half_lambda = hl + 0.5 * lam_shift
        y = to_device(x)
        rk_steps(solver, y, dt, nsteps)
        val = float(ent.evaluate_conservation(y)[0])
        if lam_shift:
            solver.close()
            solver.arrays.desc.half_lambda = hl
        return val

    h = 1e-6
    errs = []
    for t in range(args.adj_directions):
        d = np.asfortranarray(rng.standard_normal(u0.shape))  # O(1) entries: each coefficient moves by ~h
        fd = (J_of(u0 + h * d) - J_of(u0 - h * d)) / (2 * h)
        ad = float(np.sum(g * d))
        errs.append(abs(fd - ad) / max(abs(ad), 1e-300))
    fd_lam = (J_of(u0, h) - J_of(u0, -h)) / (2 * h)
    lam_err = abs(fd_lam - glam) / max(abs(glam), 1e-300)
    fd_max = max(errs) if errs else None
    fd_txt = f"{fd_max:.2e}" if errs else "n/a"
    log(f"FD check: max rel err over {len(errs)} directions {fd_txt}; dJ/dλ {glam:.6e} vs FD {fd_lam:.6e}")
    line = {
        "metric": "adjoint/forward cost (3D Euler p=3 curved tets, 200 RK steps, entropy objective)",
        "value": t_adj / t_fwd, "unit": "x", "n_gpus": 1, "steps": nsteps, "warmup": 2,
        "ms_per_step": 1e3 * t_adj / nsteps, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D Euler, curved tets, {M}^3 hexes x 6 tets, ModalTensor p={p}, "
                               "EC + LF flux differencing, CK2N54, J = entropy integral at the final time, "
                               f"checkpoints every {every} steps", "N_e": N_e, "dt": dt,
                   "t_forward_s": t_fwd, "t_adjoint_total_s": t_adj,
                   "total_gradient_over_forward": t_adj / t_fwd,
                   "reverse_only_over_forward": (t_adj - t_fwd) / t_fwd,
                   "J": J, "dJ_dlambda": glam,
                   "fd_max_rel_err": fd_max, "fd_directions": len(errs), "fd_lambda_rel_err": lam_err},
    }
    return line


def bench_advdiff(args):
    """3D linear advection-diffusion with BR1 on curved tets, 24^3 hexes x 6 tets
    (82,944 elements), ModalTensor p=5, a = (1, 0.5, 0.25), b = 5e-3, Gaussian pulse
    of width 0.1: DOF-residuals/s of the second-order StandardForm residual, plus a
    batched sweep of independent diffusivities. Prints one JSON line."""
    import torch

    from sse import semi_discrete_residual
    from sse.time_integration import to_device
    from sse.workloads import advection_diffusion_3d_config

    t0 = time.perf_counter()
    law, sd, ex, form, ode = advection_diffusion_3d_config(M=args.ad_mesh, p=5, b=5e-3, sigma=0.1, backend="hip")
    solver = ode.p
    N_p, N_c, N_e = solver.size()
    dof = N_p * N_c * N_e
    log(f"prep: {time.perf_counter() - t0:.1f} s, N_e={N_e}, DOF={dof}; kernels: {solver.describe()}")
    u = to_device(ode.u0)
    dudt = torch.empty_like(u)
    for _ in range(args.warmup):
        semi_discrete_residual(dudt, u, solver)
    solver.lib.sse_synchronize(solver.handle)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        semi_discrete_residual(dudt, u, solver)
    solver.lib.sse_synchronize(solver.handle)
    el = time.perf_counter() - t0
    sweep = advdiff_sweep(args, law, solver, u) if args.ad_members > 0 else None
    line = {
        "metric": "DOF-residuals/s (3D advection-diffusion BR1 p=5 FP64)", "value": dof * args.steps / el,
        "unit": "DOF-residuals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D linear advection-diffusion, BR1, curved tets, {args.ad_mesh}^3 hexes x 6 tets, "
                               "ModalTensor p=5, StandardForm skew-symmetric + LF, PhysicalOperator",
                   "N_e": N_e, "DOF": dof, "kernels": solver.describe()},
    }
    if sweep:
        line["sweep"] = sweep
    print(json.dumps(line), flush=True)


def bench_multi(args):
    """The dense-operator (ModalMulti) schemes of the reference's recorded notebook
    timings: Euler on M=4 curved tets, ModalMulti p=3, EC interface flux, curl
    metrics (N_e=384; examples/euler_3d.ipynb), and the isentropic vortex on M=4
    warped triangles, ModalMulti p=4, Lax-Friedrichs (N_e=32;
    examples/euler_vortex_2d.ipynb), linear advection on M=4 curved triangles,
    NodalMulti p=4, skew-symmetric StandardForm with the central flux
    (examples/advection_2d.ipynb); then the tet scheme at M=16 (24,576 elements)
    for throughput. Time per residual on the GPU beside the CPU oracle on one
    thread (the reference's Serial() runs). Prints one JSON line."""
    import torch

    from sse import semi_discrete_residual
    from sse.conservation_laws import EntropyConservativeNumericalFlux
    from sse.reference_approximation import ModalMulti, NodalMulti
    from sse.solvers import FluxDifferencingForm
    from sse.time_integration import to_device
    from sse.workloads import advection_2d_modal_tensor_tri, density_wave_3d_config, euler_vortex_2d_modal

    ec = FluxDifferencingForm(inviscid_numerical_flux=EntropyConservativeNumericalFlux())
    cases = [("tet_p3_M4_ec", lambda: density_wave_3d_config(M=4, p=3, approx=ModalMulti, form=ec), 31.8e-3,
              "examples/euler_3d.ipynb:141-143"),
             ("tri_p4_M4_lf", lambda: euler_vortex_2d_modal(M=4, p=4, approx=ModalMulti)[:5], 541e-6,
              "examples/euler_vortex_2d.ipynb:1274-1276"),
             ("adv_tri_p4_M4_nodal", lambda: advection_2d_modal_tensor_tri(M=4, p=4, approx=NodalMulti), 68.7e-3,
              "examples/advection_2d.ipynb:1260 (Threaded)"),
             ("tet_p3_M16_ec", lambda: density_wave_3d_config(M=16, p=3, approx=ModalMulti, form=ec), None, None)]
    if args.multi_cases:
        cases = [c for c in cases if c[0] in args.multi_cases.split(",")]
    out = {}
    for name, make, ref_s, where in cases:
        law, sd, ic, form, ode = make()
        solver = _solver(args, law, sd, form, ode)
        N_p, N_c, N_e = solver.size()
        dof = N_p * N_c * N_e
        u = to_device(ode.u0)
        dudt = torch.empty_like(u)
        for _ in range(args.warmup):
            semi_discrete_residual(dudt, u, solver)
        solver.lib.sse_synchronize(solver.handle)
        nrep = max(args.steps, 200 if N_e < 1000 else 20)
        t0 = time.perf_counter()
        for _ in range(nrep):
            semi_discrete_residual(dudt, u, solver)
        solver.lib.sse_synchronize(solver.handle)
        gpu_s = (time.perf_counter() - t0) / nrep
        rec = {"N_e": N_e, "N_p": N_p, "N_q": sd.reference_approximation.N_q, "DOF": dof,
               "gpu_ms_per_residual": 1e3 * gpu_s, "gpu_dof_res_per_s": dof / gpu_s, "kernels": solver.describe()}
        if ref_s is not None:
            rec["reference_recorded_ms"] = 1e3 * ref_s
            rec["reference_recorded_at"] = where
            from oracle import oracle as O
            O.set_threads(1)
            t0 = time.perf_counter()
            O.residual(solver.arrays, ode.u0)
            one = time.perf_counter() - t0
            reps = max(1, int(2.0 / max(one, 1e-4)))
            t0 = time.perf_counter()
            O.residual_repeat(solver.arrays, ode.u0, reps)
            rec["cpu_oracle_1thread_ms"] = 1e3 * (time.perf_counter() - t0) / reps
        log(name, json.dumps(rec))
        out[name] = rec
        del solver, ode, u, dudt
    big = out[list(out)[-1]]
    line = {
        "metric": "DOF-residuals/s (3D Euler ModalMulti p=3 FP64, dense operators)", "value": big["gpu_dof_res_per_s"],
        "unit": "DOF-residuals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": big["gpu_ms_per_residual"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "3D Euler density wave on curved tets, ModalMulti p=3 (collapsed degree-6 "
                               "quadrature), EC flux, curl metrics; notebook sizes and 16^3 hexes x 6 tets"},
        "cases": out,
    }
    print(json.dumps(line), flush=True)


def bench_setup(args):
    """Set-up of the 3D advection-diffusion problem (24^3 hexes x 6 curved tets, p=5):
    host (NumPy) against GPU (sse_geometric_factors, sse_apply_reference_mapping,
    sse_project_l2), stage by stage, with the largest difference of the arrays they
    produce. Prints one JSON line; value = GPU set-up seconds (the mesh connectivity,
    shared by both, is timed apart)."""
    import torch

    from sse.conservation_laws import BR1, InitialDataGaussian, LaxFriedrichsNumericalFlux
    from sse.conservation_laws import LinearAdvectionDiffusionEquation
    from sse.geometry import ChanWilcoxMetrics, make_spatial_discretization
    from sse.mesh import UniformWarping, uniform_periodic_mesh, warp_mesh
    from sse.reference_approximation import ModalTensor, ReferenceApproximationTet
    from sse.solvers import PhysicalOperator, SkewSymmetricMapping, Solver, StandardForm, initialize

    M, p = args.ad_mesh, 5
    law = LinearAdvectionDiffusionEquation((1.0, 0.5, 0.25), 5e-3)
    ic = InitialDataGaussian(1.0, 0.1, (0.5, 0.5, 0.5))
    form = StandardForm(mapping_form=SkewSymmetricMapping(), inviscid_numerical_flux=LaxFriedrichsNumericalFlux(),
                        viscous_numerical_flux=BR1())
    t0 = time.perf_counter()
    ref = ReferenceApproximationTet(ModalTensor(p), mapping_degree=p)
    t_ref = time.perf_counter() - t0
    t0 = time.perf_counter()
    mesh = warp_mesh(uniform_periodic_mesh(ref, ((0.0, 1.0),) * 3, (M, M, M)), ref,
                     UniformWarping(0.05, (1.0, 1.0, 1.0)))
    t_mesh = time.perf_counter() - t0
    torch.cuda.init()
    out, arrays = {}, {}
    for backend in ("hip", "host"):
        st = {}
        t0 = time.perf_counter()
        sd = make_spatial_discretization(mesh, ref, ChanWilcoxMetrics(), backend=backend)
        st["geometric_factors_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        u0 = initialize(ic, sd)
        st["initial_projection_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        solver = Solver(law, sd, form, PhysicalOperator())  # reference mapping + descriptor + handle
        st["solver_s"] = time.perf_counter() - t0
        st["total_s"] = sum(st.values())
        log(f"set-up {backend}: " + ", ".join(f"{k} {v:.2f}" for k, v in st.items()))
        arrays[backend] = (sd.geometric_factors, u0)
        solver.close()
        out[backend] = st
        del sd, solver
    (gh, uh), (gd, ud) = arrays["host"], arrays["hip"]
    rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
    diffs = {n: rel(getattr(gd, n), getattr(gh, n)) for n in ("J_q", "Lambda_q", "J_f", "nJf", "nJq")}
    diffs["u0"] = rel(ud, uh)
    log("max relative differences: " + ", ".join(f"{k} {v:.1e}" for k, v in diffs.items()))
    line = {
        "metric": "set-up seconds (3D advection-diffusion, 82,944 curved tets, p=5)", "value": out["hip"]["total_s"],
        "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1e3 * out["hip"]["total_s"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"set-up of the 3D advection-diffusion problem, {M}^3 hexes x 6 curved tets, "
                               "ModalTensor p=5, curl metrics, PhysicalOperator",
                   "N_e": mesh.N_e, "reference_approximation_s": t_ref, "mesh_s": t_mesh,
                   "gpu": out["hip"], "host": out["host"], "host_over_gpu": out["host"]["total_s"] / out["hip"]["total_s"],
                   "max_rel_diff": diffs},
    }
    print(json.dumps(line), flush=True)


def advdiff_sweep(args, law, solver, u):
    """Batched sweep of diffusion coefficients b ~ U[1e-3, 1e-2] (seeded) on the
    mesh of `solver`: all members advanced `ad_steps` CK2N54 steps as one ensemble,
    against a loop of plain per-problem solves (timed on `ad_loop` members), then
    the gradient of the final energy 1/2 u' M u with respect to each b by the
    discrete tangent, checked against central differences in b on two members."""
    import numpy as np
    import torch

    from sse.analysis import EnergyConservationAnalysis
    from sse.batch import BatchedSolver
    from sse.conservation_laws import LinearAdvectionDiffusionEquation
    from sse.time_integration import rk_steps

    R, n, dt = args.ad_members, args.ad_steps, args.ad_dt
    bvals = np.random.default_rng(20261202).uniform(1e-3, 1e-2, R)
    members = [solver.with_law(LinearAdvectionDiffusionEquation(law.a, float(b))) for b in bvals]
    t0 = time.perf_counter()
    batch = BatchedSolver(members)
    log(f"sweep: {R} members, set-up {time.perf_counter() - t0:.1f} s; {batch.describe()[:200]}")
    energy = EnergyConservationAnalysis(solver)  # the mass matrix does not depend on b
    dof = u.numel()
    # warm-up (kernel attributes, graph capture), then the timed sweep
    states = [u.clone() for _ in members]
    batch.rk_steps(states, [dt] * R, 2)
    states = [u.clone() for _ in members]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.rk_steps(states, [dt] * R, n)
    torch.cuda.synchronize()
    t_sweep = time.perf_counter() - t0
    J0 = float(energy.evaluate_conservation(u)[0])
    JN = [float(energy.evaluate_conservation(x)[0]) for x in states]
    log(f"sweep: {n} steps x {R} members in {t_sweep:.2f} s; energy {J0:.6e} -> [{min(JN):.6e}, {max(JN):.6e}]")
    # loop of plain per-problem solves on a few members (one handle each)
    L = min(args.ad_loop, R)
    t_loop = 0.0
    for i in range(L):
        s_i = solver.with_law(LinearAdvectionDiffusionEquation(law.a, float(bvals[i])))
        x = u.clone()
        rk_steps(s_i, x, dt, 1)  # warm-up
        x = u.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rk_steps(s_i, x, dt, n)
        torch.cuda.synchronize()
        t_loop += time.perf_counter() - t0
        err = float((x - states[i]).abs().max() / states[i].abs().max())
        if err > 1e-13:
            raise SystemExit(f"ensemble member {i} differs from its own solve: {err:.3e}")
        s_i.close()
    loop_total = t_loop / L * R
    log(f"loop: {L} members {t_loop:.2f} s -> {R} members {loop_total:.2f} s; speed-up {loop_total / t_sweep:.2f}x")
    # gradient of the final energy with respect to b: tangent pairs
    states = [u.clone() for _ in members]
    tangents = [torch.zeros_like(u) for _ in members]
    batch.rk_tangent(states, tangents, [dt] * R, 1)  # warm-up: allocates the pair buffers
    states = [u.clone() for _ in members]
    tangents = [torch.zeros_like(u) for _ in members]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.rk_tangent(states, tangents, [dt] * R, n)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    grads = [float(energy.evaluate_conservation_residual(x, v)[0]) for x, v in zip(states, tangents)]
    # central differences in b on two members (J_N is a polynomial in b)
    fd_errs = []
    for i in range(min(2, R)):
        Js = []
        for sgn in (1.0, -1.0):
            bb = float(bvals[i]) * (1.0 + sgn * 1e-4)
            s_i = solver.with_law(LinearAdvectionDiffusionEquation(law.a, bb))
            x = u.clone()
            rk_steps(s_i, x, dt, n)
            Js.append(float(energy.evaluate_conservation(x)[0]))
            s_i.close()
        fd = (Js[0] - Js[1]) / (2e-4 * float(bvals[i]))
        fd_errs.append(abs(fd - grads[i]) / abs(grads[i]))
    log(f"tangent: {t_grad:.2f} s ({t_grad / t_sweep:.2f}x the sweep); dJ/db in [{min(grads):.4e}, {max(grads):.4e}]; "
        f"FD rel err {max(fd_errs):.2e}")
    # the same gradient by the discrete adjoint (reverse sweep on the forward kernels
    # with the velocity reversed) on a few members, one at a time
    from sse.adjoint import energy_adjoint_b, energy_adjoint_helpers
    A_n = min(args.ad_adjoint, R)
    adj_errs, t_adj = [], 0.0
    helpers = energy_adjoint_helpers(solver) if A_n else None
    if A_n:  # handles, and the mass-matrix factors the first energy functional keeps
        energy_adjoint_b(members[0], u, dt, 1, helpers)
    for i in range(A_n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, g_adj, _ = energy_adjoint_b(members[i], u, dt, n, helpers)
        torch.cuda.synchronize()
        t_adj += time.perf_counter() - t0
        adj_errs.append(abs(g_adj - grads[i]) / abs(grads[i]))
    if helpers:
        for h_ in helpers:
            h_.close()
    adjoint = None
    if A_n:
        per_member_fwd = loop_total / R
        adjoint = {"members": A_n, "s_per_member": t_adj / A_n,
                   "over_forward_solve": t_adj / A_n / per_member_fwd,
                   "max_rel_diff_vs_tangent": max(adj_errs)}
        log(f"adjoint: {t_adj / A_n:.2f} s per member ({adjoint['over_forward_solve']:.2f}x its forward solve); "
            f"dJ/db vs tangent max rel diff {max(adj_errs):.2e}")
    batch.close()
    return {"members": R, "steps": n, "dt": dt, "b_range": [1e-3, 1e-2], "adjoint": adjoint,
            "sweep_s": t_sweep, "dof_steps_per_s": dof * R * n / t_sweep,
            "loop_s_extrapolated": loop_total, "loop_members_timed": L, "speedup_vs_loop": loop_total / t_sweep,
            "energy_0": J0, "energy_N_range": [min(JN), max(JN)],
            "gradient_s": t_grad, "gradient_over_sweep": t_grad / t_sweep,
            "dJ_db_range": [min(grads), max(grads)], "fd_rel_err": max(fd_errs)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mesh", type=int, default=16, help="hexes per direction")
    ap.add_argument("--p", type=int, default=4)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU sample length, each thread count")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="default line: leave out the batched-vs-loop metric")
    ap.add_argument("--no-adjoint", action="store_true", help="default line: leave out the adjoint/forward metric")
    ap.add_argument("--config", default="residual", choices=["residual", "batched", "adjoint", "advdiff", "setup", "multi"],
                    help="residual: the curved-tet Euler p=4 residual (default); batched: the "
                         "batched-vs-loop comparison over 256 vortex problems")
    ap.add_argument("--batch-problems", type=int, default=256)
    ap.add_argument("--multi-cases", default="",
                    help="multi: comma-separated subset of tet_p3_M4_ec, tri_p4_M4_lf, adv_tri_p4_M4_nodal, "
                         "tet_p3_M16_ec")
    ap.add_argument("--adj-mesh", type=int, default=12)
    ap.add_argument("--autotune", action="store_true", help="pick the residual's kernel configuration by timing")
    ap.add_argument("--deterministic", action="store_true",
                    help="residual, adjoint and multi lines: fixed-order sums (repeated calls bitwise equal)")
    ap.add_argument("--ad-mesh", type=int, default=24)
    ap.add_argument("--ad-members", type=int, default=64, help="diffusivities in the sweep (0: none)")
    ap.add_argument("--ad-steps", type=int, default=10)
    ap.add_argument("--ad-dt", type=float, default=1e-4)
    ap.add_argument("--ad-loop", type=int, default=4, help="members timed as plain per-problem solves")
    ap.add_argument("--ad-adjoint", type=int, default=2, help="members whose dJ/db is also taken by the adjoint")
    ap.add_argument("--adj-steps", type=int, default=200)
    ap.add_argument("--adj-directions", type=int, default=32)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the timed region, run 5 copy kernels (dst = 1.0 * src) over this many MiB "
                         "(a kernel of known traffic for calibrating HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X, and the "
              "multi-GPU (element-partitioned, halo-exchange) path is not built. Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.config == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.config == "adjoint":
        print(json.dumps(bench_adjoint(args)), flush=True)
        return
    if args.config == "advdiff":
        return bench_advdiff(args)
    if args.config == "setup":
        return bench_setup(args)
    if args.config == "multi":
        return bench_multi(args)

    import torch

    from sse import abi, semi_discrete_residual
    from sse.time_integration import to_device
    from sse.work_model import residual_work

    t0 = time.perf_counter()
    law, sd, ic, form, ode = build_workload(args.mesh, args.p)
    solver = _solver(args, law, sd, form, ode)
    N_p, N_c, N_e = solver.size()
    dof = N_p * N_c * N_e
    log(f"prep: {time.perf_counter() - t0:.1f} s, N_e={N_e}, N_p={N_p}, N_c={N_c}, DOF={dof}")

    u = to_device(ode.u0)
    dudt = torch.empty_like(u)
    tuned = None
    if args.autotune:  # time every compiled configuration on this state, keep the fastest
        from sse import tune
        timings = tune.autotune(solver, u, nrep=10)
        tune.select(timings[0])
        solver = solver.with_law(solver.conservation_law)
        tuned = [{"kernels": t.describe, "ms": round(t.total_ms, 4)} for t in timings]
        log("autotune: " + "; ".join(f"{t.total_ms:.3f} ms {t.describe}" for t in timings))
    log("kernels: " + solver.describe())
    for _ in range(args.warmup):
        semi_discrete_residual(dudt, u, solver)
    solver.lib.sse_synchronize(solver.handle)
    # untimed: residuals back to back for about 0.3 s, over which the GPU's clocks settle
    # (in a kernel trace the durations fall by a quarter over the first ~40 residuals)
    t_settle = time.perf_counter() + 0.3
    while time.perf_counter() < t_settle:
        for _ in range(20):
            semi_discrete_residual(dudt, u, solver)
        solver.lib.sse_synchronize(solver.handle)
    # per-kernel durations with HIP events on the library's stream over residuals run back
    # to back, taken before the timed region (untimed)
    import ctypes as C
    ms = (C.c_double * abi.SSE_NUM_KERNELS)()
    code = solver.lib.sse_profile_residual(solver.handle, u.data_ptr(), dudt.data_ptr(),
                                           max(5, args.steps), ms)
    if code != 0:
        raise RuntimeError(solver.lib.sse_last_error(solver.handle))
    solver.lib.sse_synchronize(solver.handle)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        semi_discrete_residual(dudt, u, solver)
    solver.lib.sse_synchronize(solver.handle)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = dof * args.steps / elapsed

    kms = {abi.KERNEL_NAMES[i]: ms[i] for i in range(abi.SSE_NUM_KERNELS)}
    work = residual_work(sd.reference_approximation, N_c)
    dom = max(kms, key=kms.get)
    wk = work[dom]
    achieved_tf = wk.flops * N_e / (kms[dom] * 1e-3) / 1e12
    achieved_gbs = wk.bytes * N_e / (kms[dom] * 1e-3) / 1e9
    log(f"kernels (ms): {kms}; dominant {dom}: {achieved_tf:.2f} TFLOP/s "
        f"({100 * achieved_tf / FP64_PEAK_TFLOPS:.1f}% FP64), {achieved_gbs:.0f} GB/s algorithmic "
        f"({100 * achieved_gbs / HBM_PEAK_GBS:.1f}% HBM)")
    workload = f"euler3d_tet_p{args.p}_curved_{args.mesh}cube"
    traffic = load_traffic(workload, dom)

    line = {
        "metric": "DOF-residuals/s (3D Euler p=4 FP64)",
        "value": value,
        "unit": "DOF-residuals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D compressible Euler, curved tets, FluxDifferencingForm EC + LF, "
                               "ModalTensor p=%d, %d^3 hexes x 6 tets, one residual per step" % (args.p, args.mesh),
                   "N_e": N_e, "p": args.p, "N_p": N_p, "N_q": sd.reference_approximation.N_q,
                   "N_f": sd.reference_approximation.N_f, "N_c": N_c, "DOF": dof,
                   "kernel_ms": kms, "kernels": solver.describe(), "deterministic": bool(args.deterministic)},
        "roofline": {"bound": "fp64_valu", "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tf / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": dom, "flops_per_element": wk.flops,
                     "algorithmic_bytes_per_element": wk.bytes,
                     "hbm_frac_algorithmic": achieved_gbs / HBM_PEAK_GBS},
    }
    if tuned:
        line["config"]["autotune"] = tuned
    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device=u.device)
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.mul(src, 1.0, out=dst)  # one elementwise kernel: reads src once, writes dst once
        torch.cuda.synchronize()
        log(f"copy probe: 5 x {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(solver.arrays, ode.u0, dof, args.cpu_seconds)
    # the other two metrics of the headline, each on its own workload
    solver.close()
    del solver, ode, u, dudt
    if not args.no_batched:
        b = bench_batched(args)
        c = b["config"]
        line["batched"] = {"metric": b["metric"], "value": b["value"], "unit": "x",
                           "problems": c["problems"], "steps": b["steps"], "t_loop_s": c["t_loop_s"],
                           "t_batched_s": c["t_batched_s"], "solves_per_s_batched": c["solves_per_s_batched"]}
    if not args.no_adjoint:
        a = bench_adjoint(args)
        c = a["config"]
        line["adjoint"] = {"metric": a["metric"], "value": a["value"], "unit": "x",
                           "total_gradient_over_forward": c["total_gradient_over_forward"],
                           "reverse_only_over_forward": c["reverse_only_over_forward"],
                           "t_forward_s": c["t_forward_s"], "t_gradient_s": c["t_adjoint_total_s"],
                           "N_e": c["N_e"], "steps": a["steps"], "fd_max_rel_err": c["fd_max_rel_err"],
                           "fd_directions": c["fd_directions"], "fd_lambda_rel_err": c["fd_lambda_rel_err"]}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
