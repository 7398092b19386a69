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


def cpu_baseline(arrays, u0, dof, target_s: float):
    """The CPU oracle (restatement of the reference algorithm, OpenMP over
    elements like the reference's Threaded()) on the same workload."""
    from oracle import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    O.set_threads(threads)
    t0 = time.perf_counter()
    O.residual(arrays, u0)
    one = time.perf_counter() - t0
    nrep = max(1, int(target_s / max(one, 1e-3)))
    t0 = time.perf_counter()
    O.residual_repeat(arrays, u0, nrep)
    el = time.perf_counter() - t0
    return {"value": dof * nrep / el, "unit": "DOF-residuals/s", "cores": threads, "kind": "port",
            "sample": f"{nrep} full residual evaluations of the same 24,576-element workload "
                      f"({el:.1f} s, {threads} OpenMP threads, C++ restatement of the reference algorithm)"}


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
    per-problem GPU solves. Prints one JSON line."""
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
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mesh", type=int, default=16, help="hexes per direction")
    ap.add_argument("--p", type=int, default=4)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", default="residual", choices=["residual", "batched"],
                    help="residual: the curved-tet Euler p=4 residual (default); batched: the "
                         "batched-vs-loop comparison over 256 vortex problems")
    ap.add_argument("--batch-problems", type=int, default=256)
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
        return bench_batched(args)

    import torch

    from sse import abi, semi_discrete_residual
    from sse.time_integration import to_device
    from sse.work_model import residual_work

    t0 = time.perf_counter()
    law, sd, ic, form, ode = build_workload(args.mesh, args.p)
    solver = ode.p
    N_p, N_c, N_e = solver.size()
    dof = N_p * N_c * N_e
    log(f"prep: {time.perf_counter() - t0:.1f} s, N_e={N_e}, N_p={N_p}, N_c={N_c}, DOF={dof}")

    u = to_device(ode.u0)
    dudt = torch.empty_like(u)
    log("kernels: " + solver.describe())
    for _ in range(args.warmup):
        semi_discrete_residual(dudt, u, solver)
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

    # per-kernel durations with HIP events on the library's stream
    import ctypes as C
    ms = (C.c_double * abi.SSE_NUM_KERNELS)()
    code = solver.lib.sse_profile_residual(solver.handle, u.data_ptr(), dudt.data_ptr(),
                                           max(5, args.steps), ms)
    if code != 0:
        raise RuntimeError(solver.lib.sse_last_error(solver.handle))
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
                   "kernel_ms": kms},
        "roofline": {"bound": "mfma", "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tf / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": dom, "flops_per_element": wk.flops,
                     "algorithmic_bytes_per_element": wk.bytes,
                     "hbm_frac_algorithmic": achieved_gbs / HBM_PEAK_GBS},
    }
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
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
