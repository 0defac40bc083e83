"""Benchmark of the fused SSP-RK3 march on one MI355X.

Workload: the residual-kernel microbenchmark configuration -- the synthetic
converging-diverging channel (wall y = 1 + 1.5 (x - 0.3)^2) triangulated with
1024 x 512 cells = 2^20 triangles at P3 (10 nodes per element), FP64, seeded
smooth state (seed 0), Roe flux, Zhang-Shu positivity limiter, SSP-RK3 with
the element-local CFL step.  A "step" is one SSP-RK3 step: three fused stage
kernels (residual + stage combination + limiter), with the state resident in
HBM throughout.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  ``value`` is Gdof-stage/s = 3 K E N / t / 1e9.
``roofline`` prices the fused stage kernel against HBM (algorithmic bytes per
launch, see ``workloads.residual_bytes``) with its average duration taken from
HIP events on the launch stream.  ``cpu_baseline`` times the CPU oracle (an
OpenMP restatement of the reference's compiled kernels) on a bounded sample of
the same workload.
"""

from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def _pkg(name):
    return importlib.import_module("2d_nozzle_amd." + name)


def _traffic(workload: str):
    """Measured HBM bytes per launch of the stage kernel, if profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        data = json.load(fh)
    return data.get(workload, {}).get("bytes_per_launch")


def cpu_baseline(order: int, ncx: int, ncy: int, budget_s: float) -> dict:
    """The CPU oracle on a bounded sample: the same channel at ncx x ncy cells."""
    from oracle.oracle import CpuOracle

    wl, cf = _pkg("workloads"), _pkg("config")
    ops, U = wl.channel_case(ncx, ncy, order, dense=True)
    flow = cf.FlowConditions()
    orc = CpuOracle(ops, flow)
    fac = cf.step_coefficient(order, None, "ssprk3")
    U, _ = orc.run(U, 1, "ssprk3", fac)  # warm-up
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        U, _ = orc.run(U, 1, "ssprk3", fac)
        steps += 1
    dt = time.perf_counter() - t0
    cores = len(os.sched_getaffinity(0))
    return {
        "value": 3 * steps * ops.n_elem * ops.ref.n_basis / dt / 1e9,
        "unit": "Gdof-stage/s",
        "cores": int(os.environ.get("OMP_NUM_THREADS", cores)),
        "kind": "port",
        "sample": f"{ncx}x{ncy}-cell channel, {ops.n_elem} P{order} triangles, {steps} SSP-RK3 "
                  f"steps in {dt:.1f} s (C/OpenMP restatement of the reference CPU kernels)",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ncx", type=int, default=1024)
    ap.add_argument("--ncy", type=int, default=512)
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--flux", default="roe", choices=("roe", "hllc", "ausmup"))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; multi-GPU runs "
                          "are not part of this build", "n_gpus": args.gpus}))
        return 2

    import torch

    wl, cf, be = _pkg("workloads"), _pkg("config"), _pkg("backend")
    t_setup = time.perf_counter()
    ops, U = wl.channel_case(args.ncx, args.ncy, args.order)
    flow = cf.FlowConditions(flux=args.flux)
    bk = be.HipBackend(ops, flow, cf.SolverOptions())
    print(f"setup {time.perf_counter() - t_setup:.1f} s: {ops.n_elem} elements, "
          f"{ops.n_edges_total} edges, N={ops.ref.n_basis}", file=sys.stderr, flush=True)
    E, N = ops.n_elem, ops.ref.n_basis
    cur, nxt = bk._S[0], bk._S[1]
    cur.copy_(torch.from_numpy(U).to(bk.device))
    st = torch.cuda.current_stream()
    cur = bk.advance(cur, nxt, args.warmup, "ssprk3")
    nxt = bk._S[1] if cur is bk._S[0] else bk._S[0]
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(st)
    cur = bk.advance(cur, nxt, args.steps, "ssprk3")
    ev1.record(st)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    res = float(bk._res.item())
    t = ms / 1e3
    gdof = 3 * args.steps * E * N / t / 1e9

    nb = wl.residual_bytes(ops)
    launches = 3 * args.steps
    per_launch_bytes = nb["step"] / 3.0
    avg_launch_s = t / launches
    achieved = per_launch_bytes / avg_launch_s / 1e9
    workload = f"channel-tri-p{args.order}-{args.ncx}x{args.ncy}-{args.flux}"
    out = {
        "metric": "Gdof-stage/s FP64 (fused SSP-RK3 stage: residual + stage update + limiter)",
        "value": round(gdof, 4),
        "unit": "Gdof-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "elements": E, "n_basis": N, "edges": ops.n_edges_total,
                   "scheme": "ssprk3", "limiter": "positivity", "flux": args.flux},
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": _traffic(workload),
        },
        "residual_norm_last": res,
        "kernel_avg_us": round(avg_launch_s * 1e6, 2),
        "algorithmic_bytes_per_launch": int(per_launch_bytes),
    }
    if not args.no_cpu:
        sample_cells = (64, 32) if args.order >= 2 else (256, 128)
        out["cpu_baseline"] = cpu_baseline(args.order, sample_cells[0], sample_cells[1],
                                           args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
