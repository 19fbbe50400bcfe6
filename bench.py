"""Benchmark: agent-steps/s of the Boids step at 2^20 agents, k = 7, on one
MI355X.

Workload: 1,048,576 agents uniform in a 1000^3 box with face wrap, velocities
random unit x U(0, 9), SteerLibrary flocking (separation 5/-0.707/12,
alignment 7.5/0.7/8, cohesion 9/-0.15/8), 64^3 spatial bins, exact KNN k = 7,
SimpleVehicle integrate, dt = 1/60, fp32.  A step is one pass of the whole
path over the group: wrap -> bin build (counting sort) -> exact KNN ->
flocking forces -> integrate, inputs resident in HBM.

Prints one JSON line (the driver's contract).  `roofline` is computed for the
dominant kernel (the exact KNN kernel) from its
HIP-event-timed average duration on the library's stream; `cpu_baseline`
times the CPU oracle restatement of the same step on this host's cores on a
bounded sample (a few full steps of the same 2^20-agent state).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def bytes_per_agent_step(k: int, cells: int, n: int) -> float:
    """Algorithmic bytes per agent-step of the whole path (SoA fp32):
    key+histogram 16, scan 12*C/N, scatter 64, KNN 12+4k, steering 40+4k,
    integrate 92 -> 224 + 8k + 12 C/N."""
    return 224.0 + 8.0 * k + 12.0 * cells / n


def bytes_knn_kernel(k: int) -> float:
    """Algorithmic bytes per agent of the KNN kernel: read the sorted
    position (12) and write k neighbour indices (4k)."""
    return 12.0 + 4.0 * k


def cpu_baseline(st, p, seconds: float):
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as orc

    threads = min(os.cpu_count() or 1, 16)
    orc.set_threads(threads)
    s = st.copy()
    steps = 0
    t0 = time.perf_counter()
    while True:
        orc.boids_step(s, p)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 50:
            break
    return {
        "value": st.n * steps / el,
        "unit": "agent-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} full steps of the same {st.n}-agent state (oracle restatement: stable counting-sort "
                  f"bins, exact binned KNN, SteerLibrary forces, SimpleVehicle integrate; OpenMP over agents)",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--k", type=int, default=7)
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md); run with --gpus 1.", file=sys.stderr)
        return 2

    import gpusteer_amd as gs
    from gpusteer_amd import scenarios

    st, p = scenarios.boids_box(n=args.n, k=args.k, cells=args.cells)
    n = st.n
    C = args.cells ** 3
    print(f"bench: boids_box n={n} k={args.k} cells={args.cells}^3 warmup={args.warmup} steps={args.steps}",
          file=sys.stderr, flush=True)
    with gs.Context(0) as ctx:
        A = gs.AgentGroup(ctx, st)
        bins = gs.KNNBinData(ctx, p.grid_cells, p.grid_size, origin=p.grid_origin)
        knn = gs.KNNData(ctx, n, args.k)
        gs.stepBoids(A, knn, bins, p, max(args.warmup, 1))
        ctx.synchronize()
        # Timed region: K whole steps, synchronised on both sides.
        t0 = time.perf_counter()
        gs.stepBoids(A, knn, bins, p, args.steps)
        ctx.synchronize()
        el = time.perf_counter() - t0
        # Per-stage HIP-event timing on the library's stream.
        tsteps = min(args.steps, 50)
        bin_ms, knn_ms, si_ms, step_ms = gs.stepBoidsTimed(A, knn, bins, p, tsteps)
        status, bad = A.check_finite()
        A.close()
        bins.close()
        knn.close()
    if status != 0:
        print(f"bench: non-finite state at agent {bad}", file=sys.stderr)
        return 1
    ms = el * 1e3 / args.steps
    value = n * args.steps / el
    b_knn = bytes_knn_kernel(args.k)
    achieved = n * b_knn / (knn_ms * 1e-3) / 1e9
    print(f"bench: {ms:.3f} ms/step, bin build {bin_ms:.3f} ms, knn {knn_ms:.3f} ms, steer+integrate "
          f"{si_ms:.3f} ms (event-timed step {step_ms:.3f} ms); whole-step effective "
          f"{n * bytes_per_agent_step(args.k, C, n) / (ms * 1e-3) / 1e9:.1f} GB/s", file=sys.stderr, flush=True)
    out = {
        "metric": "agent-steps/s at 1M agents k=7",
        "value": value,
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"boids_box: {n} agents, 1000^3 box with face wrap, exact KNN k={args.k}, "
                        f"{args.cells}^3 bins, SteerLibrary flocking + SimpleVehicle integrate, dt=1/60",
            "n_agents": n,
            "k": args.k,
            "cells": [args.cells] * 3,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_knn_binned (exact seeded KNN)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_agent": b_knn,
            "kernel_ms": knn_ms,
        },
        "stage_ms": {"bin_build": bin_ms, "knn": knn_ms, "steer_integrate": si_ms},
        "whole_step_effective_gbs": n * bytes_per_agent_step(args.k, C, n) / (ms * 1e-3) / 1e9,
    }
    if not args.no_cpu:
        print("bench: cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(st, p, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
