"""Benchmark: IISPH particle-updates/s on the 3D breaking dam (1,000,000 fluid
particles in a 300x150x150 box with a 3-layer closed boundary shell of
696,816 particles) on one MI355X.

One step = neighbor search + the full IISPH step (every Jacobi iteration, CFL
correction, integration) for all fluid particles; particle-updates/s =
fluid particles x steps / seconds.  Inputs are resident in HBM before timing.

Prints one JSON line (the driver contract) with `roofline` for the dominant
kernel and `cpu_baseline` from the CPU oracle timed on this host.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from fluidsolver_amd import abi, scenes  # noqa: E402
from fluidsolver_amd._native import lib  # noqa: E402
from fluidsolver_amd.solver import Context, Scene, check  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(nf, nba, L):
    """Compulsory HBM bytes per phase (per launch) of the IISPH data model:
    every per-particle FP32 array a pass touches once, neighbor lists at the
    reference codec size L bytes per fluid particle, boundary particles within
    reach of fluid (nba) read once per pass that needs them."""
    return {
        "search": nf * (12 + 2 * 32 + L) + 12 * nba,
        "density": nf * (64 + L) + 32 * nba,
        "diagonal": nf * (40 + L) + 16 * nba,
        "pressure_a": nf * (32 + L) + 16 * nba,
        "pressure_b": nf * (56 + L) + 16 * nba,
        "integrate": nf * 72,
    }


def cpu_baseline(desc, coll, steps, threads):
    from oracle import oracle as O

    if threads:
        O.set_threads(threads)
    orc = O.OracleScene(desc, coll, mode=O.MODE_FP32)
    orc.step(1)  # warm-up step (first sort from the upload order)
    t0 = time.perf_counter()
    orc.step(steps)
    dt = time.perf_counter() - t0
    nf = int(np.count_nonzero(coll.type == abi.TYPE_NORMAL))
    return nf * steps / dt, O.get_threads(), dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--side", type=int, default=100, help="fluid block side (100 -> 1,000,000 fluid)")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cpu-side", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)

    side = args.side
    interior = (3 * side, 3 * side // 2, 3 * side // 2)
    coll = scenes.config2_dam3d(seed=2, side=side, interior=interior)
    counts = coll.counts()
    nf, nb = counts["fluid"], counts["boundary"]
    desc = abi.default_desc(3, abi.SOLVER_IISPH)
    log(f"scene: {nf} fluid, {nb} boundary")

    ctx = Context(0)
    sc = Scene(ctx, desc, coll)
    t0 = time.perf_counter()
    for _ in range(args.warmup):
        sc.step(1)
    ctx.synchronize()
    log(f"warmup {args.warmup} steps: {time.perf_counter() - t0:.2f}s")

    ctx.synchronize()
    iters = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = sc.step(1)
        iters.append(st.iterations)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    value = nf * args.steps / elapsed
    ms_per_step = 1000.0 * elapsed / args.steps
    log(f"timed {args.steps} steps: {elapsed:.3f}s, {ms_per_step:.3f} ms/step, iterations {iters}")

    # per-phase device time (HIP events on the context stream) over extra steps
    sc.find_neighbors()
    ns = sc.neighbor_stats()
    L = ns["reference_codec_bytes"] / max(nf, 1)
    sc.simulate()
    ctx.set_timing(True)
    for _ in range(args.profile_steps):
        sc.step(1)
    timing = ctx.timing()
    ctx.set_timing(False)
    nba = nb  # boundary particles are all within reach of fluid in the worst case
    algo = algorithmic_bytes(nf, nba, L)
    phases = {}
    for name, (ms, cnt) in timing.items():
        if cnt:
            avg_ms = ms / cnt
            phases[name] = {
                "avg_ms": avg_ms,
                "launches": cnt,
                "GBps": algo[name] / (avg_ms * 1e-3) / 1e9,
            }
    log(json.dumps(phases))
    dom = "pressure_b" if "pressure_b" in phases else max(phases, key=lambda k: phases[k]["avg_ms"])
    achieved = phases[dom]["GBps"]
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "algorithmic_bytes_per_launch": int(algo[dom]),
        "avg_launch_ms": phases[dom]["avg_ms"],
    }

    cpu = None
    if not args.no_cpu:
        cside = args.cpu_side
        ccoll = scenes.config2_dam3d(seed=2, side=cside, interior=(3 * cside, 3 * cside // 2, 3 * cside // 2))
        v, cores, t = cpu_baseline(desc, ccoll, args.cpu_steps, 0)
        cpu = {
            "value": v,
            "unit": "particle-updates/s",
            "cores": cores,
            "kind": "port",
            "sample": f"{args.cpu_steps} IISPH step(s) of the same scene with {cside}^3 fluid after 1 warm-up step, "
            f"{t:.1f}s on {cores} threads",
        }
        log(json.dumps(cpu))

    out = {
        "metric": "IISPH particle-updates/s @1M 3D",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"3D IISPH breaking dam, {nf} fluid + {nb} boundary particles, compressed-order search",
            "fluid_particles": nf,
            "boundary_particles": nb,
            "mean_jacobi_iterations": float(np.mean(iters)),
            "reference_codec_bytes_per_list": round(L, 2),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    sc.close()
    ctx.close()


if __name__ == "__main__":
    main()
