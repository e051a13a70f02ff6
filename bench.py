"""Benchmark: IISPH particle-updates/s on the 3D breaking dam (1,000,000 fluid
particles in a 300x150x150 box with a 3-layer closed boundary shell of
696,816 particles) on one MI355X.

One step = neighbor search + the full IISPH step (every Jacobi iteration, CFL
correction, integration) for all fluid particles; particle-updates/s =
fluid particles x steps / seconds.  Inputs are resident in HBM before timing.

Prints one JSON line with `roofline` for the dominant kernel, the search and
pressure-loop `fractions` of the HBM roofline, a `full_run` of all 1000 steps of
the configuration from rest, the achievable HBM bandwidth of this GPU, and
`cpu_baseline` from the CPU oracle timed on this host.
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
COPY_BYTES_CAL = 1 << 30


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(nf, nba, L, ncells=0):
    """Compulsory HBM bytes per phase (per launch) of the IISPH data model:
    every per-particle FP32 array a pass touches once, neighbor lists at the
    reference codec size L bytes per fluid particle, boundary particles within
    reach of fluid (nba) read once per pass that needs them.  The list kernel
    alone reads each particle's key, position, type word and fluid rank once
    and each occupied cell's table entry (16 B) once, and writes the lists and
    their counts."""
    return {
        "search": nf * (12 + 2 * 32 + L) + 12 * nba + 8 * ncells,
        "lists": (nf + nba) * (4 + 16 + 4 + 4) + 16 * ncells + nf * (L + 4),
        "density": nf * (64 + 16 + L) + 32 * nba,  # + d_ii written for phase 2
        # + d_ii read, + the first Jacobi iteration's update (pressure read and
        # written, acceleration and error written)
        "diagonal": nf * (40 + 16 + 28 + L) + 16 * nba,
        "pressure_a": nf * (32 + L) + 16 * nba,
        "pressure_b": nf * (56 + L) + 16 * nba,
        "integrate": nf * 72,
        # the first Jacobi iteration's convergence decision: the block partials
        "pressure_first": 24 * ((nf + 255) // 256),
    }


# timed phases that are one kernel each in a 3D IISPH step with the compressed search
KERNEL_OF_PHASE = {"lists": "k_lists", "density": "k_iisph_prepare", "diagonal": "k_iisph_diagonal",
                   "pressure_a": "k_iisph_pass_a", "pressure_b": "k_iisph_pass_b", "integrate": "k_integrate"}

TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "iisph_dam3d_1M_pmc_traffic.md")


# main-workload searches: the reference search whose neighbor order is reproduced
SEARCHES = {"compressed": (abi.SEARCH_COMPRESSED_3D, "compressed-order"),
            "hashed": (abi.SEARCH_HASHED_3D, "hashed-3d-order"),
            "quadratic": (abi.SEARCH_QUADRATIC_3D, "quadratic-3d-order")}


def measured_traffic(phase, side):
    """HBM bytes per launch of a phase at the 1M configuration, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the counters'
    reading of a 1 GiB copy); None for other sizes or without the summary."""
    if side != 100 or not os.path.exists(TRAFFIC_SUMMARY):
        return None
    with open(TRAFFIC_SUMMARY) as fh:
        for line in fh:
            line = line.strip()
            if line.startswith("traffic-per-launch "):
                fields = dict(kv.split("=", 1) for kv in line.split()[1:])
                return int(fields[phase]) if phase in fields else None
    return None


def boundary_in_reach(sc, coll):
    """N_ba: boundary particles inside some fluid particle's search radius, i.e. the
    distinct boundary entries of the device's fluid neighbor lists."""
    _, cols = sc.export_neighbors()
    cols = np.unique(cols.astype(np.int64))
    return int(np.count_nonzero(coll.type[cols] == abi.TYPE_BOUNDARY))


def phase_byte_model(nf, nba, L, ncells):
    """The per-step byte model of the path's measurement definition: the neighbor
    search B_NS and one Jacobi iteration B_iter."""
    return {"search": nf * (12 + 2 * 32 + L) + 12 * nba + 8 * ncells, "iteration": nf * (88 + 2 * L) + 32 * nba}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def hbm_calibration(nbytes=COPY_BYTES_CAL, reps=10):
    """Achievable HBM bandwidth on this GPU: a device-to-device copy of nbytes (read and
    written once each), timed with events over reps copies after one untimed copy."""
    import torch

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


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


# neighbor-list storage of every 3D scene the bench creates (--lists)
LISTS = {"compressed": abi.LIST_COMPRESSED, "delta": abi.LIST_DELTA, "plain": abi.LIST_PLAIN}
LIST_MODE = abi.LIST_COMPRESSED


def sweep_descs(sweep):
    descs = []
    for sc in sweep:
        descs.append(abi.default_desc(3, abi.SOLVER_IISPH, iisph_gamma=sc["gamma"], iisph_viscosity=sc["viscosity"],
                                      neighbor_lists=LIST_MODE))
    return descs


def bench_batched(ctx, n_scenes, steps, alone_steps):
    """Parameter sweep of independent dam scenes (fluid counts log-uniform in
    [5k, 200k], gamma U[1,7], viscosity U[0.001, 0.05]): all scenes advanced
    together in shared launches, against the same scenes stepped one at a
    time.  Both sides run search + full IISPH step per scene-step."""
    sweep = scenes.config4_sweep(seed=4, n_scenes=n_scenes)
    descs = sweep_descs(sweep)
    nfs = [int(np.count_nonzero(s["collection"].type == abi.TYPE_NORMAL)) for s in sweep]
    nbs = [int(np.count_nonzero(s["collection"].type == abi.TYPE_BOUNDARY)) for s in sweep]
    total_f = sum(nfs)
    log(f"batched sweep: {n_scenes} scenes, {total_f} fluid + {sum(nbs)} boundary")

    # one scene at a time: each scene's steps back to back, scenes in sequence
    alone = [Scene(ctx, d, s["collection"].copy()) for d, s in zip(descs, sweep)]
    for sc in alone:
        sc.step(1)
    ctx.synchronize()
    t0 = time.perf_counter()
    for sc in alone:
        sc.step(alone_steps)
    ctx.synchronize()
    t_alone = time.perf_counter() - t0
    for sc in alone:
        sc.close()
    alone_pus = total_f * alone_steps / t_alone
    log(f"  one at a time: {alone_steps} steps each in {t_alone:.2f}s -> {alone_pus:.3e} particle-updates/s")

    batch = [Scene(ctx, d, s["collection"].copy()) for d, s in zip(descs, sweep)]
    ctx.batch_step(batch, 1)
    ctx.synchronize()
    t0 = time.perf_counter()
    st = ctx.batch_step(batch, steps)
    ctx.synchronize()
    t_batch = time.perf_counter() - t0
    for sc in batch:
        sc.close()
    pus = total_f * steps / t_batch
    log(f"  batched: {steps} steps in {t_batch:.2f}s -> {pus:.3e} particle-updates/s "
        f"(last-step iterations {sorted(set(x.iterations for x in st))})")
    return {
        "workload": f"{n_scenes} independent 3D IISPH dam scenes, fluid log-uniform in [5k, 200k] "
        f"({total_f} fluid, {sum(nbs)} boundary), gamma U[1,7], viscosity U[0.001,0.05]",
        "steps": steps,
        "scenes_per_s": n_scenes / t_batch,
        "scene_steps_per_s": n_scenes * steps / t_batch,
        "particle_updates_per_s": pus,
        "ms_per_step": 1000.0 * t_batch / steps,
        "one_at_a_time_particle_updates_per_s": alone_pus,
        "one_at_a_time_steps": alone_steps,
        "speedup_vs_one_at_a_time": pus / alone_pus,
    }


def _templates(lo, hi, buckets, seed):
    """Dam scenes at log-spaced fluid counts; arrivals reuse the nearest one."""
    sizes = np.exp(np.linspace(np.log(lo), np.log(hi), buckets))
    return [(int(n), scenes.scaled_dam3d(int(n), seed * 100 + k)) for k, n in enumerate(sizes)]


def bench_served(ctx, window_s, rate, max_running, budget_gb, buckets=12, serial_ticks=False):
    """Served scene stream: Poisson arrivals (rate/s) of 3D IISPH dam jobs with
    fluid counts log-uniform in [10k, 500k] and 100-2000 steps.  Jobs are
    submitted when their arrival time has passed, the server admits them
    against the HBM budget and slot limit, and every tick steps all running
    jobs once.  The run stops at the end of the window; unfinished jobs count
    only through their completed steps."""
    from fluidsolver_amd.serving import Server

    arrivals = scenes.config5_arrivals(seed=5, duration_s=window_s, rate=rate)
    tmpl = _templates(10000, 500000, buckets, 5)
    tsizes = np.array([np.log(n) for n, _ in tmpl])
    desc = abi.default_desc(3, abi.SOLVER_IISPH, neighbor_lists=LIST_MODE)
    srv = Server(ctx, hbm_budget_bytes=int(budget_gb * 1e9), max_running=max_running, serial_ticks=serial_ticks)
    k = 0
    offered = 0  # particle-updates the arrived jobs ask for
    open_jobs = {}
    finished = []  # (submit -> finish seconds, queue wait seconds) of completed jobs

    def harvest():
        for jid in list(open_jobs):
            info = srv.job(jid)
            if info.state == abi.JOB_DONE:
                finished.append((info.finish_s - info.submit_s, info.admit_s - info.submit_s))
                srv.release(jid)
                del open_jobs[jid]

    t0 = time.perf_counter()
    next_log = t0 + 5.0
    ticks = 0
    while True:
        now = time.perf_counter() - t0
        if now >= window_s:
            break
        while k < len(arrivals) and arrivals[k][0] <= now:
            _, nfl, steps = arrivals[k]
            n_t, coll = tmpl[int(np.argmin(np.abs(tsizes - np.log(nfl))))]
            jid = srv.submit(desc, coll, steps)
            open_jobs[jid] = True
            offered += int(np.count_nonzero(coll.type == abi.TYPE_NORMAL)) * steps
            k += 1
        if srv.tick() == 0:
            time.sleep(0.001)
        ticks += 1
        if ticks % 50 == 0:
            harvest()
        if time.perf_counter() > next_log:
            st = srv.stats()
            log(f"  served t={now:.0f}s: done {st['jobs_done']} running {st['jobs_running']} "
                f"queued {st['jobs_queued']} p99 {st['p99_step_ms']:.2f} ms")
            next_log += 5.0
    elapsed = time.perf_counter() - t0
    harvest()
    st = srv.stats()
    srv.close()
    lat = np.array([f[0] for f in finished]) if finished else np.zeros(1)
    wait = np.array([f[1] for f in finished]) if finished else np.zeros(1)
    return {
        "workload": f"Poisson arrivals {rate}/s of 3D IISPH dam jobs, fluid log-uniform in [10k, 500k] "
        f"({buckets} size templates), 100-2000 steps each",
        "window_s": window_s,
        "ticks": "serial" if serial_ticks else "overlapped",
        "max_running": max_running,
        "hbm_budget_gb": budget_gb,
        "jobs_arrived": k,
        "jobs_done": st["jobs_done"],
        "jobs_failed": st["jobs_failed"],
        "scenes_per_s": st["jobs_done"] / elapsed,
        "scene_steps_per_s": st["scene_steps"] / elapsed,
        "particle_updates_per_s": st["particle_updates"] / elapsed,
        "p50_step_ms": st["p50_step_ms"],
        "p99_step_ms": st["p99_step_ms"],
        "mean_step_ms": st["mean_step_ms"],
        "busy_fraction": st["busy_s"] / elapsed,
        "tick_seconds": {"admit": round(st["admit_s"], 3), "step": round(st["step_s"], 3),
                         "retire": round(st["retire_s"], 3), "window": round(elapsed, 3)},
        "offered_particle_updates_per_s": offered / window_s,
        "sustained_particle_updates_per_s": st["particle_updates"] / elapsed,
        "jobs_queued_at_end": st["jobs_queued"],
        "jobs_running_at_end": st["jobs_running"],
        "job_latency_p50_s": float(np.percentile(lat, 50)),
        "job_latency_p99_s": float(np.percentile(lat, 99)),
        "job_queue_wait_p50_s": float(np.percentile(wait, 50)),
        "job_queue_wait_p99_s": float(np.percentile(wait, 99)),
        "step_latency_note": "p50/p99_step_ms: wall time of a server tick, counted once per scene it stepped; "
                             "job_latency: submit to finish of the jobs that finished in the window",
    }


def bench_obstacles(ctx, steps, warmup):
    """The obstacle configuration: a 200^3 fluid block beside 80 seeded spheres
    (radius U[5,20]) and 40 tori surface-sampled as boundary particles, inside the
    3-layer shell of a 600 x 300 x 300 box (2,991,211 boundary particles); compressed search with the NeighborStorage capacity.  SESPH
    (k = 1000, the reference's linear equation of state) and IISPH run as two
    scenes from the identical initial state, each timed over `steps` steps
    after `warmup` steps."""
    coll = scenes.config3_obstacles(seed=3)
    counts = coll.counts()
    nf, nb = counts["fluid"], counts["boundary"]
    log(f"obstacles: {nf} fluid, {nb} boundary")
    out = {
        "workload": f"200^3 fluid block with 80 sphere and 40 torus obstacles in a 600x300x300 box, {nf} fluid + {nb} boundary, "
                    "compressed search with NeighborStorage capacity",
        "steps": steps,
        "warmup": warmup,
    }
    for name, solver in (("sesph", abi.SOLVER_SESPH), ("iisph", abi.SOLVER_IISPH)):
        d = abi.default_desc(3, solver, sesph_stiffness_k=1000.0)
        d.neighbor_lists = LIST_MODE if LIST_MODE != abi.LIST_PLAIN else abi.LIST_COMPRESSED
        sc = Scene(ctx, d, coll.copy())
        for _ in range(warmup):
            sc.step(1)
        ctx.synchronize()
        iters = []
        t0 = time.perf_counter()
        for _ in range(steps):
            iters.append(sc.step(1).iterations)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        sc.close()
        out[name] = {"ms_per_step": 1000.0 * dt / steps, "particle_updates_per_s": nf * steps / dt}
        if solver == abi.SOLVER_IISPH:
            out[name]["mean_jacobi_iterations"] = float(np.mean(iters))
        log(f"  {name}: {1000.0 * dt / steps:.3f} ms/step, {nf * steps / dt:.3e} particle-updates/s")
    return out


def pmc_workload(args):
    """Run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE: one plain device copy of
    COPY_BYTES (read and written once each) calibrates the counters; then the
    dam scene steps, so each kernel's per-dispatch counters can be scaled."""
    import torch

    side = args.side
    coll = scenes.config2_dam3d(seed=2, side=side, interior=(3 * side, 3 * side // 2, 3 * side // 2))
    ctx = Context(0)
    sc = Scene(ctx, abi.default_desc(3, abi.SOLVER_IISPH, neighbor_lists=LIST_MODE), coll)
    sc.step(2)
    ctx.synchronize()
    log("pmc workload: calibration copy")
    a = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log("pmc workload: 3 steps")
    sc.step(3)
    ctx.synchronize()
    log("pmc workload: done")
    sc.close()
    ctx.close()


COPY_BYTES = 1 << 30


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000,
                    help="timed steps (the dam configuration is 1000 steps)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--side", type=int, default=100, help="fluid block side (100 -> 1,000,000 fluid)")
    ap.add_argument("--search", choices=sorted(SEARCHES), default="compressed",
                    help="neighbor order of the main workload: the reference search it reproduces")
    ap.add_argument("--solver", choices=["iisph", "sesph"], default="iisph",
                    help="solver of the main workload (SESPH: stiffness 1000)")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-side", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--alone-steps", type=int, default=50)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=30.0)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-slots", type=int, default=32)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serial-ticks", action="store_true",
                    help="served leg: admit, step and retire one after another instead of overlapping the transfers")
    ap.add_argument("--obstacle-steps", type=int, default=200)
    ap.add_argument("--full-steps", type=int, default=1000,
                    help="also time this many steps of a fresh 1M scene from rest (the whole configuration; 0: skip)")
    ap.add_argument("--no-obstacles", action="store_true")
    ap.add_argument("--graph", type=int, default=0,
                    help="replay step heads from hipGraphs (1) or launch eagerly (0, the faster on MI355X)")
    ap.add_argument("--lists", choices=sorted(LISTS), default="compressed",
                    help="neighbor-list storage the solver walks read: 32-bit entries with the NeighborStorage "
                         "capacity (compressed), 16-bit delta codes with the same capacity (delta), or 32-bit "
                         "entries with the 97-neighbor capacity only (plain)")
    ap.add_argument("--pmc-workload", action="store_true",
                    help="workload for rocprofv3 --pmc: a device copy of 1 GiB (counter calibration), then 3 steps")
    args = ap.parse_args()
    global LIST_MODE
    LIST_MODE = LISTS[args.lists]
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)

    if args.pmc_workload:
        pmc_workload(args)
        return

    # torch's HIP runtime initialises before the library's (it finds no device after)
    hbm_gbps = hbm_calibration()
    log(f"achievable HBM bandwidth (1 GiB device copy): {hbm_gbps:.0f} GB/s")
    side = args.side
    interior = (3 * side, 3 * side // 2, 3 * side // 2)
    coll = scenes.config2_dam3d(seed=2, side=side, interior=interior)
    counts = coll.counts()
    nf, nb = counts["fluid"], counts["boundary"]
    if args.solver == "sesph":
        desc = abi.default_desc(3, abi.SOLVER_SESPH, sesph_stiffness_k=1000.0)
    else:
        desc = abi.default_desc(3, abi.SOLVER_IISPH)
    desc.search = SEARCHES[args.search][0]
    desc.neighbor_lists = LIST_MODE
    log(f"scene: {nf} fluid, {nb} boundary, {args.search} search")

    ctx = Context(0)
    ctx.set_graph(bool(args.graph))
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
    log(f"timed {args.steps} steps: {elapsed:.3f}s, {ms_per_step:.3f} ms/step, "
        f"iterations mean {np.mean(iters):.2f} max {max(iters)}")

    # per-phase device time (HIP events on the context stream) over extra steps
    sc.find_neighbors()
    ns = sc.neighbor_stats()
    L = ns["reference_codec_bytes"] / max(nf, 1)
    nba = boundary_in_reach(sc, coll) if args.search == "compressed" else nb
    log(f"boundary particles within reach of fluid: {nba} of {nb}")
    sc.simulate()
    ctx.set_timing(True)
    for _ in range(args.profile_steps):
        sc.step(1)
    timing = ctx.timing()
    ctx.set_timing(False)
    cells = np.floor(coll.position[:, :3] / np.float32(desc.particle_size * 2)).astype(np.int64)
    ncells = len(np.unique((cells[:, 2] * (1 << 21) + cells[:, 1]) * (1 << 21) + cells[:, 0]))
    algo = algorithmic_bytes(nf, nba, L, ncells)
    phases = {}
    for name, (ms, cnt) in timing.items():
        if cnt:
            avg_ms = ms / cnt
            phases[name] = {
                "avg_ms": avg_ms,
                "launches": cnt,
                "ms_per_step": ms / args.profile_steps,
                # the gated tail pass does work only when a loop outlasts the prediction
                "GBps": algo[name] / (avg_ms * 1e-3) / 1e9 if name in algo else None,
            }
    log(json.dumps(phases))
    # the dominant kernel: the most device time per step among the phases that
    # are one kernel each (the search phase is several and encloses the lists)
    single = [k for k in (KERNEL_OF_PHASE if args.solver == "iisph" else ()) if k in phases]
    dom = max(single, key=lambda k: phases[k]["ms_per_step"]) if single else \
        max(phases, key=lambda k: phases[k]["ms_per_step"])
    achieved = phases[dom]["GBps"]
    roofline = {
        "bound": "hbm",
        "kernel": KERNEL_OF_PHASE.get(dom, dom),
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(dom, side) if args.search == "compressed" else None,
        "traffic_source": "profiles/iisph_dam3d_1M_pmc_traffic.md (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0)",
        "algorithmic_bytes_per_launch": int(algo[dom]),
        "avg_launch_ms": phases[dom]["avg_ms"],
    }
    # every timed phase against the same peak, the pressure loop's passes included
    roofline_by_phase = {
        k: {"kernel": KERNEL_OF_PHASE.get(k, k), "avg_launch_ms": round(v["avg_ms"], 5),
            "launches_per_step": v["launches"] / args.profile_steps,
            "algorithmic_bytes_per_launch": int(algo[k]), "achieved_GBps": round(v["GBps"], 1),
            "frac": round(v["GBps"] / HBM_PEAK_GBS, 4)}
        for k, v in phases.items() if v["GBps"] is not None
    }
    if "pressure_tail" in phases:
        t = phases["pressure_tail"]
        roofline_by_phase["pressure_tail"] = {"kernel": "k_iisph_pass_a + k_iisph_pass_b, gated",
                                              "avg_launch_ms": round(t["avg_ms"], 5),
                                              "launches_per_step": t["launches"] / args.profile_steps}

    # the two fractions of the measurement definition: neighbor search B_NS / t_NS and the
    # pressure loop I * B_iter / t_loop (t_loop: the Jacobi passes and their convergence
    # decisions, the gated pass beyond the predicted ones included; the first iteration's
    # update runs inside the phase-2 kernel)
    sb = phase_byte_model(nf, nba, L, ncells)
    per_step = {k: v["ms_per_step"] for k, v in phases.items()}
    t_loop = sum(per_step.get(k, 0.0) for k in ("pressure_a", "pressure_b", "pressure_first", "pressure_tail"))
    mean_it = float(np.mean(iters))
    fractions = {
        "search": {"bytes_per_step": int(sb["search"]), "ms_per_step": per_step.get("search"),
                   "frac": round(sb["search"] / (per_step["search"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                   if per_step.get("search") else None},
        "pressure_loop": {"bytes_per_iteration": int(sb["iteration"]), "iterations": mean_it,
                          "ms_per_step": t_loop,
                          "frac": round(mean_it * sb["iteration"] / (t_loop * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                          if t_loop > 0 else None,
                          "frac_walking_iterations_only": round(
                              (mean_it - 1) * sb["iteration"] / (t_loop * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                          if t_loop > 0 else None},
    }
    log(json.dumps(fractions))
    sc.close()

    resort = None
    if args.search == "hashed":
        # the hashed search re-sorts the collection on every 101st call: the
        # wall time of each of 202 searches of a fresh scene (two re-sorts)
        hs = Scene(ctx, desc, scenes.config2_dam3d(seed=2, side=side, interior=interior))
        ctx.synchronize()
        call_ms = []
        for _ in range(202):
            t0 = time.perf_counter()
            hs.find_neighbors()
            ctx.synchronize()
            call_ms.append(1000.0 * (time.perf_counter() - t0))
        hs.close()
        resort = {"resort_call_ms": [round(call_ms[100], 3), round(call_ms[201], 3)],
                  "other_calls_median_ms": round(float(np.median(call_ms[:100] + call_ms[101:201])), 3)}
        log(f"hashed re-sort: {json.dumps(resort)}")

    full = None
    if args.full_steps > 0 and args.solver == "iisph":
        # the whole configuration: a fresh scene from rest, every step timed
        fs = Scene(ctx, desc, scenes.config2_dam3d(seed=2, side=side, interior=interior))
        ctx.synchronize()
        fit = []
        t0 = time.perf_counter()
        for _ in range(args.full_steps):
            fit.append(fs.step(1).iterations)
        ctx.synchronize()
        t_full = time.perf_counter() - t0
        fs.close()
        full = {"steps": args.full_steps, "seconds": t_full, "ms_per_step": 1000.0 * t_full / args.full_steps,
                "particle_updates_per_s": nf * args.full_steps / t_full,
                "mean_jacobi_iterations": float(np.mean(fit)), "max_jacobi_iterations": int(np.max(fit))}
        log(f"full run: {json.dumps(full)}")

    batched = None if args.no_batch else bench_batched(ctx, args.batch_scenes, args.batch_steps, args.alone_steps)
    if batched:
        log(json.dumps(batched))
    served = None
    if not args.no_serve:
        served = bench_served(ctx, args.serve_seconds, args.serve_rate, args.serve_slots, args.serve_budget_gb,
                              serial_ticks=args.serial_ticks)
        log(json.dumps(served))

    obstacles = None
    if not args.no_obstacles:
        obstacles = bench_obstacles(ctx, args.obstacle_steps, 2)
        log(json.dumps(obstacles))

    tuning = ctx.tuning()
    if tuning:
        log("autotuned neighbor-walk unroll:\n" + tuning.rstrip())

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
            "cpu_model": cpu_model(),
            "host_cpus": os.cpu_count(),
        }
        log(json.dumps(cpu))

    out = {
        "metric": "IISPH particle-updates/s @1M 3D; batched scenes/s + p99 step latency",
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
            "workload": f"3D {args.solver.upper()} breaking dam, {nf} fluid + {nb} boundary particles, "
                        f"{SEARCHES[args.search][1]} search",
            "fluid_particles": nf,
            "boundary_particles": nb,
            "mean_jacobi_iterations": float(np.mean(iters)),
            "reference_codec_bytes_per_list": round(L, 2),
            "boundary_within_reach_of_fluid": nba,
            "occupied_cells": ncells,
            "neighbor_lists": args.lists,
        },
        "fractions": fractions,
        "full_run": full,
        "hashed_resort": resort,
        "hbm_measured_GBps": round(hbm_gbps, 1),
        "roofline": roofline,
        "roofline_by_phase": roofline_by_phase,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "obstacles": obstacles,
        "graph": bool(args.graph),
        "autotune": [ln for ln in tuning.splitlines() if ln],
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
