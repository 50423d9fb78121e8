#!/usr/bin/env python3
"""Benchmark of the GPU Z-grid step on one MI355X.

Workload: 3D flocking, 1,000,000 agents from 64 Gaussian clusters (sigma 30)
in a 2000^3 box, unit-random velocities, see radius 25, separation /
alignment / cohesion weights 1.5 / 1.0 / 1.0, dt 0.1, speed clamp 2.0, cell
size 25.  A step is the grid rebuild, the see pass and the act pass.

value = see-pass pair interactions/s: P_see (agent_see calls, i.e. pairs
passing the box test, counted on the device) divided by the see kernels'
HIP-event time.  steps/s = timed steps / wall time of tay_run with agents
resident on the device (no fetch inside the timed region).  P_see comes from
a counting twin of the timed state: the GPU path is deterministic, so the
twin walks the identical trajectory.

CPU baseline: the oracle's restatement of the reference's threaded CpuGrid
(tay/cpu_grid.c) on the same input, a bounded sample of steps.
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "see-pass pair interactions/s + steps/s, batched vs plain Z-grid, 1 MI355X"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# FP32 operations every box-passing pair executes in flock_see: box test
# (3 subtractions, 6 compares), offset (3), length (3 mul, 2 add, 1 sqrt) and
# the cutoff compare (1).
F_PAIR = 19
# Compulsory bytes of one see launch per agent: the seer reads p, v and the
# three accumulators (x, y, z) plus both counts (17 words) and writes the
# accumulators and counts (11 words); the seen side reads p and v (6 words).
B_SEE_PER_AGENT = 4 * (17 + 11 + 6)
# HBM bytes per see launch on this workload, from two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE in KiB per dispatch), each scaled by what the same
# counter reads for tay_gpu_copy_probe's 1 GiB copy: FETCH x 2.000, WRITE x
# 1.000 (profiles/see_flock1M_pmc.md).  A counter run cannot be made from
# inside this process, so the figure is carried here and re-measured whenever
# the see kernel changes.
SEE_TRAFFIC = {"agents": 1_000_000, "cell_size": 25.0,
               "bytes": int((96527.0 * 2.000 + 216079.0 * 1.000) * 1024)}


def advance(sim, steps, name):
    """Untimed steps before the warm-up, in runs of 50 with a progress line
    after each (a profiler can slow them many times over)."""
    done = 0
    while done < steps:
        k = min(50, steps - done)
        sim.run(k)
        done += k
        print("%s: advanced %d of %d steps" % (name, done, steps), file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(recs, model, threads, steps):
    """The oracle's CpuGrid (the reported baseline) and CpuZGrid on the same
    agents, one after the other."""
    from tay_amd import api
    from tay_amd import workloads as W
    oracle = api.TayLib(os.path.join(ROOT, "oracle", "_build", "liboracle.so"))
    models = ctypes.CDLL(os.path.join(ROOT, "oracle", "_build", "libmodels.so"))
    oracle.tay_threads_start(threads, 0)
    runs = {}
    for name, space in (("CpuGrid", api.TAY_CPU_GRID), ("CpuZGrid", api.TAY_CPU_Z_GRID)):
        with W.Simulation(oracle, model, recs, space, funcs=models) as s:
            t0 = time.perf_counter()
            s.run(steps)
            wall = time.perf_counter() - t0
            see_s = oracle.oracle_get_see_seconds(s.state)
            _, narrow = oracle.pair_counts(s.state, 0)
        runs[name] = (narrow / see_s, wall)
    oracle.tay_threads_stop()
    value, wall = runs["CpuGrid"]
    z_value, z_wall = runs["CpuZGrid"]
    return {
        "value": value,
        "unit": "pair interactions/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d step(s) of CpuGrid (AoS, serial sort, threaded see/act, indirect call per pair) on the "
                  "same %d agents; %.1f s wall, %.3f steps/s; host %s, nproc %d"
                  % (steps, len(recs), wall, steps / wall, cpu_model(), os.cpu_count() or 0),
        "steps_per_s": steps / wall,
        "cpu_z_grid": {"value": z_value, "unit": "pair interactions/s", "steps_per_s": steps / z_wall,
                       "sample": "the same %d step(s) on CpuZGrid (power-of-two grid, dense Morton cell table); "
                                 "%.1f s wall" % (steps, z_wall)},
    }


def ensemble_cpu_baseline(ens, part, tables, threads):
    """One step of every instance on the oracle's CpuGrid, states run one
    after another as the reference allows."""
    from tay_amd import api
    from tay_amd import workloads as W
    oracle = api.TayLib(os.path.join(ROOT, "oracle", "_build", "liboracle.so"))
    models = ctypes.CDLL(os.path.join(ROOT, "oracle", "_build", "libmodels.so"))
    oracle.tay_threads_start(threads, 0)
    wall = see_s = 0.0
    pairs = 0
    for (rec, model, _), p, mb in zip(ens, part, tables):
        with W.Simulation(oracle, model, rec, api.TAY_CPU_GRID, funcs=models, shared_mb=mb, part=p) as s:
            t0 = time.perf_counter()
            s.run(1)
            wall += time.perf_counter() - t0
            see_s += oracle.oracle_get_see_seconds(s.state)
            pairs += oracle.pair_counts(s.state, 0)[1]
    oracle.tay_threads_stop()
    return {"value": pairs / see_s if see_s else 0.0, "unit": "pair interactions/s", "cores": threads,
            "kind": "port", "steps_per_s": 1.0 / wall,
            "sample": "1 step of each of the %d instances on CpuGrid, one state after another; %.1f s wall"
                      % (len(ens), wall)}


def bench_ensemble(lib, args):
    """Batched ensemble: every instance in one launch per pass (HIP-graph
    step), against one state per instance run one after another through the
    same library (tay.h, TAY_GPU_Z_GRID, same cell sizes, same results)."""
    import numpy as np
    from tay_amd import api
    from tay_amd import workloads as W
    ens = W.ensemble(args.ensemble_instances, seed=7)
    radius = [m.passes[0].radius for _, m, _ in ens]
    factors, widths, walks = (1.0, 1.5, 2.0), (64, 32), (0, 16, 64)
    # tables sized for the smallest cell factor, shared by every run below
    tables = [W.cell_table_mb(side, min(factors) * r) for (_, _, side), r in zip(ens, radius)]
    steps, warm = args.ensemble_steps, 2

    with W.Batch(lib, ens, radius, tables) as b:
        f = (ctypes.c_float * len(factors))(*factors)
        w = (ctypes.c_uint * len(widths))(*widths)
        lw = (ctypes.c_uint * len(walks))(*walks)
        table = (ctypes.c_double * (len(factors) * len(widths) * len(walks)))()
        bf, bw, bl = ctypes.c_float(), ctypes.c_uint(), ctypes.c_uint()
        if not lib.tay_gpu_batch_autotune(b.batch, f, len(factors), w, len(widths), lw, len(walks), 3, table,
                                          ctypes.byref(bf), ctypes.byref(bw), ctypes.byref(bl)):
            raise SystemExit("autotune: " + lib.tay_gpu_last_error().decode())
    part = [bf.value * r for r in radius]
    width, walk = bw.value, bl.value

    with W.Batch(lib, ens, part, tables, count_pairs=True) as twin:
        lib.tay_gpu_batch_set_tile_width(twin.batch, width)
        lib.tay_gpu_batch_set_lane_walk(twin.batch, walk)
        twin.run(warm)
        lib.tay_gpu_batch_reset_pair_counts(twin.batch)
        twin.run(steps)
        p_see = twin.pair_count(0)

    with W.Batch(lib, ens, part, tables) as b:
        lib.tay_gpu_batch_set_tile_width(b.batch, width)
        lib.tay_gpu_batch_set_lane_walk(b.batch, walk)
        b.run(warm, graph=False)
        lib.tay_gpu_batch_enable_timing(b.batch, 1)
        b.run(steps, graph=False)
        see_ms, launches = b.pass_time(0)
        act_ms, _ = b.pass_time(1)

    with W.Batch(lib, ens, part, tables) as b:
        lib.tay_gpu_batch_set_tile_width(b.batch, width)
        lib.tay_gpu_batch_set_lane_walk(b.batch, walk)
        b.run(warm, graph=True)
        lib.tay_gpu_synchronize()
        t0 = time.perf_counter()
        dev_ms = b.run(steps, graph=True)
        wall = time.perf_counter() - t0
        batched = b.agents()

    inst_wall = inst_see_ms = 0.0
    identical = True
    for (rec, model, _), p, mb, got in zip(ens, part, tables, batched):
        with W.Simulation(lib, model, rec, api.TAY_GPU_Z_GRID, shared_mb=mb, part=p) as s:
            lib.tay_set_fetch_mode(s.state, api.TAY_FETCH_ON_DEMAND)
            lib.tay_set_tile_width(s.state, width)
            lib.tay_set_lane_walk(s.state, walk)
            s.run(warm)
            lib.tay_set_pass_timing(s.state, 1)
            lib.tay_gpu_synchronize()
            t0 = time.perf_counter()
            s.run(steps)
            lib.tay_gpu_synchronize()
            inst_wall += time.perf_counter() - t0
            inst_see_ms += lib.pass_time(s.state, 0)[0]
            lib.tay_fetch_agents(s.state, None)
            identical = identical and np.array_equal(s.agents().view(np.uint8), got.view(np.uint8))

    out = {
        "workload": "flocking ensemble, %d instances" % len(ens),
        "instances": len(ens), "agents": int(sum(len(r) for r, _, _ in ens)), "steps": steps, "warmup": warm,
        "p_see_per_step": p_see / steps,
        "batched": {"pairs_per_s": p_see / (see_ms / 1e3), "steps_per_s": steps / wall,
                    "ms_per_step": wall * 1e3 / steps, "device_ms_per_step": dev_ms / steps,
                    "see_ms_per_step": see_ms / launches, "act_ms_per_step": act_ms / launches,
                    "launches_per_pass_per_step": 1},
        "per_instance": {"pairs_per_s": p_see / (inst_see_ms / 1e3), "steps_per_s": steps / inst_wall,
                         "ms_per_step": inst_wall * 1e3 / steps, "see_ms_per_step": inst_see_ms / steps,
                         "launches_per_pass_per_step": len(ens)},
        "identical": bool(identical),
        "autotune": {"cell_factors": list(factors), "tile_widths": list(widths), "lane_walks": list(walks),
                     "ms_per_step": [[[table[(i * len(widths) + j) * len(walks) + k] for k in range(len(walks))]
                                      for j in range(len(widths))] for i in range(len(factors))],
                     "chosen": {"cell_factor": bf.value, "tile_width": width, "lane_walk": walk}},
    }
    out["speedup_pairs_per_s"] = out["batched"]["pairs_per_s"] / out["per_instance"]["pairs_per_s"]
    out["speedup_steps_per_s"] = out["batched"]["steps_per_s"] / out["per_instance"]["steps_per_s"]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = ensemble_cpu_baseline(ens, part, tables, args.cpu_threads)
    return out


def bench_sph(lib, args):
    """SPH dam break: two see passes (density, forces) and two act passes
    per step on 2M particles."""
    from tay_amd import api
    from tay_amd import workloads as W
    recs, lo, hi = W.sph_particles(args.sph_particles, h=1.0)
    ctx = W.sph_context(h=1.0, dt=1e-3, box_min=lo, box_max=hi)
    model = W.sph_model(ctx, 1.0)
    steps, warm = args.sph_steps, 2
    with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID) as twin:
        lib.tay_set_fetch_mode(twin.state, api.TAY_FETCH_ON_DEMAND)
        lib.tay_set_pair_counting(twin.state, 1)
        twin.run(warm)
        twin.run(steps)
        pairs = [lib.pair_counts(twin.state, i)[1] for i in (0, 2)]
    with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID) as sim:
        lib.tay_set_fetch_mode(sim.state, api.TAY_FETCH_ON_DEMAND)
        sim.run(warm)
        lib.tay_set_pass_timing(sim.state, 1)
        lib.tay_gpu_synchronize()
        t0 = time.perf_counter()
        sim.run(steps)
        lib.tay_gpu_synchronize()
        wall = time.perf_counter() - t0
        see = [lib.pass_time(sim.state, i)[0] for i in (0, 2)]
        act = [lib.pass_time(sim.state, i)[0] for i in (1, 3)]
    return {"workload": "SPH dam break, %d particles" % len(recs), "particles": len(recs), "h": 1.0, "dt": 1e-3,
            "steps": steps, "p_see_per_step": [p / steps for p in pairs],
            "pairs_per_s": sum(pairs) / (sum(see) / 1e3), "steps_per_s": steps / wall,
            "ms_per_step": wall * 1e3 / steps, "see_ms_per_step": [t / steps for t in see],
            "act_ms_per_step": [t / steps for t in act]}


def bench_adjoint(lib, args):
    """Differentiable flocking: 65,536 agents, 50 steps, checkpoints every
    10; forward and backward device time at the same (FP32) accuracy, then
    calibration of the three weights to a target trajectory made with hidden
    weights."""
    import numpy as np
    from tay_amd import workloads as W
    from tay_amd.dflock import DiffFlock
    n, steps, every = args.adjoint_agents, 50, 10
    p0, v0 = W.smooth_flock_agents(n, seed=34)
    hidden = np.array([1.4, 0.7, 0.9], np.float32)
    w = np.array([1.0, 1.0, 1.0])
    out = {"workload": "differentiable flocking, %d agents" % n, "agents": n, "steps": steps,
           "checkpoint_every": every, "radius": 20.0, "h2": 1.0, "dt": 0.1, "damping": 0.05}
    with DiffFlock(lib, n, 20.0, 1.0, 0.1, 0.05, checkpoint_every=every) as f:
        target, _ = f.forward(p0, v0, hidden, steps)
        target = target.astype(np.float64)
        fwd = bwd = 0.0
        reps = 5
        for r in range(reps + 1):
            p, _ = f.forward(p0, v0, w.astype(np.float32), steps)
            f.backward((2.0 * (p - target) / p.size).astype(np.float32))
            fm, bm = f.times()
            if r:
                fwd += fm
                bwd += bm
        out.update({"forward_ms": fwd / reps, "backward_ms": bwd / reps, "backward_over_forward": bwd / fwd,
                    "forward_steps_per_s": steps / (fwd / reps / 1e3), "max_neighbours": f.max_neighbours()})
        m, s = np.zeros(3), np.zeros(3)
        loss0 = None
        t0 = time.perf_counter()
        iters = args.adjoint_iters
        for it in range(1, iters + 1):
            p, _ = f.forward(p0, v0, w.astype(np.float32), steps)
            e = p - target
            if loss0 is None:
                loss0 = float(np.mean(e * e))
            _, _, g = f.backward((2.0 * e / p.size).astype(np.float32))
            m = 0.9 * m + 0.1 * g
            s = 0.999 * s + 0.001 * g * g
            lr = 0.05 if it <= iters // 2 else 0.01
            w -= lr * (m / (1 - 0.9 ** it)) / (np.sqrt(s / (1 - 0.999 ** it)) + 1e-12)
        p, _ = f.forward(p0, v0, w.astype(np.float32), steps)
        out["calibration"] = {"iterations": iters, "optimizer": "Adam on the 3 weights", "start": [1.0, 1.0, 1.0],
                              "hidden": hidden.tolist(), "found": w.tolist(), "loss_start": loss0,
                              "loss_end": float(np.mean((p - target) ** 2)), "seconds": time.perf_counter() - t0}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--advance", type=int, default=0, metavar="K",
                    help="run K untimed steps of the flocking state before the warm-up (to time or profile a later "
                         "window of the trajectory; not part of the default line)")
    ap.add_argument("--agents", type=int, default=1_000_000)
    ap.add_argument("--cell-size", type=float, default=25.0, help="Z-grid cell size (min_part_sizes)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-instances", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--sph-particles", type=int, default=2_000_000)
    ap.add_argument("--sph-steps", type=int, default=500)
    ap.add_argument("--no-sph", action="store_true", help="skip the SPH measurement")
    ap.add_argument("--adjoint-agents", type=int, default=65536)
    ap.add_argument("--adjoint-iters", type=int, default=200)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-flocking measurement")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run one plain device-to-device copy of MIB MiB (reads and writes MIB MiB), the "
                         "calibration kernel for rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: the Z-grid space runs on a single MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2

    from tay_amd import api
    from tay_amd import workloads as W

    lib = api.TayLib(api.LIBRARY)
    if not lib.tay_gpu_available():
        raise SystemExit("no GPU: " + lib.tay_gpu_last_error().decode())

    recs = W.flock_agents(args.agents, seed=1, clusters=64, sigma=30.0, box=2000.0)
    model = W.flock_model(radius=25.0, weights=(1.5, 1.0, 1.0), dt=0.1, max_speed=2.0)

    # counting twin: same inputs, same trajectory, pair counts of the timed steps
    with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID, part=args.cell_size) as twin:
        lib.tay_set_fetch_mode(twin.state, api.TAY_FETCH_ON_DEMAND)
        lib.tay_set_pair_counting(twin.state, 1)
        advance(twin, args.advance, "counting twin")
        twin.run(args.warmup) if args.warmup else None
        twin.run(args.steps)
        p_see = lib.pair_counts(twin.state, 0)[1]

    with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID, part=args.cell_size) as sim:
        lib.tay_set_fetch_mode(sim.state, api.TAY_FETCH_ON_DEMAND)
        advance(sim, args.advance, "timed state")
        if args.warmup:
            sim.run(args.warmup)
        lib.tay_set_pass_timing(sim.state, 1)
        lib.tay_gpu_synchronize()
        t0 = time.perf_counter()
        sim.run(args.steps)
        lib.tay_gpu_synchronize()
        wall = time.perf_counter() - t0
        see_ms, see_launches = lib.pass_time(sim.state, 0)
        act_ms, act_launches = lib.pass_time(sim.state, 1)

    steps_per_s = args.steps / wall
    see_s = see_ms / 1e3
    pairs_per_s = p_see / see_s
    see_avg_s = see_s / see_launches
    flop_per_launch = p_see / args.steps * F_PAIR
    achieved = flop_per_launch / see_avg_s / 1e12
    hbm_achieved = B_SEE_PER_AGENT * args.agents / see_avg_s / 1e9

    out = {
        "metric": METRIC,
        "value": pairs_per_s,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "flocking-1M", "agents": args.agents, "clusters": 64, "sigma": 30.0,
                   "box": 2000.0, "radius": 25.0, "cell_size": args.cell_size, "weights": [1.5, 1.0, 1.0], "dt": 0.1,
                   "max_speed": 2.0, "structure": "TAY_GPU_Z_GRID",
                   **({"advanced_steps": args.advance} if args.advance else {})},
        "steps_per_s": steps_per_s,
        "p_see_per_step": p_see / args.steps,
        "see_ms_per_step": see_ms / see_launches,
        "act_ms_per_step": act_ms / max(act_launches, 1),
        "build_ms_per_step": max(wall * 1e3 / args.steps - see_ms / see_launches - act_ms / max(act_launches, 1),
                                 0.0),
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": (SEE_TRAFFIC["bytes"] if (args.agents, args.cell_size) ==
                        (SEE_TRAFFIC["agents"], SEE_TRAFFIC["cell_size"]) else None),
            "kernel": "see (tay_pass_0, flock_see)",
            "note": "FP32 vector-ALU bound, no matrix instructions: the peak is MI355X's FP32 157.3 TFLOP/s "
                    "(FMA counted as 2, packed dual-issue); F = P_see x %d FP32 ops" % F_PAIR,
        },
        "roofline_hbm": {"achieved": hbm_achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": hbm_achieved / PEAK_HBM_GBS, "bytes_per_agent": B_SEE_PER_AGENT},
    }
    if args.copy_probe:
        probe_ms = ctypes.c_double()
        if not lib.tay_gpu_copy_probe(args.copy_probe << 20, ctypes.byref(probe_ms)):
            raise SystemExit("copy probe: " + lib.tay_gpu_last_error().decode())
        out["copy_probe"] = {"bytes": args.copy_probe << 20, "ms": probe_ms.value,
                             "GB/s": 2 * (args.copy_probe << 20) / probe_ms.value / 1e6}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(recs, model, args.cpu_threads, args.cpu_steps)
    if not args.no_ensemble:
        out["ensemble"] = bench_ensemble(lib, args)
    if not args.no_sph:
        out["sph"] = bench_sph(lib, args)
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint(lib, args)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
