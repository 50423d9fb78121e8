#!/usr/bin/env python3
"""Benchmark of the GPU Z-grid step on one MI355X.

Headline workload: 3D flocking, 1,000,000 agents from 64 Gaussian clusters
(sigma 30) in a 2000^3 box, unit-random velocities, see radius 25,
separation / alignment / cohesion weights 1.5 / 1.0 / 1.0, dt 0.1, speed
clamp 2.0, cell size 25.  A step is the grid rebuild, the see pass and the
act pass.

value = see-pass pair interactions/s: P_see (agent_see calls, i.e. pairs
passing the box test, counted on the device) divided by the see kernels'
HIP-event time.  P_see comes from a counting twin of the timed state: the GPU
path is deterministic, so the twin walks the identical trajectory.
steps_per_s = timed steps / wall time of one tay_run in the default fetch
mode, which ends by copying every agent back to host storage, as the
reference's ms-per-step timer does (tay/state.c:196-320); steps_per_s_resident
leaves that copy out.  full_run repeats the measurement over the configured
200 steps from the initial agents.

roofline: the see kernel is bound by FP32 vector issue (no matrix
instructions: a pair interaction is not a contraction).  F = P_see x f_pair,
f_pair counted from flock_see's source per branch and weighted by the
branch shares measured on the window's first and last agents (see-only
passes: P_in from the plain counting model, P_front and P_sep from
flock_see's own counters).  The fraction is given against MI355X's
published 157.3 TFLOP/s (an FMA as two operations, packed dual issue) and
against 39.3 T operations/s, one FP32 operation per lane per clock, the
ceiling of code built without contraction or packed FP32.

Other workloads, each with its roofline and CPU baseline: c1 (10k plain
agents, radius 40), sph (2M-particle dam break), ensemble (256 flocking
instances, batched vs one launch per instance), adjoint (65,536-agent
differentiable flocking, forward and backward).

CPU baselines: the oracle's restatements of the reference's CpuGrid,
CpuZGrid and CpuSimple (tay/cpu_grid.c, tay/cpu_z_grid.c, tay/cpu_simple.c;
AoS records, serial sort, threaded see and act, an indirect call per pair)
on the same inputs, bounded samples, at --cpu-threads and at one thread.
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
PEAK_FP32_TFLOPS = 157.3          # published: FMA as two operations, packed dual issue
PEAK_FP32_SINGLE_TOPS = 39.3      # one FP32 operation per lane per clock (no FMA, no packed FP32)
PEAK_HBM_GBS = 8000.0
CONFIGURED_STEPS = 200            # the 1M flocking workload's length

# FP32 operations per box-passing pair, per branch, counted from the model
# sources (tay_amd/agent_src).  Every pair that passes Tay's per-axis box test
# has executed the test (3 subtractions, 6 compares per pair of a 3D space).
F_BOX_TEST = 9
# flock_see: every pair: offset (3), length (3 mul, 2 add, 1 sqrt), cutoff
# compare; inside r: dot(v, d) (3 mul, 2 add), front compare; in front:
# alignment (3), cohesion (3), separation compare; closer than
# separation_r: near-zero compare, separation_r - dl, one division, 3 mul,
# 3 sub (the length is the one computed before).
FLOCK_F = {"all": 10, "inside": 6, "front": 7, "separation": 9}
# plain_see: offset (3), d^2 (3 mul, 2 add), compare; inside: 3 adds
PLAIN_F = {"all": 9, "inside": 3}
# sph_density: offset (3), r^2 (5), h^2 - r^2, compare; inside: poly6 q^3 (3 mul), add
SPH_DENSITY_F = {"all": 10, "inside": 4}
# sph_forces: offset (3), length (5 + sqrt), compare; inside: h - rl, 1e-5
# compare, spiky w w (2), pa + pb, product, 2 rho_b, two divisions, r s (3),
# add (3), v_b - v_a (3), visc w, / rho_b, dv s (3), add (3)
SPH_FORCES_F = {"all": 10, "inside": 27}
# the differentiable forward (k_df_step) per neighbour pair (r^2 < R^2):
# offset (3), r^2 (5), compare, weight (1 - r^2/R^2)^3 (4), q = r^2 + h2,
# weight sum, per axis v_j - v_i, w (v_j - v_i), add, w d, add, / q, sub (21)
SMOOTH_F_PAIR = 36

# Compulsory bytes of one flocking see launch per agent: the seer reads p, v
# and the three accumulators (x, y, z) plus both counts (17 words) and writes
# the accumulators and counts (11 words); the seen side reads p and v (6).
B_SEE_PER_AGENT = 4 * (17 + 11 + 6)
# HBM bytes per see launch on the headline workload, from two rocprofv3
# --pmc runs (FETCH_SIZE, WRITE_SIZE in KiB per dispatch, means of the timed
# state's dispatches), each scaled by what the same counter reads for
# tay_gpu_copy_probe's 1 GiB copy (FETCH x 2.000, WRITE x 1.000;
# profiles/see_flock1M_pmc.md).  A counter run cannot be made from inside
# this process: the figure is carried here, with the commit of the kernel it
# was measured on, and re-measured whenever the see kernel changes.
SEE_TRAFFIC = {"agents": 1_000_000, "cell_size": 25.0,
               "reference": {"kernel": "see kernel of commit c7e55b7",
                             "bytes": int((43705.6 * 2.000 + 54675.3 * 1.000) * 1024)},
               # split over 4 lanes, lane walk 0 (profiles/see_flock1M_split_pmc.md)
               "split4_walk0": {"kernel": "split see kernel of commit be0ef4a, 8 wavefronts per SIMD",
                                "bytes": int((147273.2 * 2.000 + 404267.9 * 1.000) * 1024)}}
# The SPH line's see passes on 2M particles (profiles/see_sph2M_pmc.md),
# same calibration.  Algorithmic bytes per particle: forces, the seer reads
# p, v, pressure and both accumulators (15 words) and writes the
# accumulators (8), the seen side reads p, v, density and pressure (8);
# density, the seer reads p and density (4) and writes density (1), the seen
# side reads p (3).
SPH_TRAFFIC = {"particles": 2_000_000,
               "reference": {"kernel": "see kernels of commit c7e55b7",
                             "forces": int((69859.7 * 2.000 + 46875.0 * 1.000) * 1024),
                             "density": int((19950.6 * 2.000 + 7812.5 * 1.000) * 1024)},
               "balanced": {"kernel": "see kernels of commit aaf7979",
                            "forces": int((128419.3 * 2.000 + 124785.2 * 1.000) * 1024),
                            "density": int((19950.9 * 2.000 + 7812.5 * 1.000) * 1024)},
               # the density pass split over 4 lanes (profiles/see_flock1M_split_pmc.md)
               "density_split4": int((19968.6 * 2.000 + 7812.7 * 1.000) * 1024)}
B_SPH_FORCES_PER_PARTICLE = 4 * (15 + 8 + 8)
B_SPH_DENSITY_PER_PARTICLE = 4 * (4 + 1 + 3)


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


def progress(msg):
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------ measuring

def see_only(model):
    from tay_amd import workloads as W
    return W.Model(model.dtype, model.passes[:1], model.part, model.dims)


def pair_counts_at(lib, positions, radius, cell):
    """(box pairs, pairs within Euclidean distance `radius`) among these
    positions, ordered, self excluded: the plain counting model's see pass,
    its box-test counter and its neighbour counts."""
    import numpy as np
    from tay_amd import agents as A
    from tay_amd import api
    from tay_amd import workloads as W
    rec = np.zeros(len(positions), A.PLAIN_AGENT)
    rec["p"] = positions
    m = W.plain_model(radius)
    with W.Simulation(lib, W.Model(m.dtype, m.passes[:1], cell), rec, api.TAY_GPU_Z_GRID, part=cell) as s:
        lib.tay_set_fetch_mode(s.state, api.TAY_FETCH_ON_DEMAND)
        lib.tay_set_pair_counting(s.state, 1)
        s.run(1)
        box = lib.pair_counts(s.state, 0)[1]
        lib.tay_fetch_agents(s.state, None)
        return box, int(s.agents()["ncount"].sum(dtype=np.int64))


def flock_shares(lib, recs, model, cell):
    """Branch shares of flock_see on these agents: P_in / P_see (inside r),
    P_front / P_see (in front), P_sep / P_see (closer than separation_r)."""
    import numpy as np
    from tay_amd import api
    from tay_amd import workloads as W
    with W.Simulation(lib, see_only(model), recs, api.TAY_GPU_Z_GRID, part=cell) as s:
        lib.tay_set_fetch_mode(s.state, api.TAY_FETCH_ON_DEMAND)
        lib.tay_set_pair_counting(s.state, 1)
        s.run(1)
        p_see = lib.pair_counts(s.state, 0)[1]
        lib.tay_fetch_agents(s.state, None)
        a = s.agents()
    _, p_in = pair_counts_at(lib, recs["p"], model.passes[0].radius, cell)
    front = int(a["cohesion_count"].sum(dtype=np.int64))
    sep = int(a["separation_count"].sum(dtype=np.int64))
    return {"p_see": p_see, "inside": p_in / p_see, "front": front / p_see, "separation": sep / p_see}


def f_pair_of(counts, shares):
    f = F_BOX_TEST + counts["all"]
    for k, v in counts.items():
        if k != "all":
            f += v * shares[k]
    return f


def roofline(flop_per_launch, launch_s, kernel, f_pair, note, traffic=None, **extra):
    achieved = flop_per_launch / launch_s / 1e12
    out = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
           "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic, "kernel": kernel, "f_pair": f_pair,
           "peak_single_issue": PEAK_FP32_SINGLE_TOPS, "frac_single_issue": achieved / PEAK_FP32_SINGLE_TOPS,
           "note": note}
    out.update(extra)
    return out


def timed_run(lib, sim, steps, pass_timing=True):
    """Wall time of one tay_run of `steps` steps in the default fetch mode
    (it ends by copying every agent to host storage), the see and act pass
    times (HIP events around every pass launch, when pass_timing is set),
    and the time of that final copy alone."""
    from tay_amd import api
    lib.tay_set_fetch_mode(sim.state, api.TAY_FETCH_EVERY_RUN)
    lib.tay_set_pass_timing(sim.state, 1 if pass_timing else 0)
    lib.tay_gpu_synchronize()
    t0 = time.perf_counter()
    sim.run(steps)
    wall = time.perf_counter() - t0
    times = [lib.pass_time(sim.state, i) for i in range(len(sim.model.passes))] if pass_timing else None
    t1 = time.perf_counter()
    lib.tay_fetch_agents(sim.state, None)
    fetch = time.perf_counter() - t1
    return wall, times, fetch


# ---------------------------------------------------------- CPU baselines

def oracle_libs():
    from tay_amd import api
    oracle = api.TayLib(os.path.join(ROOT, "oracle", "_build", "liboracle.so"))
    models = ctypes.CDLL(os.path.join(ROOT, "oracle", "_build", "libmodels.so"))
    return oracle, models


def oracle_run(recs, model, space, threads, steps, seer_limit=0, part=None, shared_mb=250):
    """(see-pass pairs/s, steps/s, wall s, pairs) of the oracle's restatement
    of a reference structure on these agents; seer_limit runs only that many
    seers of every see pass (a bounded sample)."""
    from tay_amd import workloads as W
    oracle, models = oracle_libs()
    oracle.tay_threads_start(threads, 0)
    try:
        with W.Simulation(oracle, model, recs, space, funcs=models, part=part, shared_mb=shared_mb) as s:
            if seer_limit:
                for i, ps in enumerate(model.passes):
                    if ps.kind == "see":
                        oracle.oracle_limit_seers(s.state, i, seer_limit)
            t0 = time.perf_counter()
            s.run(steps)
            wall = time.perf_counter() - t0
            see_s = oracle.oracle_get_see_seconds(s.state)
            pairs = sum(oracle.pair_counts(s.state, i)[1] for i, ps in enumerate(model.passes) if ps.kind == "see")
    finally:
        oracle.tay_threads_stop()
    return pairs / see_s if see_s else 0.0, steps / wall, wall, pairs


def cpu_baseline(recs, model, threads, steps, single_sample):
    """The oracle's CpuGrid (the reported baseline) and CpuZGrid on the same
    agents at `threads` threads, then both at one thread on the first
    `single_sample` seers of one step."""
    from tay_amd import api
    out = {}
    for name, space in (("CpuGrid", api.TAY_CPU_GRID), ("CpuZGrid", api.TAY_CPU_Z_GRID)):
        progress("cpu baseline: %s, %d threads" % (name, threads))
        out[name] = oracle_run(recs, model, space, threads, steps)
    single = {}
    for name, space in (("CpuGrid", api.TAY_CPU_GRID), ("CpuZGrid", api.TAY_CPU_Z_GRID)):
        progress("cpu baseline: %s, 1 thread, %d seers" % (name, single_sample))
        rate, _, wall, pairs = oracle_run(recs, model, space, 1, 1, seer_limit=single_sample)
        single[name] = {"value": rate, "unit": "pair interactions/s", "cores": 1, "pairs": pairs,
                        "sample": "the see pass of the first %d seers (after the sort) of one step; %.1f s wall"
                                  % (single_sample, wall)}
    value, sps, wall, _ = out["CpuGrid"]
    z_value, z_sps, z_wall, _ = out["CpuZGrid"]
    return {
        "value": value,
        "unit": "pair interactions/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d step(s) of CpuGrid (AoS, serial sort, threaded see/act, indirect call per pair) on the "
                  "same %d agents; %.1f s wall, %.3f steps/s; host %s, nproc %d"
                  % (steps, len(recs), wall, sps, cpu_model(), os.cpu_count() or 0),
        "steps_per_s": sps,
        "cpu_z_grid": {"value": z_value, "unit": "pair interactions/s", "steps_per_s": z_sps,
                       "sample": "the same %d step(s) on CpuZGrid (power-of-two grid, dense Morton cell table); "
                                 "%.1f s wall" % (steps, z_wall)},
        "single_thread": single,
    }


# -------------------------------------------------------------- workloads

def bench_flocking(lib, args):
    import numpy as np
    from tay_amd import api
    from tay_amd import workloads as W
    recs = W.flock_agents(args.agents, seed=1, clusters=64, sigma=30.0, box=2000.0)
    model = W.flock_model(radius=25.0, weights=(1.5, 1.0, 1.0), dt=0.1, max_speed=2.0)

    def setup(s, split):
        lib.tay_set_fetch_mode(s.state, api.TAY_FETCH_ON_DEMAND)
        if args.order_bits >= 0:
            lib.tay_set_order_bits(s.state, s.group, args.order_bits)
        if split:
            s.see_split(0, args.flock_split)
            lib.tay_set_lane_walk(s.state, args.flock_lane_walk)

    def window(advance_steps, warmup, steps, name, split=True):
        split = split and args.flock_split > 1
        # counting twin: same inputs, same trajectory, pair counts of the timed steps
        with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID, part=args.cell_size) as twin:
            setup(twin, split)
            lib.tay_set_pair_counting(twin.state, 1)
            advance(twin, advance_steps, name + " counting twin")
            if warmup:
                twin.run(warmup)
            twin.run(steps)
            p_see = lib.pair_counts(twin.state, 0)[1]
        with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID, part=args.cell_size) as sim:
            setup(sim, split)
            advance(sim, advance_steps, name + " timed state")
            if warmup:
                sim.run(warmup)
            lib.tay_fetch_agents(sim.state, None)
            first = sim.agents()
            wall, times, fetch = timed_run(lib, sim, steps)
            last = sim.agents()
        return p_see, wall, times, fetch, first, last

    p_see, wall, ((see_ms, see_launches), (act_ms, act_launches)), fetch, first, last = window(
        args.advance, args.warmup, args.steps, "flocking")
    progress("flocking: timed window done")
    shares_first = flock_shares(lib, first, model, args.cell_size)
    shares_last = flock_shares(lib, last, model, args.cell_size)
    shares = {k: 0.5 * (shares_first[k] + shares_last[k]) for k in ("inside", "front", "separation")}
    f_pair = f_pair_of(FLOCK_F, shares)
    see_avg_s = see_ms / 1e3 / see_launches
    flop_per_launch = p_see / args.steps * f_pair
    hbm = B_SEE_PER_AGENT * args.agents / see_avg_s / 1e9
    setting = "split%d_walk%d" % (args.flock_split, args.flock_lane_walk) if args.flock_split > 1 else "reference"
    traffic = SEE_TRAFFIC[setting]["bytes"] if (args.agents, args.cell_size, args.advance) == (
        SEE_TRAFFIC["agents"], SEE_TRAFFIC["cell_size"], 0) and setting in SEE_TRAFFIC else None

    out = {
        "metric": METRIC,
        "value": p_see / (see_ms / 1e3),
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
                   "see_split": {"lanes": args.flock_split, "lane_walk": args.flock_lane_walk}
                   if args.flock_split > 1 else None,
                   **({"order_bits": args.order_bits} if args.order_bits >= 0 else {}),
                   **({"advanced_steps": args.advance} if args.advance else {})},
        "steps_per_s": args.steps / wall,
        "steps_per_s_resident": args.steps / (wall - fetch),
        "fetch_ms_per_run": fetch * 1e3,
        "p_see_per_step": p_see / args.steps,
        "see_ms_per_step": see_ms / see_launches,
        "act_ms_per_step": act_ms / max(act_launches, 1),
        "build_ms_per_step": max((wall - fetch) * 1e3 / args.steps - see_ms / see_launches
                                 - act_ms / max(act_launches, 1), 0.0),
        "roofline": roofline(
            flop_per_launch, see_avg_s, "see (tay_pass_0, flock_see)", f_pair,
            "FP32 vector-issue bound, no matrix instructions; F = P_see x f_pair, f_pair = %d (box test) + %d "
            "(every pair) + %d x inside + %d x front + %d x separation, shares measured on the window's first "
            "and last agents; traffic from rocprofv3 FETCH_SIZE/WRITE_SIZE (profiles/see_flock1M_split_pmc.md "
            "for the split pass, profiles/see_flock1M_pmc.md unsplit)"
            % (F_BOX_TEST, FLOCK_F["all"], FLOCK_F["inside"], FLOCK_F["front"], FLOCK_F["separation"]),
            traffic=traffic, branch_shares=shares, branch_shares_first=shares_first,
            branch_shares_last=shares_last, traffic_bytes_algorithmic=B_SEE_PER_AGENT * args.agents),
        "roofline_hbm": {"achieved": hbm, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": hbm / PEAK_HBM_GBS,
                         "bytes_per_agent": B_SEE_PER_AGENT},
    }
    if args.flock_split > 1 and not args.no_reference_order:
        # the unsplit pass (reference sweep order, default lane walk) on the same window
        rp, rw, ((rsee, rl), _), rfetch, _, _ = window(args.advance, args.warmup, args.steps, "unsplit", split=False)
        out["reference_see_order"] = {"pairs_per_s": rp / (rsee / 1e3), "steps_per_s": args.steps / rw,
                                      "steps_per_s_resident": args.steps / (rw - rfetch),
                                      "ms_per_step": rw * 1e3 / args.steps, "see_ms_per_step": rsee / rl,
                                      "p_see_per_step": rp / args.steps,
                                      "note": "the see pass unsplit, each seer's pairs in the reference CPU "
                                              "grids' sweep order; same pairs and counts"}
        progress("flocking: unsplit window done")
    if args.steps < CONFIGURED_STEPS and not args.advance and not args.no_full_run:
        fp, fw, ((fsee, fl), _), ffetch, _, _ = window(0, 0, CONFIGURED_STEPS, "full run")
        out["full_run"] = {"steps": CONFIGURED_STEPS, "pairs_per_s": fp / (fsee / 1e3),
                           "steps_per_s": CONFIGURED_STEPS / fw, "steps_per_s_resident": CONFIGURED_STEPS / (fw - ffetch),
                           "ms_per_step": fw * 1e3 / CONFIGURED_STEPS, "see_ms_per_step": fsee / fl,
                           "p_see_per_step": fp / CONFIGURED_STEPS}
        progress("flocking: full run done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(recs, model, args.cpu_threads, args.cpu_steps,
                                           args.cpu_single_sample)
    return out


def bench_c1(lib, args):
    """10k plain agents: positions U[0,1000)^3, velocities U[-1,1]^3, radius
    40 (Euclidean, self excluded), cells of 80, 100 steps in one tay_run;
    CPU baselines CpuSimple, CpuGrid and CpuZGrid."""
    from tay_amd import api
    from tay_amd import workloads as W
    n, steps, radius = 10_000, 100, 40.0
    recs = W.plain_agents(n, seed=1)
    model = W.plain_model(radius)

    def state(split):
        sim = W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID)
        if split > 1:
            sim.see_split(0, split)
        return sim

    def runs(split):
        with state(split) as twin:
            lib.tay_set_fetch_mode(twin.state, api.TAY_FETCH_ON_DEMAND)
            lib.tay_set_pair_counting(twin.state, 1)
            twin.run(2)
            twin.run(steps)
            p_see = lib.pair_counts(twin.state, 0)[1]
        # a step of this state takes about 0.1 ms, so the events that time
        # the passes (a few microseconds each) would show in its wall time:
        # steps/s comes from a run without them, the pass times from a
        # second run of the same steps
        with state(split) as sim:
            sim.run(2)
            wall, _, fetch = timed_run(lib, sim, steps, pass_timing=False)
        with state(split) as sim:
            sim.run(2)
            _, ((see_ms, launches), _), _ = timed_run(lib, sim, steps)
            p_in = int(sim.agents()["last_ncount"].sum())
        return p_see, wall, fetch, see_ms, launches, p_in

    p_see, wall, fetch, see_ms, launches, p_in = runs(args.c1_split)
    shares = {"inside": p_in / (p_see / steps)}
    f_pair = f_pair_of(PLAIN_F, shares)
    out = {"workload": "plain-10k", "agents": n, "radius": radius, "cell_size": 2 * radius, "steps": steps,
           "see_split": args.c1_split if args.c1_split > 1 else None, "p_see_per_step": p_see / steps,
           "pairs_per_s": p_see / (see_ms / 1e3), "steps_per_s": steps / wall,
           "steps_per_s_resident": steps / (wall - fetch), "ms_per_step": wall * 1e3 / steps,
           "see_ms_per_step": see_ms / launches,
           "roofline": roofline(p_see / steps * f_pair, see_ms / 1e3 / launches, "see (tay_pass_0, plain_see)",
                                f_pair, "launch-latency bound: the whole state is 0.8 MB; F = P_see x f_pair, "
                                        "share inside r from the last step's counts", branch_shares=shares)}
    if args.c1_split > 1 and not args.no_reference_order:
        rp, rwall, rfetch, rsee, rl, _ = runs(1)
        out["reference_see_order"] = {"pairs_per_s": rp / (rsee / 1e3), "steps_per_s": steps / rwall,
                                      "steps_per_s_resident": steps / (rwall - rfetch),
                                      "ms_per_step": rwall * 1e3 / steps, "see_ms_per_step": rsee / rl,
                                      "note": "the see pass unsplit (reference sweep order); same pairs"}
    if not args.no_cpu_baseline:
        base = {}
        for name, space, k in (("CpuSimple", api.TAY_CPU_SIMPLE, 10), ("CpuGrid", api.TAY_CPU_GRID, steps),
                               ("CpuZGrid", api.TAY_CPU_Z_GRID, steps)):
            progress("c1 cpu baseline: %s" % name)
            rate, sps, w, _ = oracle_run(recs, model, space, args.cpu_threads, k)
            rate1, sps1, w1, _ = oracle_run(recs, model, space, 1, k)
            base[name] = {"pairs_per_s": rate, "steps_per_s": sps, "cores": args.cpu_threads,
                          "single_thread": {"pairs_per_s": rate1, "steps_per_s": sps1},
                          "sample": "%d steps; %.2f s wall (%.2f s at one thread)" % (k, w, w1)}
        out["cpu_baseline"] = {"value": base["CpuGrid"]["pairs_per_s"], "unit": "pair interactions/s",
                               "cores": args.cpu_threads, "kind": "port", "structures": base,
                               "sample": "CpuSimple 10 steps, CpuGrid and CpuZGrid %d steps, same inputs" % steps}
    return out


def ensemble_cpu_baseline(ens, part, tables, threads):
    """One step of every instance on the oracle's CpuGrid, states run one
    after another as the reference allows."""
    from tay_amd import api
    wall = see_s = 0.0
    pairs = 0
    for (rec, model, _), p, mb in zip(ens, part, tables):
        rate, sps, w, n = oracle_run(rec, model, api.TAY_CPU_GRID, threads, 1, part=p, shared_mb=mb)
        wall += w
        see_s += n / rate if rate else 0.0
        pairs += n
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
    # cell sizes per instance by density: about `target` agents per cell,
    # between 1 and 4 radii (dense instances get fine grids, sparse ones
    # coarse grids; below one radius a seer's box spans three or more cells
    # per axis)
    targets, fmin, fmax = (4.0, 8.0, 16.0), 1.0, 4.0
    widths, walks = (64, 32), (16, 64)
    # tables sized for the smallest target's cells, shared by every run below
    tables = [W.cell_table_mb(side, W.density_part(rec, r, min(targets), fmin, fmax))
              for (rec, _, side), r in zip(ens, radius)]
    steps, warm = args.ensemble_steps, 2

    parts = np.zeros((len(ens), 4), np.float32)
    with W.Batch(lib, ens, radius, tables) as b:
        tg = (ctypes.c_float * len(targets))(*targets)
        w = (ctypes.c_uint * len(widths))(*widths)
        lw = (ctypes.c_uint * len(walks))(*walks)
        table = (ctypes.c_double * (len(targets) * len(widths) * len(walks)))()
        bt, bw, bl = ctypes.c_float(), ctypes.c_uint(), ctypes.c_uint()
        if not lib.tay_gpu_batch_autotune_density(b.batch, tg, len(targets), fmin, fmax, w, len(widths), lw,
                                                  len(walks), 3, table, ctypes.byref(bt),
                                                  parts.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                  ctypes.byref(bw), ctypes.byref(bl)):
            raise SystemExit("autotune: " + lib.tay_gpu_last_error().decode())
    part = [float(x) for x in parts[:, 0]]
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
    progress("ensemble: batched runs done")

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
    progress("ensemble: per-instance runs done")

    f_pair = F_BOX_TEST + FLOCK_F["all"]
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
        "autotune": {"agents_per_cell": list(targets), "cell_factor_bounds": [fmin, fmax],
                     "tile_widths": list(widths), "lane_walks": list(walks),
                     "ms_per_step": [[[table[(i * len(widths) + j) * len(walks) + k] for k in range(len(walks))]
                                      for j in range(len(widths))] for i in range(len(targets))],
                     "chosen": {"agents_per_cell": bt.value, "tile_width": width, "lane_walk": walk,
                                "cell_factors": {"min": float(min(p / r for p, r in zip(part, radius))),
                                                 "median": float(np.median([p / r for p, r in zip(part, radius)])),
                                                 "max": float(max(p / r for p, r in zip(part, radius)))}}},
        "roofline": roofline(p_see / steps * f_pair, see_ms / 1e3 / launches, "batched see (tay_batch_pass_0)",
                             f_pair, "f_pair = the operations every box-passing pair of flock_see executes (box "
                                     "test and the always-taken part), a lower bound"),
    }
    out["speedup_pairs_per_s"] = out["batched"]["pairs_per_s"] / out["per_instance"]["pairs_per_s"]
    out["speedup_steps_per_s"] = out["batched"]["steps_per_s"] / out["per_instance"]["steps_per_s"]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = ensemble_cpu_baseline(ens, part, tables, args.cpu_threads)
    return out


SEE_ORDERS = {"reference": 0, "balanced": 1}   # TAY_SEE_ORDER_REFERENCE / _BALANCED


def bench_sph(lib, args):
    """SPH dam break: two see passes (density, forces) and two act passes
    per step on 2M particles, in the see order args.sph_see_order (the
    timed line) and, for comparison, in the other order."""
    from tay_amd import api
    from tay_amd import workloads as W
    recs, lo, hi = W.sph_particles(args.sph_particles, h=1.0)
    ctx = W.sph_context(h=1.0, dt=1e-3, box_min=lo, box_max=hi)
    model = W.sph_model(ctx, 1.0)
    steps, warm = args.sph_steps, 2
    order = SEE_ORDERS[args.sph_see_order]
    dsplit = args.sph_density_split

    def split(sim):
        # the density pass split over dsplit lanes (it then runs in the
        # sweep order); the force pass in the state's see order
        if dsplit > 1:
            sim.see_split(0, dsplit)

    with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID) as twin:
        lib.tay_set_fetch_mode(twin.state, api.TAY_FETCH_ON_DEMAND)
        lib.tay_set_see_order(twin.state, order)
        split(twin)
        lib.tay_set_pair_counting(twin.state, 1)
        twin.run(warm)
        twin.run(steps)
        pairs = [lib.pair_counts(twin.state, i)[1] for i in (0, 2)]
    runs = {}
    for name in [args.sph_see_order] + [n for n in SEE_ORDERS if n != args.sph_see_order]:
        with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID) as sim:
            lib.tay_set_fetch_mode(sim.state, api.TAY_FETCH_ON_DEMAND)
            lib.tay_set_see_order(sim.state, SEE_ORDERS[name])
            if name == args.sph_see_order:
                split(sim)
            sim.run(warm)
            lib.tay_fetch_agents(sim.state, None)
            first = sim.agents()
            runs[name] = timed_run(lib, sim, steps) + (first, sim.agents())
        progress("sph: timed run done (%s see order)" % name)
    wall, times, fetch, first, last = runs[args.sph_see_order]
    see = [times[i][0] for i in (0, 2)]
    act = [times[i][0] for i in (1, 3)]
    other = [n for n in runs if n != args.sph_see_order][0]
    o_wall, o_times, _, _, _ = runs[other]
    inside = []
    for a in (first, last):
        box, ball = pair_counts_at(lib, a["p"], 1.0, 1.0)
        # the density pass sees every particle itself too (inside h); the
        # force pass does not
        inside.append(((ball + len(a)) / (box + len(a)), ball / box))
    share_d = 0.5 * (inside[0][0] + inside[1][0])
    share_f = 0.5 * (inside[0][1] + inside[1][1])
    measured = len(recs) == SPH_TRAFFIC["particles"]
    traffic = dict(SPH_TRAFFIC[args.sph_see_order])
    if dsplit > 1:
        traffic["density"] = SPH_TRAFFIC.get("density_split%d" % dsplit)
    f_d = f_pair_of(SPH_DENSITY_F, {"inside": share_d})
    f_f = f_pair_of(SPH_FORCES_F, {"inside": share_f})
    out = {"workload": "SPH dam break, %d particles" % len(recs), "particles": len(recs), "h": 1.0, "dt": 1e-3,
           "steps": steps, "see_order": args.sph_see_order,
           "density_split": dsplit if dsplit > 1 else None, "p_see_per_step": [p / steps for p in pairs],
           "pairs_per_s": sum(pairs) / (sum(see) / 1e3), "steps_per_s": steps / wall,
           "steps_per_s_resident": steps / (wall - fetch), "ms_per_step": wall * 1e3 / steps,
           "see_ms_per_step": [t / steps for t in see], "act_ms_per_step": [t / steps for t in act],
           "roofline": roofline(pairs[1] / steps * f_f, see[1] / 1e3 / steps, "see (tay_pass_2, sph_forces)", f_f,
                                "the force pass, the longer of the two see kernels; share of box pairs inside h "
                                "measured on the window's first and last particles; traffic from rocprofv3 "
                                "FETCH_SIZE/WRITE_SIZE (profiles/see_sph2M_pmc.md)",
                                traffic=traffic["forces"] if measured else None,
                                traffic_bytes_algorithmic=B_SPH_FORCES_PER_PARTICLE * len(recs),
                                branch_shares={"inside": share_f},
                                density=roofline(pairs[0] / steps * f_d, see[0] / 1e3 / steps,
                                                 "see (tay_pass_0, sph_density)", f_d, "",
                                                 traffic=traffic["density"] if measured and traffic["density"]
                                                 else None,
                                                 traffic_bytes_algorithmic=B_SPH_DENSITY_PER_PARTICLE * len(recs),
                                                 branch_shares={"inside": share_d})),
           other + "_see_order": {"note": "both see passes unsplit in this order",
                                  "steps_per_s": steps / o_wall, "ms_per_step": o_wall * 1e3 / steps,
                                  "see_ms_per_step": [o_times[i][0] / steps for i in (0, 2)],
                                  "pairs_per_s": sum(pairs) / (sum(o_times[i][0] for i in (0, 2)) / 1e3)}}
    if not args.no_cpu_baseline:
        base = {}
        for threads in (args.cpu_threads, 1):
            progress("sph cpu baseline: CpuZGrid, %d threads" % threads)
            limit = 0 if threads > 1 else args.cpu_single_sample
            rate, sps, w, n = oracle_run(recs, model, api.TAY_CPU_Z_GRID, threads, 1, seer_limit=limit)
            base[threads] = (rate, sps, w, limit)
        rate, sps, w, _ = base[args.cpu_threads]
        rate1, _, w1, limit = base[1]
        out["cpu_baseline"] = {"value": rate, "unit": "pair interactions/s", "cores": args.cpu_threads,
                               "kind": "port", "steps_per_s": sps, "structure": "CpuZGrid (cells of h, as "
                               "m_sph/main.c:70)",
                               "sample": "1 step of CpuZGrid on the same %d particles; %.1f s wall" % (len(recs), w),
                               "single_thread": {"value": rate1, "cores": 1,
                                                 "sample": "both see passes of the first %d seers of one step; "
                                                           "%.1f s wall" % (limit, w1)}}
    return out


def dflock_oracle():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_build", "libdflock.so"))

    class Prm(ctypes.Structure):
        _fields_ = [("radius", ctypes.c_double), ("h2", ctypes.c_double), ("dt", ctypes.c_double),
                    ("damping", ctypes.c_double)]
    d = ctypes.POINTER(ctypes.c_double)
    lib.oracle_dflock_forward.argtypes = [ctypes.c_uint, d, d, d, ctypes.POINTER(Prm), ctypes.c_uint, d, d]
    lib.oracle_dflock_gradient.argtypes = [ctypes.c_uint, d, d, d, ctypes.POINTER(Prm), ctypes.c_uint, d, d, d,
                                           d, d]
    return lib, Prm


def bench_adjoint(lib, args):
    """Differentiable flocking: 65,536 agents, 50 steps, checkpoints every
    10; forward and backward device time at the same (FP32) accuracy, then
    calibration of the three weights to a target trajectory made with hidden
    weights; the FP64 CPU forward and adjoint timed on the same problem."""
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
            pairs = f.pair_count()
            f.backward((2.0 * (p - target) / p.size).astype(np.float32))
            fm, bm = f.times()
            if r:
                fwd += fm
                bwd += bm
        fwd_ms = fwd / reps
        out.update({"forward_ms": fwd_ms, "backward_ms": bwd / reps, "backward_over_forward": bwd / fwd,
                    "forward_steps_per_s": steps / (fwd_ms / 1e3), "max_neighbours": f.max_neighbours(),
                    "pairs_per_step": pairs / steps, "forward_pairs_per_s": pairs / (fwd_ms / 1e3),
                    "roofline": roofline(pairs * SMOOTH_F_PAIR, fwd_ms / 1e3, "forward (50 steps: grid build + "
                                         "k_df_step)", SMOOTH_F_PAIR,
                                         "F = neighbour pairs x %d over the whole forward call, grid builds "
                                         "included in the time" % SMOOTH_F_PAIR)})
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
    progress("adjoint: GPU runs done")
    if not args.no_cpu_baseline:
        o, Prm = dflock_oracle()
        prm = Prm(20.0, 1.0, 0.1, 0.05)
        dp = ctypes.POINTER(ctypes.c_double)
        P0, V0 = np.ascontiguousarray(p0, np.float64), np.ascontiguousarray(v0, np.float64)
        W64 = np.ones(3)
        pf, vf = np.zeros((n, 3)), np.zeros((n, 3))
        t0 = time.perf_counter()
        o.oracle_dflock_forward(n, P0.ctypes.data_as(dp), V0.ctypes.data_as(dp), W64.ctypes.data_as(dp),
                                ctypes.byref(prm), steps, pf.ctypes.data_as(dp), vf.ctypes.data_as(dp))
        t_fwd = time.perf_counter() - t0
        loss = ctypes.c_double()
        gp, gv, gw = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(3)
        tgt = np.ascontiguousarray(target)
        t0 = time.perf_counter()
        o.oracle_dflock_gradient(n, P0.ctypes.data_as(dp), V0.ctypes.data_as(dp), W64.ctypes.data_as(dp),
                                 ctypes.byref(prm), steps, tgt.ctypes.data_as(dp), ctypes.byref(loss),
                                 gp.ctypes.data_as(dp), gv.ctypes.data_as(dp), gw.ctypes.data_as(dp))
        t_grad = time.perf_counter() - t0
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {"value": pairs / t_fwd, "unit": "pair interactions/s", "cores": threads,
                               "kind": "port", "forward_ms": t_fwd * 1e3, "gradient_ms": t_grad * 1e3,
                               "backward_ms": (t_grad - t_fwd) * 1e3,
                               "sample": "the FP64 oracle (oracle/dflock.c, OpenMP) on the same 50 steps: one "
                                         "forward, then forward + reverse with all states kept"}
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
    ap.add_argument("--order-bits", type=int, default=-1,
                    help="sub-cell bits of the storage order inside a cell (0-3; default: the library's)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cpu-single-sample", type=int, default=100_000,
                    help="seers of the one-thread CPU baselines of the large workloads")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-full-run", action="store_true", help="skip the configured 200-step flocking run")
    ap.add_argument("--ensemble-instances", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--no-c1", action="store_true", help="skip the 10k plain-agent measurement")
    ap.add_argument("--sph-particles", type=int, default=2_000_000)
    ap.add_argument("--sph-steps", type=int, default=500)
    ap.add_argument("--flock-split", type=int, choices=(1, 2, 4, 8), default=4,
                    help="lanes per seer of the flocking see pass (tay_set_see_split; 1: unsplit)")
    ap.add_argument("--flock-lane-walk", type=int, default=0,
                    help="tay_set_lane_walk for the split flocking pass (0: lane groups halve, never walk)")
    ap.add_argument("--no-reference-order", action="store_true",
                    help="skip the unsplit flocking and 10k plain runs timed beside the split ones")
    ap.add_argument("--c1-split", type=int, choices=(1, 2, 4, 8), default=8,
                    help="lanes per seer of the 10k plain see pass (1: unsplit)")
    ap.add_argument("--sph-density-split", type=int, choices=(1, 2, 4, 8), default=4,
                    help="lanes per seer of the SPH density pass of the timed state (1: unsplit)")
    ap.add_argument("--sph-see-order", choices=sorted(SEE_ORDERS), default="balanced",
                    help="see order of the timed SPH run (the other order is timed too, for comparison)")
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
    lib = api.TayLib(api.LIBRARY)
    if not lib.tay_gpu_available():
        raise SystemExit("no GPU: " + lib.tay_gpu_last_error().decode())

    out = bench_flocking(lib, args)
    if args.copy_probe:
        probe_ms = ctypes.c_double()
        if not lib.tay_gpu_copy_probe(args.copy_probe << 20, ctypes.byref(probe_ms)):
            raise SystemExit("copy probe: " + lib.tay_gpu_last_error().decode())
        out["copy_probe"] = {"bytes": args.copy_probe << 20, "ms": probe_ms.value,
                             "GB/s": 2 * (args.copy_probe << 20) / probe_ms.value / 1e6}
    if not args.no_c1:
        out["c1"] = bench_c1(lib, args)
        progress("c1 done")
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
