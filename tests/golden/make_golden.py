"""Writes the golden fixtures of tests/golden from the reference's own code,
compiled from its sources in this container by `make -C oracle ref`.

Simulation fixtures (oracle/_ref/ref_driver: the reference's CpuSimple,
CpuGrid and CpuZGrid sorts, sweeps and pairing loops with the test/agent.c
models), per structure and step: storage order, tags (part_i,
cell_agent_i), the see passes' sums and counts, positions, velocities,
f_buffer and the grid geometry.
  ref_test_model_{uniform,clump}.npz  800 point agents, one group
  ref_test_model_c1.npz               10,000 point agents in [0, 1000)^3,
                                      radius 40, cells of 80 (the 10k plain
                                      workload's shape)
  ref_nonpoint.npz                    box agents (test/test_nonpoint.c), Simple
  ref_combo.npz                       two groups seeing themselves and each
                                      other (test/test_combo.c:114-119): point
                                      groups in every structure, box groups,
                                      and a point group with a box group

Model fixtures (oracle/_ref/libref_*.so: the reference plug-ins' model
functions, m_flocking/agent.c, m_sph/agent.c, test/agent.c), on seeded
agents whose inputs reach every branch: per pair, the seer after one
agent_see call; per agent, the record after one act call.
  ref_models.npz

The unused fourth lane of float4 outputs is zeroed before saving: the
reference's helpers leave it undefined (m_flocking/taystd.c:137-143,
test/taystd.c:91-178) and it is never compared.

usage: python tests/golden/make_golden.py
"""
import ctypes as C
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tay_amd import agents as A  # noqa: E402
from tay_amd import workloads as W  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
DRIVER = os.path.join(REF, "ref_driver")

# name: (agents, seed, radius, partition size, fraction clumped, steps)
SCENARIOS = {
    "uniform": (800, 1, 100.0, 200.0, 0.0, 3),
    "clump": (800, 2, 60.0, 120.0, 0.2, 3),
    "c1": (10000, 3, 40.0, 80.0, 0.0, 1),
}
STRUCTURES = ["simple", "grid", "zgrid"]

# the per-pair harness: pair k sits at site k of a 10^3 lattice of spacing
# SITE_SPACING centred on the origin, far beyond every see radius used
SITES = 10
SITE_SPACING = 200.0
PAIRS = SITES ** 3


def run_driver(steps, radius, groups):
    """groups: list of (kind, structure, records, part).  Returns per group
    (dumps[steps, 3, n], geometry[steps, 12])."""
    with tempfile.TemporaryDirectory() as tmp:
        args = [DRIVER, str(steps), repr(radius), "250", os.path.join(tmp, "out.bin")]
        for gi, (kind, structure, recs, part) in enumerate(groups):
            path = os.path.join(tmp, "in%d.bin" % gi)
            recs.tofile(path)
            args += [kind, structure, str(len(recs)), repr(part), path]
        subprocess.run(args, check=True)
        raw = np.fromfile(os.path.join(tmp, "out.bin"), dtype=np.uint8)
    sizes = [len(r) * r.dtype.itemsize for _, _, r, _ in groups]
    per_step = 3 * sum(sizes)
    out = [np.zeros((steps, 3, len(r)), r.dtype) for _, _, r, _ in groups]
    for step in range(steps):
        off = step * per_step
        for phase in range(3):
            for gi, (_, _, r, _) in enumerate(groups):
                out[gi][step, phase] = np.frombuffer(raw[off:off + sizes[gi]].tobytes(), dtype=r.dtype)
                off += sizes[gi]
    geo = np.frombuffer(raw[steps * per_step:].tobytes(), dtype=np.float32).reshape(steps, len(groups), 12)
    return [(out[gi], geo[:, gi]) for gi in range(len(groups))]


def store(out, prefix, dumps, geometry, structure, point=True):
    sort, see, act = dumps[:, 0], dumps[:, 1], dumps[:, 2]
    pos = "p" if point else "min"
    out[prefix + "_order"] = act["result_index"].astype(np.int32)
    out[prefix + "_part_i"] = act["part_i"]
    out[prefix + "_cell_agent_i"] = act["cell_agent_i"]
    out[prefix + "_b_buffer"] = see["b_buffer"][:, :, :3]
    out[prefix + "_b_count"] = see["b_buffer_count"]
    out[prefix + "_" + pos] = act[pos][:, :, :3]
    if not point:
        out[prefix + "_max"] = act["max"][:, :, :3]
    out[prefix + "_v"] = act["v"][:, :, :3]
    out[prefix + "_f_buffer"] = act["f_buffer"][:, :, :3]
    if structure != "simple":
        out[prefix + "_origin"] = geometry[:, 0:3]
        out[prefix + "_cell_size"] = geometry[:, 4:7]
        out[prefix + "_cell_count"] = geometry[:, 8:11].view(np.int32)
    assert np.array_equal(sort["result_index"], act["result_index"])


def simulation_fixtures():
    for name, (n, seed, radius, part, clump, steps) in SCENARIOS.items():
        recs = W.test_agents(n, seed=seed, box=1000.0, clump=clump)
        out = {"p": recs["p"], "v": recs["v"], "radius": np.float32(radius), "part": np.float32(part),
               "steps": np.int32(steps)}
        for structure in STRUCTURES:
            (dumps, geometry), = run_driver(steps, radius, [("point", structure, recs, part)])
            store(out, structure, dumps, geometry, structure)
        save("ref_test_model_%s.npz" % name, out)

    # non-point agents: Simple only (the reference's grids take points only,
    # tay/state_compile.c:25-27)
    recs = W.box_agents(1000, seed=41)
    radius, steps = 20.0, 3
    out = {"min": recs["min"], "max": recs["max"], "v": recs["v"], "radius": np.float32(radius),
           "steps": np.int32(steps)}
    (dumps, geometry), = run_driver(steps, radius, [("box", "simple", recs, radius)])
    store(out, "simple", dumps, geometry, "simple", point=False)
    save("ref_nonpoint.npz", out)

    # two groups: a sees a, a sees b, b sees b, b sees a, then both act
    a = W.test_agents(600, seed=42)
    b = W.test_agents(400, seed=43, clump=0.2)
    ba = W.box_agents(500, seed=44)
    bb = W.box_agents(300, seed=45)
    radius, steps = 50.0, 3
    out = {"a_p": a["p"], "a_v": a["v"], "b_p": b["p"], "b_v": b["v"],
           "ba_min": ba["min"], "ba_max": ba["max"], "ba_v": ba["v"],
           "bb_min": bb["min"], "bb_max": bb["max"], "bb_v": bb["v"],
           "radius": np.float32(radius), "steps": np.int32(steps)}
    for structure in STRUCTURES:
        (da, ga), (db, gb) = run_driver(steps, radius, [("point", structure, a, radius),
                                                       ("point", structure, b, 2.0 * radius)])
        store(out, "pp_%s_a" % structure, da, ga, structure)
        store(out, "pp_%s_b" % structure, db, gb, structure)
    (da, ga), (db, gb) = run_driver(steps, 20.0, [("box", "simple", ba, 20.0), ("box", "simple", bb, 20.0)])
    store(out, "bb_a", da, ga, "simple", point=False)
    store(out, "bb_b", db, gb, "simple", point=False)
    (da, ga), (db, gb) = run_driver(steps, 20.0, [("point", "simple", a, 20.0), ("box", "simple", ba, 20.0)])
    store(out, "pb_a", da, ga, "simple")
    store(out, "pb_b", db, gb, "simple", point=False)
    save("ref_combo.npz", out)


# ------------------------------------------------------------ model fixtures

def sites():
    """Pair-site centres: a 10^3 lattice of spacing 200 centred on 0."""
    i = np.arange(PAIRS)
    idx = np.stack([i % SITES, (i // SITES) % SITES, i // (SITES * SITES)], 1)
    return ((idx - SITES // 2) * SITE_SPACING).astype(np.float32)


def f4(xyz):
    out = np.zeros((len(xyz), 4), np.float32)
    out[:, :3] = xyz
    return out


def offsets(g, n, radius, box):
    """Seen-minus-seer offsets inside the per-axis box [-box, box]^3: a
    quarter uniform in the box, a third inside the radius, a quarter inside
    half the radius, the rest a few units of the last place or zero."""
    k1, k2, k3 = n // 4, n // 3, n // 4
    k4 = n - k1 - k2 - k3
    box_u = g.uniform(-box, box, size=(k1, 3))
    ball = g.normal(size=(k2 + k3, 3))
    ball /= np.linalg.norm(ball, axis=1, keepdims=True)
    rad = np.concatenate([radius * g.uniform(0, 1, k2) ** (1 / 3), 0.5 * radius * g.uniform(0, 1, k3) ** (1 / 3)])
    tiny = g.uniform(-1e-6, 1e-6, size=(k4, 3)) * (g.uniform(size=(k4, 1)) < 0.5)
    off = np.concatenate([box_u, ball * rad[:, None], tiny]).astype(np.float32)
    return off[g.permutation(n)]


def zero_w(rec):
    for name in rec.dtype.names:
        if rec.dtype[name].shape == (4,):
            rec[name][:, 3] = 0.0
    return rec


def apply_see(fn, seers, seens, ctx):
    out = seers.copy()
    for k in range(len(out)):
        fn(out[k:k + 1].ctypes.data, seens[k:k + 1].ctypes.data, ctx.ctypes.data)
    return zero_w(out)


def apply_act(fn, agents, ctx):
    out = agents.copy()
    for k in range(len(out)):
        fn(out[k:k + 1].ctypes.data, ctx.ctypes.data)
    return zero_w(out)


def apply_act1(fn, agents):
    out = agents.copy()
    for k in range(len(out)):
        fn(out[k:k + 1].ctypes.data)
    return zero_w(out)


def load_ref(name, see=(), act=(), act1=()):
    """A reference plug-in library with its functions' pointer signatures."""
    lib = C.CDLL(os.path.join(REF, "libref_%s.so" % name))
    for names, nargs in ((see, 3), (act, 2), (act1, 1)):
        for fname in names:
            fn = getattr(lib, fname)
            fn.argtypes = [C.c_void_p] * nargs
            fn.restype = None
    return lib


def boid_cases(g):
    """m_flocking/agent.c inputs.  Seer accumulators and counts start
    nonzero, so the adds are checked, not only the first term."""
    s = sites()
    a = np.zeros(PAIRS, A.BOID)
    a["p"] = f4(s + g.uniform(-5, 5, size=(PAIRS, 3)))
    a["dir"] = f4(W.unit_vectors(PAIRS, g))
    a["color"] = np.arange(PAIRS)
    a["speed"] = 1.0
    for f in ("separation", "alignment", "cohesion"):
        a[f] = f4(g.normal(scale=5.0, size=(PAIRS, 3)))
    a["cohesion_count"] = g.integers(0, 20, PAIRS)
    a["separation_count"] = g.integers(0, 20, PAIRS)
    b = np.zeros(PAIRS, A.BOID)
    b["p"] = f4(a["p"][:, :3] + offsets(g, PAIRS, 20.0, 25.0))
    b["dir"] = f4(W.unit_vectors(PAIRS, g))
    b["color"] = np.arange(PAIRS) + PAIRS
    b["speed"] = 1.0
    # the near-zero branch of normalize_to: the pair at the origin site, in
    # front of the seer at a distance below 1e-6
    k = int(np.argmin(np.abs(s).sum(1)))
    a["p"][k] = 0.0
    a["dir"][k] = [1.0, 0.0, 0.0, 0.0]
    b["p"][k] = [1e-7, 0.0, 0.0, 0.0]
    ctx = np.zeros(1, A.BOID_SEE_CONTEXT)
    ctx["r"] = 20.0
    ctx["separation_r"] = np.float32(20.0) * np.float32(0.5)

    act = np.zeros(PAIRS, A.BOID)
    act["p"] = f4(g.uniform(-700, 700, size=(PAIRS, 3)))
    act["dir"] = f4(W.unit_vectors(PAIRS, g))
    act["speed"] = g.uniform(0.5, 2.0, PAIRS)
    for f in ("separation", "alignment", "cohesion"):
        act[f] = f4(g.normal(scale=3.0, size=(PAIRS, 3)))
    act["cohesion_count"] = g.integers(0, 30, PAIRS) * (g.uniform(size=PAIRS) < 0.8)
    act["separation_count"] = g.integers(0, 30, PAIRS) * (g.uniform(size=PAIRS) < 0.6)
    act["color"] = np.arange(PAIRS)
    # normalize's near-zero branch: heading and steering cancel
    act["p"][:16] = f4(g.uniform(-100, 100, size=(16, 3)))
    act["dir"][:16] = 0.0
    act["cohesion_count"][:16] = 0
    act["separation_count"][:16] = 0
    act_ctx = np.zeros(1, A.BOID_ACT_CONTEXT)
    return a, b, ctx, act, act_ctx


def sph_contexts():
    """Two kernel scales: h = 1 (the dam-break workload) and the reference
    plug-in's own h from 20 neighbours per influence volume
    (m_sph/main.c:44-61: 100,000 particles of 0.05 in water)."""
    out = []
    particles, m, rho = 100000, np.float32(0.05), np.float32(998.29)
    volume = np.float32(particles) * m / rho
    h_ref = np.float32(math.pow(float(np.float32(3.0) * (np.float32(20) * (volume / np.float32(particles)))
                                      / (np.float32(4.0) * np.float32(math.pi))), 1.0 / 3.0))
    for h, dt, lo, hi in ((1.0, 1e-3, [0, 0, 0, 0], [20, 20, 20, 0]),
                          (float(h_ref), 0.004, [-3, -2, -1, 0], [3, 2, 1, 0])):
        c = W.sph_context(h=h, dt=dt, box_min=np.array(lo, np.float32), box_max=np.array(hi, np.float32))
        c["surface_tension"] = 0.0728
        c["surface_tension_threshold"] = 7.065
        out.append(c)
    return out


def sph_cases(g, ctx):
    h = float(ctx["h"][0])
    s = sites()
    a = np.zeros(PAIRS, A.SPH_PARTICLE)
    a["p"] = f4(s + g.uniform(-5, 5, size=(PAIRS, 3)))
    a["density"] = g.uniform(900.0, 1100.0, PAIRS)
    a["pressure"] = g.uniform(-50.0, 500.0, PAIRS)
    a["v"] = f4(g.normal(size=(PAIRS, 3)))
    a["vh"] = a["v"]
    a["pressure_accum"] = f4(g.normal(scale=10.0, size=(PAIRS, 3)))
    a["viscosity_accum"] = f4(g.normal(scale=10.0, size=(PAIRS, 3)))
    b = np.zeros(PAIRS, A.SPH_PARTICLE)
    b["p"] = f4(a["p"][:, :3] + offsets(g, PAIRS, h, h))
    b["density"] = g.uniform(900.0, 1100.0, PAIRS)
    b["pressure"] = g.uniform(-50.0, 500.0, PAIRS)
    b["v"] = f4(g.normal(size=(PAIRS, 3)))
    # coincident partners (the density pass sees self) and r < 1e-5
    b["p"][:32] = a["p"][:32]

    lo, hi = ctx["min"][0][:3], ctx["max"][0][:3]
    act = np.zeros(PAIRS, A.SPH_PARTICLE)
    span = hi - lo
    act["p"] = f4(lo + g.uniform(-0.1, 1.1, size=(PAIRS, 3)) * span)
    act["v"] = f4(g.normal(size=(PAIRS, 3)))
    act["vh"] = f4(g.normal(size=(PAIRS, 3)))
    act["pressure_accum"] = f4(g.normal(scale=1000.0, size=(PAIRS, 3)))
    act["viscosity_accum"] = f4(g.normal(scale=100.0, size=(PAIRS, 3)))
    act["density"] = g.uniform(0.5, 2.0, PAIRS) * ctx["density"][0]
    act["pressure"] = g.uniform(-50.0, 500.0, PAIRS)
    return a, b, act


def test_model_cases(g):
    s = sites()
    a = np.zeros(PAIRS, A.AGENT)
    a["p"] = f4(s + g.uniform(-5, 5, size=(PAIRS, 3)))
    a["b_buffer"] = f4(g.normal(scale=50.0, size=(PAIRS, 3)))
    a["b_buffer_count"] = g.integers(0, 40, PAIRS)
    a["result_index"] = np.arange(PAIRS)
    b = np.zeros(PAIRS, A.AGENT)
    b["p"] = f4(a["p"][:, :3] + offsets(g, PAIRS, 20.0, 25.0))
    b["result_index"] = np.arange(PAIRS) + PAIRS
    ba = np.zeros(PAIRS, A.BOX_AGENT)
    ba["min"] = f4(s + g.uniform(-5, 5, size=(PAIRS, 3)))
    ba["max"] = f4(ba["min"][:, :3] + g.uniform(0.5, 10.0, size=(PAIRS, 3)))
    ba["b_buffer"] = f4(g.normal(scale=50.0, size=(PAIRS, 3)))
    ba["b_buffer_count"] = g.integers(0, 40, PAIRS)
    ba["result_index"] = np.arange(PAIRS)
    bb = np.zeros(PAIRS, A.BOX_AGENT)
    bb["min"] = f4(ba["min"][:, :3] + offsets(g, PAIRS, 10.0, 12.0))
    bb["max"] = f4(bb["min"][:, :3] + g.uniform(0.5, 10.0, size=(PAIRS, 3)))
    bb["result_index"] = np.arange(PAIRS) + PAIRS
    see_ctx = np.zeros(1, A.SEE_CONTEXT)
    see_ctx["radii"] = [25.0, 25.0, 25.0, 0.0]

    act = np.zeros(PAIRS, A.AGENT)
    act["p"] = f4(g.uniform(-20.0, 1020.0, size=(PAIRS, 3)))
    act["v"] = f4(g.normal(scale=5.0, size=(PAIRS, 3)))
    act["b_buffer"] = f4(g.normal(scale=50.0, size=(PAIRS, 3)))
    act["b_buffer_count"] = g.integers(0, 40, PAIRS) * (g.uniform(size=PAIRS) < 0.8)
    act["f_buffer"] = f4(g.normal(scale=5.0, size=(PAIRS, 3)))
    act["result_index"] = np.arange(PAIRS)
    bact = np.zeros(PAIRS, A.BOX_AGENT)
    bact["min"] = f4(g.uniform(-20.0, 1000.0, size=(PAIRS, 3)))
    bact["max"] = f4(bact["min"][:, :3] + g.uniform(0.5, 40.0, size=(PAIRS, 3)))
    bact["v"] = f4(g.normal(scale=5.0, size=(PAIRS, 3)))
    bact["b_buffer"] = f4(g.normal(scale=50.0, size=(PAIRS, 3)))
    bact["b_buffer_count"] = g.integers(0, 40, PAIRS) * (g.uniform(size=PAIRS) < 0.8)
    bact["f_buffer"] = f4(g.normal(scale=5.0, size=(PAIRS, 3)))
    bact["result_index"] = np.arange(PAIRS)
    act_ctx = np.zeros(1, A.ACT_CONTEXT)
    act_ctx["max"] = [1000.0, 1000.0, 1000.0, 0.0]
    return a, b, ba, bb, see_ctx, act, bact, act_ctx


def model_fixtures():
    g = W.rng(2024)
    out = {"site_spacing": np.float32(SITE_SPACING)}

    flock = load_ref("m_flocking", see=["agent_see"], act=["agent_act"])
    a, b, ctx, act, act_ctx = boid_cases(g)
    out.update(boid_see_seer=zero_w(a.copy()), boid_see_seen=zero_w(b.copy()), boid_see_ctx=ctx,
               boid_see_out=apply_see(flock.agent_see, a, b, ctx),
               boid_act_in=zero_w(act.copy()), boid_act_ctx=act_ctx,
               boid_act_out=apply_act(flock.agent_act, act, act_ctx))

    sph = load_ref("m_sph", see=["sph_particle_density", "sph_force_terms"],
                   act=["sph_particle_pressure", "sph_particle_leapfrog"], act1=["sph_particle_reset"])
    for k, ctx in enumerate(sph_contexts()):
        a, b, act = sph_cases(g, ctx)
        p = "sph%d_" % k
        out[p + "ctx"] = ctx
        out[p + "seer"] = zero_w(a.copy())
        out[p + "seen"] = zero_w(b.copy())
        out[p + "density_out"] = apply_see(sph.sph_particle_density, a, b, ctx)
        out[p + "forces_out"] = apply_see(sph.sph_force_terms, a, b, ctx)
        out[p + "act_in"] = zero_w(act.copy())
        out[p + "pressure_out"] = apply_act(sph.sph_particle_pressure, act, ctx)
        out[p + "leapfrog_out"] = apply_act(sph.sph_particle_leapfrog, act, ctx)
        out[p + "reset_out"] = apply_act1(sph.sph_particle_reset, act)

    test = load_ref("test", see=["agent_see", "box_agent_see"], act=["agent_act", "box_agent_act"])
    a, b, ba, bb, see_ctx, act, bact, act_ctx = test_model_cases(g)
    out.update(agent_see_seer=zero_w(a.copy()), agent_see_seen=zero_w(b.copy()), see_ctx=see_ctx,
               agent_see_out=apply_see(test.agent_see, a, b, see_ctx),
               box_see_seer=zero_w(ba.copy()), box_see_seen=zero_w(bb.copy()),
               box_see_out=apply_see(test.box_agent_see, ba, bb, see_ctx),
               agent_act_in=zero_w(act.copy()), act_ctx=act_ctx,
               agent_act_out=apply_act(test.agent_act, act, act_ctx),
               box_act_in=zero_w(bact.copy()),
               box_act_out=apply_act(test.box_agent_act, bact, act_ctx))
    save("ref_models.npz", out)


def save(name, out):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


def main():
    if not os.path.exists(DRIVER):
        raise SystemExit("build the reference code first: make -C oracle ref")
    simulation_fixtures()
    model_fixtures()


if __name__ == "__main__":
    main()
