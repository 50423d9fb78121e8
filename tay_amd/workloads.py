"""Synthetic workloads and a state builder over the Tay API.

Inputs come from numpy's PCG64 generator with a fixed seed, so the GPU
library and the CPU oracle receive identical agents.  A Simulation builds one
state on any library exporting the Tay API, the way the reference harness
builds its test states (test/test_basic.c:12-85): add sources, one group,
configure the space, register passes, add agents, start.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import agents as A
from .api import float4
from .api import TayGpuBatchPass as api_batch_pass


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def unit_vectors(n, gen):
    v = gen.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v.astype(np.float32)


def uniform_points(n, lo, hi, gen):
    return gen.uniform(lo, hi, size=(n, 3)).astype(np.float32)


def clustered_points(n, clusters, sigma, box, gen):
    """n points from `clusters` isotropic Gaussians of deviation sigma whose
    centres are uniform in [3 sigma, box - 3 sigma)^3."""
    centres = gen.uniform(3 * sigma, box - 3 * sigma, size=(clusters, 3))
    which = gen.integers(0, clusters, size=n)
    return (centres[which] + gen.normal(scale=sigma, size=(n, 3))).astype(np.float32)


def _xyz0(a):
    out = np.zeros((len(a), 4), np.float32)
    out[:, :3] = a
    return out


def plain_agents(n, seed=1, box=1000.0):
    """Plain workload: positions U[0, box)^3, velocities U[-1, 1]^3."""
    g = rng(seed)
    rec = np.zeros(n, A.PLAIN_AGENT)
    rec["p"] = _xyz0(uniform_points(n, 0.0, box, g))
    rec["v"] = _xyz0(g.uniform(-1.0, 1.0, size=(n, 3)).astype(np.float32))
    rec["result_index"] = np.arange(n)
    return rec


def test_agents(n, seed=1, box=1000.0, clump=0.0):
    """The reference harness's agents (test/test.c:317-348): positions
    uniform in [0, box)^3, unit-speed random velocities; with clump > 0 that
    fraction sits in [0, 0.05 box)^3 (test/test_basic.c:57-69)."""
    g = rng(seed)
    rec = np.zeros(n, A.AGENT)
    k = int(n * clump)
    pos = np.concatenate([uniform_points(n - k, 0.0, box, g), uniform_points(k, 0.0, 0.05 * box, g)])
    rec["p"] = _xyz0(pos)
    rec["v"] = _xyz0(unit_vectors(n, g))
    rec["result_index"] = np.arange(n)
    return rec


def box_agents(n, seed=1, box=1000.0, size=(1.0, 50.0)):
    """Non-point test agents (test/agent.h BoxAgent): min corners uniform in
    [0, box)^3, extents uniform in `size` per axis, unit-random velocities
    (test/test.c make_randomized_direction_cluster_nonpoint with exponent 0)."""
    g = rng(seed)
    rec = np.zeros(n, A.BOX_AGENT)
    lo = g.uniform(0.0, box, size=(n, 3)).astype(np.float32)
    ext = g.uniform(size[0], size[1], size=(n, 3)).astype(np.float32)
    rec["min"][:, :3] = lo
    rec["max"][:, :3] = lo + ext
    rec["v"][:, :3] = unit_vectors(n, g)
    rec["result_index"] = np.arange(n)
    return rec


def flock_agents(n, seed=1, clusters=64, sigma=30.0, box=2000.0):
    """Flocking workload: Gaussian clusters, unit-random velocities."""
    g = rng(seed)
    rec = np.zeros(n, A.FLOCK_AGENT)
    rec["p"] = _xyz0(clustered_points(n, clusters, sigma, box, g))
    rec["v"] = _xyz0(unit_vectors(n, g))
    rec["result_index"] = np.arange(n)
    return rec


def smooth_flock_agents(n, seed=1, clusters=16, sigma=30.0, box=500.0):
    """Differentiable-flocking inputs: (positions, velocities) as (n, 3)
    float32 arrays, Gaussian clusters and unit-random velocities."""
    g = rng(seed)
    return clustered_points(n, clusters, sigma, box, g), unit_vectors(n, g)


def boids(n, seed=1, half=300.0):
    """m_flocking's boids (m_flocking/main.c:161-194): positions U[-half,
    half)^3, unit headings, speed 1."""
    g = rng(seed)
    rec = np.zeros(n, A.BOID)
    rec["p"] = _xyz0(uniform_points(n, -half, half, g))
    rec["dir"] = _xyz0(unit_vectors(n, g))
    rec["speed"] = 1.0
    return rec


def sph_particles(n, h=1.0, fill=0.4, seed=1, jitter=0.05):
    """Dam break: a water block filling `fill` of a box on a lattice of
    spacing 0.5 h, jittered by +-jitter h.  Returns (records, box_min,
    box_max)."""
    g = rng(seed)
    spacing = 0.5 * h
    per_axis = int(np.ceil(n ** (1.0 / 3.0)))
    side = per_axis * spacing
    box = side / fill ** (1.0 / 3.0)
    idx = np.arange(per_axis ** 3)[:n]
    lattice = np.stack([idx % per_axis, (idx // per_axis) % per_axis, idx // (per_axis * per_axis)], 1)
    pos = (lattice + 0.5) * spacing + g.uniform(-jitter * h, jitter * h, size=(n, 3))
    rec = np.zeros(n, A.SPH_PARTICLE)
    rec["p"] = _xyz0(pos.astype(np.float32))
    return rec, np.array([0, 0, 0, 0], np.float32), np.array([box, box, box, 0], np.float32)


def sph_context(h=1.0, dt=1e-3, box_min=None, box_max=None, m=0.05, rest_density=998.29):
    """Kernel constants of Mueller, Charypar and Gross (2003), computed as in
    m_sph/main.c:172-187."""
    c = np.zeros(1, A.SPH_CONTEXT)
    pi = np.float32(3.14159265358979323846)
    h = np.float32(h)
    h2 = h * h
    h6 = h2 * h2 * h2
    h9 = h6 * h2 * h
    m = np.float32(m)
    c["h"] = h
    c["h2"] = h2
    c["dt"] = dt
    c["m"] = m
    c["dynamic_viscosity"] = 3.5
    c["K"] = np.float32(101325.0) / np.float32(rest_density)
    c["density"] = rest_density
    c["poly6"] = m * np.float32(315.0) / (np.float32(64.0) * pi * h9)
    c["poly6_gradient"] = m * np.float32(-945.0) / (np.float32(32.0) * pi * h9)
    c["poly6_laplacian"] = c["poly6_gradient"]
    c["spiky"] = m * np.float32(-45.0) / (pi * h6)
    c["viscosity"] = -c["spiky"]
    c["min"] = box_min if box_min is not None else 0.0
    c["max"] = box_max if box_max is not None else 0.0
    return c


@dataclass
class Pass:
    kind: str                   # "see" or "act"
    func: str
    context: np.ndarray = None  # one-element structured array, or None
    radius: float = 0.0
    self_see: int = 0
    radii: tuple = None         # per-axis see radii (x, y, z); None: radius on every axis


@dataclass
class Model:
    dtype: np.dtype
    passes: list
    part: float                 # suggested cell size
    dims: int = 3
    extra: dict = field(default_factory=dict)


def plain_model(radius=40.0, box=1000.0):
    see = np.zeros(1, A.PLAIN_SEE_CONTEXT)
    see["r2"] = np.float32(radius) * np.float32(radius)
    act = np.zeros(1, A.ACT_CONTEXT)
    act["max"] = [box, box, box, 0.0]
    return Model(A.PLAIN_AGENT, [Pass("see", "plain_see", see, radius), Pass("act", "plain_act", act)],
                 part=2.0 * radius)


def test_model(radius=50.0, box=1000.0, part=None):
    see = np.zeros(1, A.SEE_CONTEXT)
    see["radii"] = [radius, radius, radius, 0.0]
    act = np.zeros(1, A.ACT_CONTEXT)
    act["max"] = [box, box, box, 0.0]
    return Model(A.AGENT, [Pass("see", "agent_see", see, radius), Pass("act", "agent_act", act)],
                 part=2.0 * radius if part is None else part)


def flock_model(radius=25.0, weights=(1.5, 1.0, 1.0), dt=0.1, max_speed=2.0, part=None):
    see = np.zeros(1, A.FLOCK_SEE_CONTEXT)
    see["r"] = radius
    see["separation_r"] = np.float32(radius) * np.float32(0.5)
    act = np.zeros(1, A.FLOCK_ACT_CONTEXT)
    act["separation_w"], act["alignment_w"], act["cohesion_w"] = weights
    act["dt"] = dt
    act["max_speed"] = max_speed
    return Model(A.FLOCK_AGENT, [Pass("see", "flock_see", see, radius), Pass("act", "flock_act", act)],
                 part=radius if part is None else part)


def smooth_model(radius=20.0, weights=(1.0, 1.0, 1.0), h2=1.0, dt=0.1, damping=0.05, part=None):
    """The differentiable flocking model (smooth_see / smooth_act)."""
    see = np.zeros(1, A.SMOOTH_SEE_CONTEXT)
    r = np.float32(radius)
    see["r2"] = r * r
    see["inv_r2"] = np.float32(1.0) / (r * r)
    see["h2"] = h2
    act = np.zeros(1, A.SMOOTH_ACT_CONTEXT)
    act["separation_w"], act["alignment_w"], act["cohesion_w"] = weights
    act["dt"] = dt
    act["damping"] = damping
    return Model(A.SMOOTH_AGENT, [Pass("see", "smooth_see", see, radius), Pass("act", "smooth_act", act)],
                 part=radius if part is None else part)


def boid_model(radius=20.0):
    see = np.zeros(1, A.BOID_SEE_CONTEXT)
    see["r"] = radius
    see["separation_r"] = np.float32(radius) * np.float32(0.5)
    act = np.zeros(1, A.BOID_ACT_CONTEXT)
    return Model(A.BOID, [Pass("see", "boid_see", see, radius), Pass("act", "boid_act", act)], part=radius)


def sph_model(ctx, h=1.0):
    return Model(A.SPH_PARTICLE,
                 [Pass("see", "sph_density", ctx, h, self_see=1), Pass("act", "sph_pressure", ctx),
                  Pass("see", "sph_forces", ctx, h, self_see=0), Pass("act", "sph_integrate", ctx)],
                 part=h)


# Clear and combine functions of the see functions that only add to their
# seer (tay_amd/agent_src), for split see passes.
SPLIT_FUNCS = {
    "flock_see": ("flock_see_clear", "flock_see_combine"),
    "plain_see": ("plain_see_clear", "plain_see_combine"),
    "agent_see": ("agent_see_clear", "agent_see_combine"),
    "box_agent_see": ("box_agent_see_clear", "box_agent_see_combine"),
    "point_box_see": ("agent_see_clear", "agent_see_combine"),
    "box_point_see": ("box_agent_see_clear", "box_agent_see_combine"),
    "sph_density": ("sph_density_clear", "sph_density_combine"),
    "sph_forces": ("sph_forces_clear", "sph_forces_combine"),
}


class Simulation:
    """One Tay state with one group running `model` on `lib`.

    funcs is a loaded library holding the model's C callbacks (the CPU
    oracle needs them; the GPU library compiles the sources instead)."""

    def __init__(self, lib, model, records, space_type, funcs=None, shared_mb=250, capacity=None,
                 sources=None, part=None, box=None):
        self.lib = lib
        self.model = model
        self.funcs = funcs
        self.state = lib.tay_create_state()
        if not self.state:
            raise RuntimeError("tay_create_state failed")
        for text in (A.source_texts() if sources is None else sources):
            lib.ocl_add_source(self.state, text.encode())
        n = len(records)
        self.group = lib.tay_add_group(self.state, model.dtype.itemsize, capacity or max(n, 1), 1)
        p = model.part if part is None else part
        lib.tay_configure_space(self.state, self.group, space_type, model.dims, float4(p, p, p, p), shared_mb)
        if box is not None:
            lib.tay_fix_space_box(self.state, self.group, float4(*box[0]), float4(*box[1]))
        self._keep = []
        for ps in model.passes:
            ctx = ps.context
            ptr = ctx.ctypes.data if ctx is not None else None
            size = ctx.itemsize if ctx is not None else 0
            self._keep.append(ctx)
            fn = C.cast(getattr(funcs, ps.func), C.c_void_p) if funcs is not None else None
            if ps.kind == "see":
                r = ps.radius
                rad = float4(*ps.radii, r) if ps.radii is not None else float4(r, r, r, r)
                lib.tay_add_see(self.state, self.group, self.group, fn, ps.func.encode(), rad,
                                ps.self_see, ptr, size)
            else:
                lib.tay_add_act(self.state, self.group, fn, ps.func.encode(), ptr, size)
        lib.add_agents(self.state, self.group, records)
        self.count = n
        lib.tay_simulation_start(self.state)

    def run(self, steps):
        done = self.lib.tay_run(self.state, steps)
        if done != steps:
            raise RuntimeError("tay_run failed: error %d %s" % (self.lib.tay_get_error(self.state),
                                                                 self.lib.error_message(self.state)))
        return done

    def see_split(self, pass_index, lanes):
        """Splits see pass `pass_index` over `lanes` lanes per seer
        (tay_set_see_split) with the see function's clear and combine
        functions (SPLIT_FUNCS)."""
        names = SPLIT_FUNCS[self.model.passes[pass_index].func]
        fns = [C.cast(getattr(self.funcs, f), C.c_void_p) if self.funcs is not None else None for f in names]
        if not self.lib.tay_set_see_split(self.state, pass_index, lanes, fns[0], names[0].encode(), fns[1],
                                          names[1].encode()):
            raise RuntimeError("tay_set_see_split failed: %s" % self.lib.error_message(self.state))

    def agents(self):
        return self.lib.get_agents(self.state, self.group, self.count, self.model.dtype)

    def ms_per_step(self):
        return self.lib.tay_get_ms_per_step_for_last_run(self.state)

    def close(self):
        if self.state:
            self.lib.tay_simulation_end(self.state)
            self.lib.tay_destroy_state(self.state)
            self.state = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def ensemble(instances=256, seed=7, counts=(1000, 64000), boxes=(200.0, 1500.0), radii=(10.0, 40.0),
             weights=(0.5, 2.0), dt=0.1, max_speed=2.0):
    """Batched-ensemble workload: per instance an agent count log-uniform in
    `counts`, a box side uniform in `boxes`, a see radius uniform in `radii`
    and three flocking weights uniform in `weights`; agents uniform in the
    instance's box with unit-random velocities.  Returns a list of
    (records, model, box side) per instance."""
    g = rng(seed)
    n = np.exp(g.uniform(np.log(counts[0]), np.log(counts[1]), instances)).astype(np.int64)
    side = g.uniform(boxes[0], boxes[1], instances)
    r = g.uniform(radii[0], radii[1], instances).astype(np.float32)
    w = g.uniform(weights[0], weights[1], (instances, 3)).astype(np.float32)
    out = []
    for b in range(instances):
        gb = rng(seed * 100003 + b)
        rec = np.zeros(int(n[b]), A.FLOCK_AGENT)
        rec["p"] = _xyz0(uniform_points(int(n[b]), 0.0, float(side[b]), gb))
        rec["v"] = _xyz0(unit_vectors(int(n[b]), gb))
        rec["result_index"] = np.arange(int(n[b]))
        out.append((rec, flock_model(radius=float(r[b]), weights=tuple(float(x) for x in w[b]), dt=dt,
                                     max_speed=max_speed), float(side[b])))
    return out


def density_part(records, radius, target, min_factor, max_factor):
    """Cell side for about `target` agents per cell over the bounding box of
    the records' positions, clamped to [min_factor, max_factor] x radius: the
    rule tay_gpu_batch_density_part_sizes applies on the device (3D)."""
    p = np.stack([records["p"][:, k] for k in range(3)], 1).astype(np.float64)
    vol = float(np.prod(np.maximum(p.max(0) - p.min(0), 0.0))) if len(p) else 0.0
    side = (target * vol / len(p)) ** (1.0 / 3.0) if len(p) and vol > 0 else 0.0
    return min(max(side, min_factor * radius), max_factor * radius)


def cell_table_mb(side, part, headroom=2.0):
    """Whole megabytes of cell table for a cubic box of `side` at cell size
    `part`, with `headroom` per axis for the box to grow."""
    cells = (np.ceil(side * headroom / part) + 1.0) ** 3
    return int(max(1, np.ceil(cells * 8 / (1 << 20))))


class Batch:
    """A batched ensemble on the GPU library: every instance runs `model`'s
    passes (same functions, own contexts and radii) over its own records.

    part[b] is instance b's cell size and table_mb[b] its cell table in whole
    megabytes, so a state built with tay_configure_space(..., part[b],
    table_mb[b]) computes the same grid."""

    def __init__(self, lib, instances, part, table_mb, count_pairs=False, sources=None):
        self.lib = lib
        self.models = [m for _, m, _ in instances]
        self.records = [r for r, _, _ in instances]
        m0 = self.models[0]
        self.dtype = m0.dtype
        texts = [t.encode() for t in (A.source_texts() if sources is None else sources)]
        self._src = (C.c_char_p * len(texts))(*texts)
        self.passes = (api_batch_pass * len(m0.passes))(
            *[api_batch_pass(0 if ps.kind == "see" else 1, ps.func.encode(), ps.self_see,
                             ps.context.itemsize if ps.context is not None else 0) for ps in m0.passes])
        self.counts = np.array([len(r) for r in self.records], np.uint32)
        B = len(self.records)
        tables = (C.c_ulonglong * B)(*[int(mb) << 20 for mb in table_mb])
        log = C.create_string_buffer(1 << 16)
        self.batch = lib.tay_gpu_batch_create(self._src, len(texts), self.passes, len(m0.passes),
                                              m0.dtype.itemsize, m0.dims, B,
                                              self.counts.ctypes.data_as(C.POINTER(C.c_uint)), tables,
                                              1 if count_pairs else 0, log, len(log))
        if not self.batch:
            raise RuntimeError("tay_gpu_batch_create: %s\n%s" % (lib.tay_gpu_last_error().decode(), log.value.decode()))
        self.set_part(part)
        for i, ps in enumerate(m0.passes):
            if ps.kind == "see":
                radii = np.zeros((B, 4), np.float32)
                radii[:, :3] = np.array([m.passes[i].radius for m in self.models], np.float32)[:, None]
                radii[:, 3] = radii[:, 0]
                self._ok(lib.tay_gpu_batch_set_radii(self.batch, i, radii.ctypes.data_as(C.POINTER(C.c_float))))
            if ps.context is not None:
                ctx = np.concatenate([m.passes[i].context for m in self.models])
                self._ok(lib.tay_gpu_batch_set_contexts(self.batch, i, ctx.ctypes.data))
        allrec = np.concatenate(self.records)
        self._ok(lib.tay_gpu_batch_push(self.batch, allrec.ctypes.data))

    def _ok(self, r):
        if not r:
            raise RuntimeError(self.lib.tay_gpu_last_error().decode())
        return r

    def set_part(self, part):
        parts = np.zeros((len(self.records), 4), np.float32)
        parts[:] = np.asarray(part, np.float32)[:, None]
        self._ok(self.lib.tay_gpu_batch_set_part_sizes(self.batch, parts.ctypes.data_as(C.POINTER(C.c_float))))

    def run(self, steps, graph=True):
        ms = C.c_double()
        self._ok(self.lib.tay_gpu_batch_run(self.batch, steps, 1 if graph else 0, C.byref(ms)))
        return ms.value

    def agents(self):
        out = np.zeros(int(self.counts.sum()), self.dtype)
        self._ok(self.lib.tay_gpu_batch_fetch(self.batch, out.ctypes.data))
        return np.split(out, np.cumsum(self.counts)[:-1].astype(np.int64))

    def pair_count(self, pass_index):
        n = C.c_ulonglong()
        self._ok(self.lib.tay_gpu_batch_pair_count(self.batch, pass_index, C.byref(n)))
        return n.value

    def pass_time(self, pass_index):
        ms, launches = C.c_double(), C.c_uint()
        self._ok(self.lib.tay_gpu_batch_pass_time(self.batch, pass_index, C.byref(ms), C.byref(launches)))
        return ms.value, launches.value

    def close(self):
        if self.batch:
            self.lib.tay_gpu_batch_destroy(self.batch)
            self.batch = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
