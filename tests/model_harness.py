"""Per-pair and per-agent harnesses for model functions.

Host side: the models of tay_amd/agent_src built as C callbacks
(oracle/_build/libmodels.so), called once per pair or agent.

GPU side: a state of two groups through tay.h.  Pair k's seer and seen agent
sit at site k of a lattice whose spacing is far beyond the see radius, the
seer group in a Simple space (its storage order is kept, so outputs line up
with inputs) and the seen group in a Z-grid or Simple space; one see pass,
one step: every seer meets exactly its own partner.  Act functions run on a
Simple group with only the act pass."""
import ctypes as C

import numpy as np

from tay_amd import agents as A
from tay_amd import api
from tay_amd.api import float4

from conftest import MODELS_LIB

TAG = ("part_i", "cell_agent_i")


def host_models():
    return C.CDLL(MODELS_LIB)


def _fn(lib, name, nargs):
    f = getattr(lib, name)
    f.argtypes = [C.c_void_p] * nargs
    f.restype = None
    return f


def host_see(lib, name, seer, seen, ctx):
    out = np.ascontiguousarray(seer).copy()
    seen = np.ascontiguousarray(seen)
    f = _fn(lib, name, 3)
    for k in range(len(out)):
        f(out[k:k + 1].ctypes.data, seen[k:k + 1].ctypes.data, ctx.ctypes.data)
    return out


def host_act(lib, name, agents, ctx=None):
    out = np.ascontiguousarray(agents).copy()
    f = _fn(lib, name, 2 if ctx is not None else 1)
    for k in range(len(out)):
        if ctx is not None:
            f(out[k:k + 1].ctypes.data, ctx.ctypes.data)
        else:
            f(out[k:k + 1].ctypes.data)
    return out


def mismatches(a, b, skip=TAG):
    """Fields (with the number of records) where a and b differ bit for bit;
    the tag and the unused fourth lane of float4 fields are not compared."""
    bad = []
    for f in a.dtype.names:
        if f in skip:
            continue
        x, y = a[f], b[f]
        if x.ndim == 2:
            x, y = x[:, :3], y[:, :3]
        if x.dtype.kind == "f":
            x, y = np.ascontiguousarray(x).view(np.uint32), np.ascontiguousarray(y).view(np.uint32)
        n = int((x != y).reshape(len(a), -1).any(axis=1).sum())
        if n:
            bad.append((f, n))
    return bad


def _state(lib, sources):
    state = lib.tay_create_state()
    for t in (A.source_texts() if sources is None else sources):
        lib.ocl_add_source(state, t.encode())
    return state


def _finish(lib, state, groups):
    lib.tay_simulation_start(state)
    done = lib.tay_run(state, 1)
    assert done == 1, (lib.tay_get_error(state), lib.error_message(state))
    out = [lib.get_agents(state, g, n, dt) for g, n, dt in groups]
    lib.tay_simulation_end(state)
    lib.tay_destroy_state(state)
    return out


def gpu_see_pairs(lib, name, seer, seen, ctx, radius, seen_space=api.TAY_GPU_Z_GRID, seer_point=True,
                  seen_point=True, sources=None):
    """The seer records after one see pass in which seer k meets only seen
    agent k (pair sites far apart; per-axis box radius `radius`)."""
    state = _state(lib, sources)
    n = len(seer)
    ga = lib.tay_add_group(state, seer.dtype.itemsize, n, 1 if seer_point else 0)
    gb = lib.tay_add_group(state, seen.dtype.itemsize, n, 1 if seen_point else 0)
    lib.tay_configure_space(state, ga, api.TAY_GPU_SIMPLE, 3, float4(radius, radius, radius, radius), 0)
    lib.tay_configure_space(state, gb, seen_space, 3, float4(radius, radius, radius, radius), 64)
    lib.tay_add_see(state, ga, gb, None, name.encode(), float4(radius, radius, radius, radius), 0,
                    ctx.ctypes.data, ctx.itemsize)
    lib.add_agents(state, ga, seer)
    lib.add_agents(state, gb, seen)
    return _finish(lib, state, [(ga, n, seer.dtype)])[0]


def gpu_act(lib, name, agents, ctx, sources=None):
    """The records after one act pass on a Simple group."""
    state = _state(lib, sources)
    n = len(agents)
    g = lib.tay_add_group(state, agents.dtype.itemsize, n, 1)
    lib.tay_configure_space(state, g, api.TAY_GPU_SIMPLE, 3, float4(1, 1, 1, 1), 0)
    lib.tay_add_act(state, g, None, name.encode(), ctx.ctypes.data if ctx is not None else None,
                    ctx.itemsize if ctx is not None else 0)
    lib.add_agents(state, g, agents)
    return _finish(lib, state, [(g, n, agents.dtype)])[0]


def site_pairs(dtype, seed, radius, box, count=1000, spacing=200.0, pos="p"):
    """count seer/seen pairs of `dtype` at the sites of a cubic lattice of
    `spacing` (far beyond the radius): seer k at site k (jittered), seen k
    at an offset inside the per-axis box [-box, box]^3, half of them inside
    the radius; velocities normal where the record has one."""
    from tay_amd import workloads as W
    g = W.rng(seed)
    side = int(round(count ** (1.0 / 3.0)))
    assert side ** 3 == count
    i = np.arange(count)
    s = (np.stack([i % side, (i // side) % side, i // (side * side)], 1) - side // 2) * spacing
    a = np.zeros(count, dtype)
    b = np.zeros(count, dtype)
    a[pos][:, :3] = s + g.uniform(-5, 5, size=(count, 3))
    off = np.where(g.uniform(size=(count, 1)) < 0.5, g.uniform(-box, box, size=(count, 3)),
                   g.normal(size=(count, 3)) * radius / 3.0)
    b[pos][:, :3] = (a[pos][:, :3] + off).astype(np.float32)
    if "v" in dtype.names:
        a["v"][:, :3] = g.normal(size=(count, 3))
        b["v"][:, :3] = g.normal(size=(count, 3))
    return a, b, g
