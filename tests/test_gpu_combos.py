"""Multi-group states on the GPU, as the reference's combo harness builds
them (test/test_combo.c:64-141): two groups of different sizes, each seeing
itself and the other, in every combination of Z-grid and Simple spaces;
non-point (box) agents in Simple spaces; a Z-grid with a fixed box smaller
than the agents' extent (tay_fix_space_box, which the reference's OpenCL
grid ignores).  Every case is bit-exact against the oracle."""
import ctypes as C

import numpy as np
import pytest

from tay_amd import agents as A
from tay_amd import api
from tay_amd import workloads as W
from tay_amd.api import float4

from test_gpu_parity import assert_bitwise

pytestmark = pytest.mark.gpu

Z, S = api.TAY_GPU_Z_GRID, api.TAY_CPU_SIMPLE


def box_agents(n, seed, box=1000.0, size=(1.0, 50.0)):
    g = W.rng(seed)
    rec = np.zeros(n, A.BOX_AGENT)
    lo = g.uniform(0.0, box, size=(n, 3)).astype(np.float32)
    ext = g.uniform(size[0], size[1], size=(n, 3)).astype(np.float32)
    rec["min"][:, :3] = lo
    rec["max"][:, :3] = lo + ext
    rec["v"][:, :3] = W.unit_vectors(n, g)
    rec["result_index"] = np.arange(n)
    return rec


def run_state(lib, groups, see, radius, steps, funcs=None, fixed=None):
    """groups: list of (records, space); see: list of (seer, seen) group
    indices; every group also gets the model's act pass."""
    point = groups[0][0].dtype == A.AGENT
    see_name, act_name = ("agent_see", "agent_act") if point else ("box_agent_see", "box_agent_act")
    state = lib.tay_create_state()
    if space_needs_switch(groups):
        lib.tay_switch_to_ocl(state)
    for t in A.source_texts():
        lib.ocl_add_source(state, t.encode())
    act = np.zeros(1, A.ACT_CONTEXT)
    act["max"] = [1000.0, 1000.0, 1000.0, 0.0]
    see_ctx = np.zeros(1, A.SEE_CONTEXT)
    see_ctx["radii"] = [radius, radius, radius, 0.0]
    gs = []
    for k, (rec, space) in enumerate(groups):
        g = lib.tay_add_group(state, rec.dtype.itemsize, len(rec), 1 if point else 0)
        lib.tay_configure_space(state, g, space, 3, float4(radius, radius, radius, radius), 250)
        if fixed is not None and fixed[k] is not None:
            lo, hi = fixed[k]
            lib.tay_fix_space_box(state, g, float4(*lo), float4(*hi))
        gs.append(g)
    fs = C.cast(getattr(funcs, see_name), C.c_void_p) if funcs is not None else None
    fa = C.cast(getattr(funcs, act_name), C.c_void_p) if funcs is not None else None
    for a, b in see:
        lib.tay_add_see(state, gs[a], gs[b], fs, see_name.encode(), float4(radius, radius, radius, radius), 0,
                        see_ctx.ctypes.data, see_ctx.itemsize)
    for g in gs:
        lib.tay_add_act(state, g, fa, act_name.encode(), act.ctypes.data, act.itemsize)
    for g, (rec, _) in zip(gs, groups):
        lib.add_agents(state, g, rec)
    lib.tay_simulation_start(state)
    done = lib.tay_run(state, steps)
    assert done == steps, (lib.tay_get_error(state), lib.error_message(state))
    out = [lib.get_agents(state, g, len(rec), rec.dtype) for g, (rec, _) in zip(gs, groups)]
    lib.tay_simulation_end(state)
    lib.tay_destroy_state(state)
    return out


def space_needs_switch(groups):
    return any(space in (api.TAY_CPU_SIMPLE, api.TAY_CPU_Z_GRID) for _, space in groups)


ALL_SEES = [(0, 0), (0, 1), (1, 1), (1, 0)]


@pytest.mark.parametrize("spaces", [(Z, Z), (Z, S), (S, Z), (S, S)], ids=["zz", "zs", "sz", "ss"])
def test_two_point_groups_all_space_combinations(gpu, oracle, models, spaces):
    a = W.test_agents(3000, seed=51)
    b = W.test_agents(1700, seed=52, clump=0.2)
    groups = [(a, spaces[0]), (b, spaces[1])]
    print("case: two point groups %s" % (spaces,), flush=True)
    got = run_state(gpu, groups, ALL_SEES, 50.0, 3)
    want = run_state(oracle, groups, ALL_SEES, 50.0, 3, funcs=models)
    for x, y in zip(got, want):
        assert_bitwise(x, y)


def test_non_point_groups_in_simple_spaces(gpu, oracle, models):
    a = box_agents(1200, seed=53)
    b = box_agents(900, seed=54)
    groups = [(a, S), (b, S)]
    print("case: two box groups in Simple spaces", flush=True)
    got = run_state(gpu, groups, ALL_SEES, 20.0, 3)
    want = run_state(oracle, groups, ALL_SEES, 20.0, 3, funcs=models)
    for x, y in zip(got, want):
        assert_bitwise(x, y)


def test_fixed_box_smaller_than_the_agents(gpu, oracle, models):
    a = W.test_agents(5000, seed=55)
    groups = [(a, Z)]
    fixed = [((200.0, 200.0, 200.0, 0.0), (800.0, 800.0, 800.0, 0.0))]
    print("case: fixed box", flush=True)
    got = run_state(gpu, groups, [(0, 0)], 50.0, 3, fixed=fixed)
    want = run_state(oracle, groups, [(0, 0)], 50.0, 3, funcs=models, fixed=fixed)
    assert_bitwise(got[0], want[0])
