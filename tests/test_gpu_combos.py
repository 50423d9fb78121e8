"""Multi-group states, mixed pairings, the reference harness's own
configuration and the fetch paths on the GPU.

* Two groups of different sizes, each seeing itself and the other
  (test/test_combo.c:64-141), in every combination of Z-grid and Simple
  spaces; two box groups; a point group with a box group (the PN and NP box
  tests between groups of different agent types and sizes), bit-exact
  against the oracle; in Simple spaces, where no sort reorders agents, also
  bit-exact against the reference's own outputs (ref_combo.npz).
* TAY_CPU_Z_GRID after tay_switch_to_ocl, the configuration of the
  reference's harness (test/test_basic.c:35-44): runs as the GPU Z-grid,
  bit-exact against the oracle's, neighbour counts equal to the reference
  CpuZGrid's.
* The 10k plain workload's shape (10,000 test agents in [0, 1000)^3,
  radius 40, cells of 80) against the reference CpuGrid: with the previous
  storage order inside cells (0 sub-cell bits) the first step meets every
  seer's pairs in the reference's order, so sums and counts are bit-exact.
* A fixed box smaller than the agents (tay_fix_space_box, which the
  reference's OpenCL grid ignores).
* Fetch paths: TAY_FETCH_ON_DEMAND keeps agents on the device until
  tay_fetch_agents or the end of the simulation; tay_fetch_agent_fields
  copies selected fields only."""
import ctypes as C
import os

import numpy as np
import pytest

from tay_amd import agents as A
from tay_amd import api
from tay_amd import workloads as W

from combo_harness import ALL_SEES, run_groups
from conftest import GOLDEN
from test_gpu_parity import assert_bitwise

pytestmark = pytest.mark.gpu

Z, S = api.TAY_GPU_Z_GRID, api.TAY_CPU_SIMPLE
ORACLE_SPACE = {api.TAY_GPU_Z_GRID: api.TAY_GPU_Z_GRID, api.TAY_CPU_SIMPLE: api.TAY_CPU_SIMPLE,
                api.TAY_GPU_SIMPLE: api.TAY_CPU_SIMPLE, api.TAY_CPU_Z_GRID: api.TAY_GPU_Z_GRID}


def against_oracle(gpu, oracle, models, groups, sees, radius, steps, **kw):
    got = run_groups(gpu, groups, sees, radius, steps, **kw)
    ogroups = [(rec, ORACLE_SPACE[space], part) for rec, space, part in groups]
    want = run_groups(oracle, ogroups, sees, radius, steps, funcs=models, **kw)
    for x, y in zip(got, want):
        assert_bitwise(x, y)
    return got


@pytest.mark.parametrize("spaces", [(Z, Z), (Z, S), (S, Z), (S, S)], ids=["zz", "zs", "sz", "ss"])
def test_two_point_groups_all_space_combinations(gpu, oracle, models, spaces):
    a = W.test_agents(3000, seed=51)
    b = W.test_agents(1700, seed=52, clump=0.2)
    print("case: two point groups %s" % (spaces,), flush=True)
    against_oracle(gpu, oracle, models, [(a, spaces[0], 50.0), (b, spaces[1], 50.0)], ALL_SEES, 50.0, 3)


def test_non_point_groups_in_simple_spaces(gpu, oracle, models):
    a = W.box_agents(1200, seed=53)
    b = W.box_agents(900, seed=54)
    print("case: two box groups in Simple spaces", flush=True)
    against_oracle(gpu, oracle, models, [(a, S, 20.0), (b, S, 20.0)], ALL_SEES, 20.0, 3)


@pytest.mark.parametrize("seen_point_space", [Z, S], ids=["points_in_zgrid", "points_in_simple"])
def test_point_group_with_box_group(gpu, oracle, models, seen_point_space):
    """Groups of different agent types and sizes (Agent, 80 bytes; BoxAgent,
    96 bytes): the point group sees the box group through the PN box test,
    the box group the point group through NP."""
    a = W.test_agents(2000, seed=56)
    b = W.box_agents(1500, seed=57)
    print("case: point group with box group", flush=True)
    got = against_oracle(gpu, oracle, models, [(a, seen_point_space, 20.0), (b, S, 20.0)], ALL_SEES, 20.0, 3)
    assert got[0]["f_buffer"][:, :3].any() and got[1]["f_buffer"][:, :3].any()


@pytest.mark.parametrize("case", ["one_group", "two_groups", "with_points"])
def test_box_agents_in_z_grid(gpu, oracle, models, case):
    """Non-point agents in the GPU Z-grid (binned by min corner, the seen
    range grown by the group's largest extent): cells of 20 against extents
    up to 50, bit-exact against the oracle's statement over three steps;
    after one see pass every agent has met exactly the pairs a Simple space
    gives it (the all-pairs set; tests/test_oracle_golden.py pins the
    statement's pairs to the reference's nonpoint loop)."""
    a = W.box_agents(3000, seed=59)
    b = W.box_agents(2000, seed=60, size=(0.5, 5.0))
    p = W.test_agents(2500, seed=61, clump=0.2)
    groups = {"one_group": [(a, Z, 20.0)], "two_groups": [(a, Z, 20.0), (b, Z, 10.0)],
              "with_points": [(p, Z, 20.0), (a, Z, 20.0)]}[case]
    sees = [(0, 0)] if case == "one_group" else ALL_SEES
    print("case: box agents in Z-grid, %s" % case, flush=True)
    got = against_oracle(gpu, oracle, models, groups, sees, 20.0, 3)
    assert all(g["f_buffer"][:, :3].any() for g in got)
    zgrid = run_groups(gpu, groups, sees, 20.0, 1, acts=False)
    simple = run_groups(gpu, [(rec, api.TAY_GPU_SIMPLE, part) for rec, _, part in groups], sees, 20.0, 1,
                        acts=False)
    for x, y in zip(zgrid, simple):
        x = x[np.argsort(x["result_index"])]
        y = y[np.argsort(y["result_index"])]
        assert np.array_equal(x["b_buffer_count"], y["b_buffer_count"])
        assert x["b_buffer_count"].sum() > len(x)


def combo_inputs(d, prefix, box=False):
    n = len(d[prefix + ("_min" if box else "_p")])
    rec = np.zeros(n, A.BOX_AGENT if box else A.AGENT)
    if box:
        rec["min"], rec["max"] = d[prefix + "_min"], d[prefix + "_max"]
    else:
        rec["p"] = d[prefix + "_p"]
    rec["v"] = d[prefix + "_v"]
    rec["result_index"] = np.arange(n)
    return rec


def check_fixture(a, d, prefix, step, point=True):
    assert np.array_equal(a["result_index"], d[prefix + "_order"][step])
    for field in (("p",) if point else ("min", "max")) + ("v", "f_buffer"):
        x = np.ascontiguousarray(a[field][:, :3]).view(np.uint32)
        y = np.ascontiguousarray(d[prefix + "_" + field][step]).view(np.uint32)
        assert np.array_equal(x, y), (prefix, field, step)


@pytest.mark.parametrize("space", [api.TAY_GPU_SIMPLE, api.TAY_CPU_SIMPLE], ids=["gpu_simple", "switched"])
def test_simple_space_combos_match_reference_outputs(gpu, space):
    """Simple spaces keep storage order and meet pairs in it, as the
    reference's CpuSimple does: every step of the reference's combo runs
    (point groups, box groups, point with box) reproduced bit for bit."""
    d = np.load(os.path.join(GOLDEN, "ref_combo.npz"))
    a, b = combo_inputs(d, "a"), combo_inputs(d, "b")
    ba, bb = combo_inputs(d, "ba", box=True), combo_inputs(d, "bb", box=True)
    r, steps = float(d["radius"]), int(d["steps"])
    cases = [([(a, space, r), (b, space, 2 * r)], r, ("pp_simple_a", True), ("pp_simple_b", True)),
             ([(ba, space, 20.0), (bb, space, 20.0)], 20.0, ("bb_a", False), ("bb_b", False)),
             ([(a, space, 20.0), (ba, space, 20.0)], 20.0, ("pb_a", True), ("pb_b", False))]
    for groups, radius, ga, gb in cases:
        print("case: reference combo %s / %s" % (ga[0], gb[0]), flush=True)
        out = run_groups(gpu, groups, ALL_SEES, radius, steps, every_step=True)
        for step, (x, y) in enumerate(out):
            check_fixture(x, d, ga[0], step, ga[1])
            check_fixture(y, d, gb[0], step, gb[1])


def test_non_point_agents_match_reference_outputs(gpu):
    """Box agents in a Simple space (test/test_nonpoint.c) against the
    reference's own outputs, every step."""
    d = np.load(os.path.join(GOLDEN, "ref_nonpoint.npz"))
    rec = np.zeros(len(d["min"]), A.BOX_AGENT)
    rec["min"], rec["max"], rec["v"] = d["min"], d["max"], d["v"]
    rec["result_index"] = np.arange(len(rec))
    print("case: reference non-point run", flush=True)
    out = run_groups(gpu, [(rec, api.TAY_GPU_SIMPLE, float(d["radius"]))], [(0, 0)], float(d["radius"]),
                     int(d["steps"]), every_step=True)
    for step, (x,) in enumerate(out):
        check_fixture(x, d, "simple", step, point=False)


def test_cpu_z_grid_after_switch_to_ocl(gpu, oracle, models):
    """The reference harness configures TAY_CPU_Z_GRID and switches to the
    device (test/test_basic.c:35-44; tay/state_compile.c:94-107 then runs it
    on the device): here it is the GPU Z-grid, bit-exact with the oracle's,
    and its pair set (per-agent counts after one see pass) is the reference
    CpuZGrid's."""
    d = np.load(os.path.join(GOLDEN, "ref_test_model_uniform.npz"))
    rec = np.zeros(len(d["p"]), A.AGENT)
    rec["p"], rec["v"] = d["p"], d["v"]
    rec["result_index"] = np.arange(len(rec))
    r, part = float(d["radius"]), float(d["part"])
    print("case: CPU Z-grid after switch", flush=True)
    against_oracle(gpu, oracle, models, [(rec, api.TAY_CPU_Z_GRID, part)], [(0, 0)], r, 3)
    model = W.test_model(radius=r, part=part)
    model.passes = model.passes[:1]

    def switched_run(lib):
        state = lib.tay_create_state()
        assert lib.tay_switch_to_ocl(state)
        for t in A.source_texts():
            lib.ocl_add_source(state, t.encode())
        g = lib.tay_add_group(state, A.AGENT.itemsize, len(rec), 1)
        lib.tay_configure_space(state, g, api.TAY_CPU_Z_GRID, 3, api.float4(part, part, part, part), 250)
        ctx = model.passes[0].context
        lib.tay_add_see(state, g, g, None, b"agent_see", api.float4(r, r, r, r), 0, ctx.ctypes.data, ctx.itemsize)
        lib.add_agents(state, g, rec)
        lib.tay_simulation_start(state)
        assert lib.tay_run(state, 1) == 1
        out = lib.get_agents(state, g, len(rec), A.AGENT)
        lib.tay_simulation_end(state)
        lib.tay_destroy_state(state)
        return out

    a = switched_run(gpu)
    a = a[np.argsort(a["result_index"])]
    want = np.zeros(len(rec), np.int32)
    want[d["zgrid_order"][0]] = d["zgrid_b_count"][0]
    assert np.array_equal(a["b_buffer_count"], want)


def test_ten_thousand_agents_match_reference_cpu_grid(gpu):
    """The 10k plain workload's shape on the GPU Z-grid with 0 sub-cell bits:
    same geometry as the reference CpuGrid, and after one step every seer
    has met its pairs in the reference's sweep order (cells x, y, z, each in
    previous storage order), so the see sums, counts, f_buffer, positions and
    velocities equal the reference's bit for bit (matched by identity: the
    two grids lay cells out in different orders)."""
    d = np.load(os.path.join(GOLDEN, "ref_test_model_c1.npz"))
    rec = np.zeros(len(d["p"]), A.AGENT)
    rec["p"], rec["v"] = d["p"], d["v"]
    rec["result_index"] = np.arange(len(rec))
    r, part = float(d["radius"]), float(d["part"])
    model = W.test_model(radius=r, part=part)
    print("case: 10k test agents vs reference CpuGrid", flush=True)
    for see_only in (True, False):
        m = W.Model(model.dtype, model.passes[:1] if see_only else model.passes, part)
        with W.Simulation(gpu, m, rec, api.TAY_GPU_Z_GRID) as s:
            assert gpu.tay_set_order_bits(s.state, s.group, 0)
            s.run(1)
            a = s.agents()
            origin, size, count, _ = gpu.grid_info(s.state, s.group)
        assert np.array_equal(origin[:3].view(np.uint32), d["grid_origin"][0].view(np.uint32))
        assert np.array_equal(size[:3].view(np.uint32), d["grid_cell_size"][0].view(np.uint32))
        assert np.array_equal(count[:3], d["grid_cell_count"][0])
        a = a[np.argsort(a["result_index"])]
        order = d["grid_order"][0]
        if see_only:
            want_count = np.zeros(len(a), np.int32)
            want_count[order] = d["grid_b_count"][0]
            want_sum = np.zeros((len(a), 3), np.float32)
            want_sum[order] = d["grid_b_buffer"][0]
            assert np.array_equal(a["b_buffer_count"], want_count)
            assert np.array_equal(a["b_buffer"][:, :3].view(np.uint32), want_sum.view(np.uint32))
        else:
            for field in ("p", "v", "f_buffer"):
                want = np.zeros((len(a), 3), np.float32)
                want[order] = d["grid_" + field][0]
                assert np.array_equal(np.ascontiguousarray(a[field][:, :3]).view(np.uint32), want.view(np.uint32))


def test_fixed_box_smaller_than_the_agents(gpu, oracle, models):
    a = W.test_agents(5000, seed=55)
    print("case: fixed box", flush=True)
    against_oracle(gpu, oracle, models, [(a, Z, 50.0)], [(0, 0)], 50.0, 3,
                   fixed={0: ((200.0, 200.0, 200.0, 0.0), (800.0, 800.0, 800.0, 0.0))})


def test_fetch_on_demand_and_selected_fields(gpu, oracle, models):
    """TAY_FETCH_ON_DEMAND: host storage keeps the pushed agents through the
    runs; tay_fetch_agent_fields copies the positions (and then the
    velocities) of the current storage order without touching host
    storage; tay_fetch_agents brings every field back, equal to the
    oracle's; agents of a last run that was never fetched reach host
    storage when the simulation ends."""
    recs = W.flock_agents(20000, seed=58, clusters=8, sigma=30.0, box=600.0)
    model = W.flock_model(25.0)
    print("case: fetch on demand", flush=True)
    with W.Simulation(oracle, model, recs, api.TAY_GPU_Z_GRID, funcs=models) as o:
        o.run(3)
        want3 = o.agents()
        o.run(2)
        want5 = o.agents()
    s = W.Simulation(gpu, model, recs, api.TAY_GPU_Z_GRID)
    try:
        gpu.tay_set_fetch_mode(s.state, api.TAY_FETCH_ON_DEMAND)
        s.run(3)
        stale = s.agents()
        assert np.array_equal(stale["p"], recs["p"])      # not fetched yet
        pos = np.zeros((len(recs), 3), np.float32)
        assert gpu.tay_fetch_agent_fields(s.state, s.group, 8, 12, pos.ctypes.data) == len(recs)
        assert np.array_equal(pos.view(np.uint32), np.ascontiguousarray(want3["p"][:, :3]).view(np.uint32))
        vel = np.zeros((len(recs), 4), np.float32)
        assert gpu.tay_fetch_agent_fields(s.state, s.group, 24, 16, vel.ctypes.data) == len(recs)
        assert np.array_equal(vel.view(np.uint32), want3["v"].view(np.uint32))
        assert np.array_equal(s.agents()["p"], recs["p"])  # host storage untouched
        assert gpu.tay_fetch_agent_fields(s.state, s.group, 6, 4, vel.ctypes.data) == 0
        assert gpu.tay_fetch_agent_fields(s.state, s.group, 96, 8, vel.ctypes.data) == 0
        assert gpu.tay_fetch_agents(s.state, s.group)
        assert_bitwise(s.agents(), want3)
        s.run(2)
        gpu.tay_simulation_end(s.state)
        assert_bitwise(gpu.get_agents(s.state, s.group, len(recs), model.dtype), want5)
    finally:
        gpu.tay_destroy_state(s.state)
        s.state = None
