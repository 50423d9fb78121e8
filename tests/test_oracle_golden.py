"""The CPU oracle against fixtures written by the reference's own compiled
CpuSimple / CpuGrid / CpuZGrid (tests/golden/make_golden.py).

Same inputs, same model (test/agent.c semantics), same pair order: storage
order, tags, see sums and counts, positions, velocities, f_buffer and the
grid geometry must match bit for bit at every step."""
import os

import numpy as np
import pytest

from tay_amd import agents as A
from tay_amd import api
from tay_amd import workloads as W

from conftest import GOLDEN

SCENARIOS = ["uniform", "clump", "c1"]
STRUCTURES = {"simple": api.TAY_CPU_SIMPLE, "grid": api.TAY_CPU_GRID, "zgrid": api.TAY_CPU_Z_GRID}


def load(name):
    return np.load(os.path.join(GOLDEN, "ref_test_model_%s.npz" % name))


def inputs(d):
    n = len(d["p"])
    rec = np.zeros(n, A.AGENT)
    rec["p"] = d["p"]
    rec["v"] = d["v"]
    rec["result_index"] = np.arange(n)
    return rec


def bits(x):
    return np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("scenario", SCENARIOS)
@pytest.mark.parametrize("structure", list(STRUCTURES))
def test_steps_match_reference(oracle, models, scenario, structure):
    d = load(scenario)
    model = W.test_model(radius=float(d["radius"]), part=float(d["part"]))
    with W.Simulation(oracle, model, inputs(d), STRUCTURES[structure], funcs=models) as sim:
        for step in range(int(d["steps"])):
            sim.run(1)
            a = sim.agents()
            assert np.array_equal(a["result_index"], d[structure + "_order"][step])
            assert np.array_equal(a["part_i"], d[structure + "_part_i"][step])
            assert np.array_equal(a["cell_agent_i"], d[structure + "_cell_agent_i"][step])
            for field in ("p", "v", "f_buffer"):
                assert np.array_equal(bits(a[field][:, :3]), bits(d[structure + "_" + field][step])), field
            if structure != "simple":
                origin, size, count, _ = oracle.grid_info(sim.state, sim.group)
                assert np.array_equal(bits(origin[:3]), bits(d[structure + "_origin"][step]))
                assert np.array_equal(bits(size[:3]), bits(d[structure + "_cell_size"][step]))
                assert np.array_equal(count[:3], d[structure + "_cell_count"][step])


@pytest.mark.parametrize("scenario", SCENARIOS)
@pytest.mark.parametrize("structure", list(STRUCTURES))
def test_see_pass_matches_reference(oracle, models, scenario, structure):
    """A state with only the see pass exposes the sums and counts of step 1."""
    d = load(scenario)
    model = W.test_model(radius=float(d["radius"]), part=float(d["part"]))
    model.passes = model.passes[:1]
    with W.Simulation(oracle, model, inputs(d), STRUCTURES[structure], funcs=models) as sim:
        sim.run(1)
        a = sim.agents()
        assert np.array_equal(a["b_buffer_count"], d[structure + "_b_count"][0])
        assert np.array_equal(bits(a["b_buffer"][:, :3]), bits(d[structure + "_b_buffer"][0]))


@pytest.mark.parametrize("scenario", SCENARIOS)
def test_gpu_z_grid_statement_against_reference(oracle, models, scenario):
    """The oracle's statement of the GPU Z-grid finds the reference's pair set
    (per-agent counts equal CpuSimple's) and the reference CpuGrid's geometry
    (same formula, no memory cap at these sizes); its sums differ from the
    reference's only by summation order."""
    d = load(scenario)
    model = W.test_model(radius=float(d["radius"]), part=float(d["part"]))
    see_only = W.test_model(radius=float(d["radius"]), part=float(d["part"]))
    see_only.passes = see_only.passes[:1]
    with W.Simulation(oracle, see_only, inputs(d), api.TAY_GPU_Z_GRID, funcs=models) as sim:
        sim.run(1)
        a = sim.agents()
        origin, size, count, _ = oracle.grid_info(sim.state, sim.group)
    assert np.array_equal(bits(origin[:3]), bits(d["grid_origin"][0]))
    assert np.array_equal(bits(size[:3]), bits(d["grid_cell_size"][0]))
    assert np.array_equal(count[:3], d["grid_cell_count"][0])
    a = a[np.argsort(a["result_index"])]
    ref = np.zeros(len(a), np.int32)
    ref[d["simple_order"][0]] = d["simple_b_count"][0]
    assert np.array_equal(a["b_buffer_count"], ref)

    with W.Simulation(oracle, model, inputs(d), api.TAY_GPU_Z_GRID, funcs=models) as sim:
        sim.run(int(d["steps"]))
        a = sim.agents()
    a = a[np.argsort(a["result_index"])]
    f_ref = np.zeros((len(a), 3), np.float32)
    f_ref[d["simple_order"][-1]] = d["simple_f_buffer"][-1]
    # f_buffer is a sum of per-step means of offsets bounded by the radius
    np.testing.assert_allclose(a["f_buffer"][:, :3], f_ref, rtol=0, atol=1e-3)


def check_step(a, d, prefix, step, point=True):
    """One group's records after a step against the fixture's."""
    assert np.array_equal(a["result_index"], d[prefix + "_order"][step])
    assert np.array_equal(a["part_i"], d[prefix + "_part_i"][step])
    assert np.array_equal(a["cell_agent_i"], d[prefix + "_cell_agent_i"][step])
    for field in (("p",) if point else ("min", "max")) + ("v", "f_buffer"):
        assert np.array_equal(bits(a[field][:, :3]), bits(d[prefix + "_" + field][step])), (prefix, field)


def test_non_point_agents_match_reference(oracle, models):
    """Box agents in a Simple space (test/test_nonpoint.c): the reference's
    nonpoint-nonpoint pairing loop and box_agent_* models."""
    from combo_harness import run_groups
    d = np.load(os.path.join(GOLDEN, "ref_nonpoint.npz"))
    rec = np.zeros(len(d["min"]), A.BOX_AGENT)
    rec["min"], rec["max"], rec["v"] = d["min"], d["max"], d["v"]
    rec["result_index"] = np.arange(len(rec))
    steps = run_groups(oracle, [(rec, api.TAY_CPU_SIMPLE, float(d["radius"]))], [(0, 0)], float(d["radius"]),
                       int(d["steps"]), funcs=models, every_step=True)
    for step, (a,) in enumerate(steps):
        check_step(a, d, "simple", step, point=False)


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


@pytest.mark.parametrize("structure", list(STRUCTURES))
def test_two_point_groups_match_reference(oracle, models, structure):
    """Two point groups seeing themselves and each other
    (test/test_combo.c:114-119), partition sizes r and 2r."""
    from combo_harness import ALL_SEES, run_groups
    d = np.load(os.path.join(GOLDEN, "ref_combo.npz"))
    r = float(d["radius"])
    a, b = combo_inputs(d, "a"), combo_inputs(d, "b")
    space = STRUCTURES[structure]
    steps = run_groups(oracle, [(a, space, r), (b, space, 2.0 * r)], ALL_SEES, r, int(d["steps"]), funcs=models,
                       every_step=True)
    for step, (ga, gb) in enumerate(steps):
        check_step(ga, d, "pp_%s_a" % structure, step)
        check_step(gb, d, "pp_%s_b" % structure, step)


def test_two_box_groups_and_mixed_groups_match_reference(oracle, models):
    """Two box groups, and a point group with a box group (the reference's
    point-nonpoint and nonpoint-point pairing loops), in Simple spaces."""
    from combo_harness import ALL_SEES, run_groups
    d = np.load(os.path.join(GOLDEN, "ref_combo.npz"))
    ba, bb = combo_inputs(d, "ba", box=True), combo_inputs(d, "bb", box=True)
    a = combo_inputs(d, "a")
    S = api.TAY_CPU_SIMPLE
    for groups, (pa, pb, pa_point, pb_point) in (
            ([(ba, S, 20.0), (bb, S, 20.0)], ("bb_a", "bb_b", False, False)),
            ([(a, S, 20.0), (ba, S, 20.0)], ("pb_a", "pb_b", True, False))):
        steps = run_groups(oracle, groups, ALL_SEES, 20.0, int(d["steps"]), funcs=models, every_step=True)
        for step, (ga, gb) in enumerate(steps):
            check_step(ga, d, pa, step, point=pa_point)
            check_step(gb, d, pb, step, point=pb_point)


@pytest.mark.parametrize("scenario", SCENARIOS)
def test_gpu_z_grid_statement_with_previous_order_is_reference_cpu_grid(oracle, models, scenario):
    """With 0 sub-cell bits the GPU Z-grid keeps the previous storage order
    inside a cell and sweeps cells x, y, z, as the reference CpuGrid does, on
    the CpuGrid's geometry: in the first step every seer meets its pairs in
    the reference's order, so the see sums (and then f_buffer) are the
    reference CpuGrid's bit for bit."""
    d = load(scenario)
    model = W.test_model(radius=float(d["radius"]), part=float(d["part"]))
    for see_only in (True, False):
        m = W.Model(model.dtype, model.passes[:1] if see_only else model.passes, model.part)
        with W.Simulation(oracle, m, inputs(d), api.TAY_GPU_Z_GRID, funcs=models) as sim:
            assert oracle.tay_set_order_bits(sim.state, sim.group, 0)
            sim.run(1)
            a = sim.agents()
        a = a[np.argsort(a["result_index"])]
        order = d["grid_order"][0]
        field, key = ("b_buffer", "grid_b_buffer") if see_only else ("f_buffer", "grid_f_buffer")
        want = np.zeros((len(a), 3), np.float32)
        want[order] = d[key][0]
        assert np.array_equal(bits(a[field][:, :3]), bits(want))


def test_gpu_z_grid_statement_with_box_agents_finds_reference_pairs(oracle, models):
    """Box agents binned by min corner in the GPU Z-grid statement, cells of
    20 against extents up to 50: every agent meets exactly the pairs of the
    reference's all-pairs nonpoint loop (per-agent counts equal the
    reference's), and f_buffer after three steps equals it up to summation
    order."""
    from combo_harness import run_groups
    d = np.load(os.path.join(GOLDEN, "ref_nonpoint.npz"))
    rec = np.zeros(len(d["min"]), A.BOX_AGENT)
    rec["min"], rec["max"], rec["v"] = d["min"], d["max"], d["v"]
    rec["result_index"] = np.arange(len(rec))
    r = float(d["radius"])
    assert (rec["max"][:, :3] - rec["min"][:, :3]).max() > 2 * r
    state = oracle.tay_create_state()
    g = oracle.tay_add_group(state, A.BOX_AGENT.itemsize, len(rec), 0)
    oracle.tay_configure_space(state, g, api.TAY_GPU_Z_GRID, 3, api.float4(r, r, r, r), 250)
    ctx = np.zeros(1, A.SEE_CONTEXT)
    import ctypes as C
    oracle.tay_add_see(state, g, g, C.cast(models.box_agent_see, C.c_void_p), b"box_agent_see",
                       api.float4(r, r, r, r), 0, ctx.ctypes.data, ctx.itemsize)
    oracle.add_agents(state, g, rec)
    oracle.tay_simulation_start(state)
    assert oracle.tay_run(state, 1) == 1
    a = oracle.get_agents(state, g, len(rec), A.BOX_AGENT)
    oracle.tay_simulation_end(state)
    oracle.tay_destroy_state(state)
    a = a[np.argsort(a["result_index"])]
    want = np.zeros(len(rec), np.int32)
    want[d["simple_order"][0]] = d["simple_b_count"][0]
    assert np.array_equal(a["b_buffer_count"], want) and want.sum() > 500

    (out,) = run_groups(oracle, [(rec, api.TAY_GPU_Z_GRID, r)], [(0, 0)], r, int(d["steps"]), funcs=models)
    out = out[np.argsort(out["result_index"])]
    f_ref = np.zeros((len(out), 3), np.float32)
    f_ref[d["simple_order"][-1]] = d["simple_f_buffer"][-1]
    np.testing.assert_allclose(out["f_buffer"][:, :3], f_ref, rtol=0, atol=1e-3)
