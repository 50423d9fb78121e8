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

SCENARIOS = ["uniform", "clump"]
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
