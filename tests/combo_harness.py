"""Multi-group states of the reference test models, built as the reference's
harnesses build them (test/test_combo.c:64-141, test/test_nonpoint.c:28-58):
groups of point agents (test/agent.h Agent) or box agents (BoxAgent), see
passes between given pairs of groups with the model of each pairing (point
sees point: agent_see, box sees box: box_agent_see, and the mixed
point_box_see / box_point_see), then every group's act pass."""
import ctypes as C

import numpy as np

from tay_amd import agents as A
from tay_amd import api
from tay_amd.api import float4

SEE = {(True, True): "agent_see", (False, False): "box_agent_see", (True, False): "point_box_see",
       (False, True): "box_point_see"}
ACT = {True: "agent_act", False: "box_agent_act"}


def is_point(rec):
    return rec.dtype == A.AGENT


def run_groups(lib, groups, sees, radius, steps, funcs=None, fixed=None, switch=None, shared_mb=250,
               every_step=False, setup=None, acts=True):
    """groups: list of (records, space type, partition size); sees: list of
    (seer, seen) group indices.  Returns the groups' records after the run,
    or, with every_step, a list of them after each step.  switch (default:
    when a group has a CPU space type) calls tay_switch_to_ocl first; fixed
    maps a group index to its (min, max) box; setup(lib, state, group
    handles) runs before the simulation starts; acts=False leaves out the
    act passes (the see sums and counts then stay in the records)."""
    state = lib.tay_create_state()
    if switch is None:
        switch = any(space in (api.TAY_CPU_SIMPLE, api.TAY_CPU_Z_GRID) for _, space, _ in groups)
    if switch:
        lib.tay_switch_to_ocl(state)
    for t in A.source_texts():
        lib.ocl_add_source(state, t.encode())
    act = np.zeros(1, A.ACT_CONTEXT)
    act["max"] = [1000.0, 1000.0, 1000.0, 0.0]
    see_ctx = np.zeros(1, A.SEE_CONTEXT)
    see_ctx["radii"] = [radius, radius, radius, 0.0]
    handles = []
    for k, (rec, space, part) in enumerate(groups):
        g = lib.tay_add_group(state, rec.dtype.itemsize, max(len(rec), 1), 1 if is_point(rec) else 0)
        lib.tay_configure_space(state, g, space, 3, float4(part, part, part, part), shared_mb)
        if fixed is not None and k in fixed:
            lo, hi = fixed[k]
            lib.tay_fix_space_box(state, g, float4(*lo), float4(*hi))
        handles.append(g)

    def fn(name):
        return C.cast(getattr(funcs, name), C.c_void_p) if funcs is not None else None

    for a, b in sees:
        name = SEE[(is_point(groups[a][0]), is_point(groups[b][0]))]
        lib.tay_add_see(state, handles[a], handles[b], fn(name), name.encode(),
                        float4(radius, radius, radius, radius), 0, see_ctx.ctypes.data, see_ctx.itemsize)
    for g, (rec, _, _) in zip(handles, groups):
        name = ACT[is_point(rec)]
        if acts:
            lib.tay_add_act(state, g, fn(name), name.encode(), act.ctypes.data, act.itemsize)
    for g, (rec, _, _) in zip(handles, groups):
        lib.add_agents(state, g, rec)
    if setup is not None:
        setup(lib, state, handles)
    lib.tay_simulation_start(state)
    out = []
    for _ in range(steps if every_step else 1):
        n = 1 if every_step else steps
        done = lib.tay_run(state, n)
        assert done == n, (lib.tay_get_error(state), lib.error_message(state))
        out.append([lib.get_agents(state, g, len(rec), rec.dtype) for g, (rec, _, _) in zip(handles, groups)])
    lib.tay_simulation_end(state)
    lib.tay_destroy_state(state)
    return out if every_step else out[0]


ALL_SEES = [(0, 0), (0, 1), (1, 1), (1, 0)]
