"""Split see passes (tay_set_see_split) on the GPU against the oracle's
statement of the same order.

A split pass gives each seer K = 2, 4 or 8 lanes of the see kernel's
wavefront: the candidate at position pos of the lane group's sweep goes to
the seer's partial pos mod K, and at the end of the tile the partials merge
in a fixed tree, ((0 + 1) + (2 + 3)) for K = 4.  The oracle states that order independently
(oracle/oracle.c, gpu_z_grid_split_sweep), so every field of every agent
must match bit for bit.  Against the reference order the pairs are the
same: neighbour counts are equal and float sums differ only in rounding
(within 1e-5 of their magnitude).
"""
import ctypes as C

import numpy as np
import pytest

from tay_amd import api
from tay_amd import workloads as W

from combo_harness import ALL_SEES, SEE, is_point, run_groups
from test_gpu_parity import assert_bitwise

pytestmark = pytest.mark.gpu


def run(lib, funcs, model, recs, steps, lanes, counting=False, **kw):
    with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID, funcs=funcs, **kw) as s:
        for k, p in enumerate(model.passes):
            if p.kind == "see" and lanes:
                s.see_split(k, lanes)
        if counting:
            lib.tay_set_pair_counting(s.state, 1)
        s.run(steps)
        pairs = [lib.pair_counts(s.state, i)[1] for i, p in enumerate(model.passes)
                 if p.kind == "see" and (counting or lib.has("oracle_get_pair_counts"))]
        return s.agents(), pairs


def check(gpu, oracle, models, model, recs, steps, lanes):
    print("case: split %d %s n=%d steps=%d" % (lanes, model.passes[0].func, len(recs), steps), flush=True)
    ga, gp = run(gpu, None, model, recs, steps, lanes, counting=True)
    oa, op = run(oracle, models, model, recs, steps, lanes)
    assert_bitwise(ga, oa)
    assert gp == op
    return ga


@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_flocking_clustered(gpu, oracle, models, lanes):
    recs = W.flock_agents(30000, seed=5, clusters=8, sigma=30.0, box=600.0)
    check(gpu, oracle, models, W.flock_model(25.0), recs, 3, lanes)


@pytest.mark.parametrize("lanes", [2, 4])
def test_plain_trajectory(gpu, oracle, models, lanes):
    recs = W.plain_agents(10000, seed=12)
    check(gpu, oracle, models, W.plain_model(40.0), recs, 10, lanes)


@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_reference_test_model_clumped(gpu, oracle, models, lanes):
    """Clumped agents: lane groups that halve down to single seers beside
    groups that sweep together."""
    recs = W.test_agents(12000, seed=31, clump=0.2)
    check(gpu, oracle, models, W.test_model(40.0), recs, 2, lanes)


@pytest.mark.parametrize("lanes", [2, 4])
def test_sph_two_see_passes(gpu, oracle, models, lanes):
    """Density (self-see: every particle meets itself, in one of its
    partials) and forces (self skipped through the tag's cell), both split."""
    recs, lo, hi = W.sph_particles(8000, h=1.0)
    ctx = W.sph_context(h=1.0, dt=1e-3, box_min=lo, box_max=hi)
    check(gpu, oracle, models, W.sph_model(ctx, 1.0), recs, 3, lanes)


@pytest.mark.parametrize("n", [1, 2, 3, 17, 63, 65])
def test_ragged_sizes(gpu, oracle, models, n):
    """Partial tiles: seers whose lanes lie past the last seer."""
    recs = W.plain_agents(n, seed=100 + n, box=60.0)
    check(gpu, oracle, models, W.plain_model(40.0, box=60.0), recs, 2, 4)


def test_dense_cells(gpu, oracle, models):
    recs = W.plain_agents(6000, seed=14)
    recs["p"][:, :3] = (recs["p"][:, :3] * 0.05).astype(np.float32)
    check(gpu, oracle, models, W.plain_model(40.0), recs, 2, 2)


@pytest.mark.parametrize("dims", [1, 2])
def test_lower_dimensional_grids(gpu, oracle, models, dims):
    m = W.test_model(30.0)
    model = W.Model(m.dtype, m.passes, m.part, dims)
    recs = W.test_agents(6000 if dims == 2 else 3000, seed=70 + dims, clump=0.1)
    if dims == 1:
        recs["p"][:, 1:3] = 0.0
        recs["v"][:, 1:3] = 0.0
    check(gpu, oracle, models, model, recs, 3, 4)


@pytest.mark.parametrize("case", ["two_point_groups", "box_groups", "with_points"])
def test_groups(gpu, oracle, models, case):
    """Multi-group passes (seer and seen group different, point and box
    agents), every see pass split over four lanes."""
    a = W.test_agents(3000, seed=51)
    b = W.test_agents(1700, seed=52, clump=0.2)
    x = W.box_agents(2500, seed=59)
    y = W.box_agents(1500, seed=60, size=(0.5, 5.0))
    Z = api.TAY_GPU_Z_GRID
    groups = {"two_point_groups": [(a, Z, 50.0), (b, Z, 50.0)], "box_groups": [(x, Z, 20.0), (y, Z, 10.0)],
              "with_points": [(a, Z, 20.0), (x, Z, 20.0)]}[case]
    radius = 50.0 if case == "two_point_groups" else 20.0
    print("case: split groups %s" % case, flush=True)

    def setup(funcs):
        def go(lib, state, handles):
            for k, (s, t) in enumerate(ALL_SEES):
                clear, combine = W.SPLIT_FUNCS[SEE[(is_point(groups[s][0]), is_point(groups[t][0]))]]
                fc = C.cast(getattr(funcs, clear), C.c_void_p) if funcs is not None else None
                fm = C.cast(getattr(funcs, combine), C.c_void_p) if funcs is not None else None
                assert lib.tay_set_see_split(state, k, 4, fc, clear.encode(), fm, combine.encode())
        return go

    got = run_groups(gpu, groups, ALL_SEES, radius, 3, setup=setup(None))
    want = run_groups(oracle, groups, ALL_SEES, radius, 3, funcs=models, setup=setup(models))
    for g, w in zip(got, want):
        assert_bitwise(g, w)


@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_same_pairs_as_the_reference_order(gpu, lanes):
    """One see pass of the flocking model split and not split on the GPU:
    every neighbour count is equal and every float sum agrees within 1e-5 of
    the sum of its terms' magnitudes."""
    recs = W.flock_agents(60000, seed=8, clusters=16, sigma=30.0, box=1000.0)
    model = W.flock_model(25.0)
    see_only = W.Model(model.dtype, model.passes[:1], model.part, model.dims)
    out = []
    for k in (0, lanes):
        a, _ = run(gpu, None, see_only, recs, 1, k)
        out.append(a[np.argsort(a["result_index"])])
    ref, spl = out
    for f in ("cohesion_count", "separation_count"):
        assert np.array_equal(ref[f], spl[f]), f
    n = ref["cohesion_count"].astype(np.float64)
    scale = {"alignment": n, "cohesion": n * 25.0, "separation": ref["separation_count"] * 25.0}
    for f, mag in scale.items():
        d = np.abs(ref[f][:, :3].astype(np.float64) - spl[f][:, :3].astype(np.float64)).max(axis=1)
        assert np.all(d <= 1e-5 * np.maximum(mag, 1.0)), (f, d.max())
    assert (ref["cohesion_count"] > 0).mean() > 0.5
    assert not np.array_equal(ref["cohesion"], spl["cohesion"])   # the order did change


def test_split_switched_off_between_runs(gpu, oracle, models):
    """Split on for one run, off for the next (the program is compiled
    again each time), both sides alike."""
    recs = W.flock_agents(20000, seed=9, clusters=4, sigma=30.0, box=400.0)
    model = W.flock_model(25.0)
    out = []
    for lib, funcs in ((gpu, None), (oracle, models)):
        with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID, funcs=funcs) as s:
            s.see_split(0, 2)
            s.run(1)
            s.see_split(0, 1)
            s.run(1)
            s.see_split(0, 8)
            s.run(1)
            out.append(s.agents())
    assert_bitwise(*out)


def test_bad_split_is_refused(gpu):
    recs = W.plain_agents(100, seed=45)
    with W.Simulation(gpu, W.plain_model(40.0), recs, api.TAY_GPU_Z_GRID) as s:
        assert not gpu.tay_set_see_split(s.state, 1, 2, None, b"plain_see_clear", None, b"plain_see_combine")
        assert not gpu.tay_set_see_split(s.state, 0, 3, None, b"plain_see_clear", None, b"plain_see_combine")
        assert not gpu.tay_set_see_split(s.state, 0, 2, None, b"", None, b"")
        assert gpu.tay_set_see_split(s.state, 0, 4, None, b"plain_see_clear", None, b"plain_see_combine")
        s.run(1)
