"""Pins of the oracle's split see passes (oracle/oracle.c,
gpu_z_grid_split_sweep and gpu_z_grid_see_tiles), on the CPU.

In a split pass (tay_set_see_split) K lanes share each seer: tiles hold
64 / K seers, a lane group sweeps its cells in the reference sweep order and
the candidate at position pos of the group's sweep goes to the seer's
partial pos mod K (partial 0 is the seer's own record, partials 1..K-1 start
as cleared copies); in a lane group that walks per seer, the seer's cell
rows (x, y) of its own box go to partial (row index in the box) mod K; at
the end of the tile the partials merge in a tree, ((0 + 1) + (2 + 3)) for
K = 4.

* A numpy restatement of that order, written from the definition above,
  accumulates the plain model's neighbour sums in float32: bit-exact against
  the oracle for K = 2, 4 and 8.  Planted changes (partials by blocks of the
  sweep instead of by position mod K, a sequential merge, positions shifted
  by one, every pair on partial 0, walked rows all on partial 0) fail it.
* The pairs and neighbour counts are those of the reference order, the sums
  agree within rounding, and one lane per seer is the reference order.
"""
import ctypes as C

import numpy as np
import pytest

from tay_amd import api
from tay_amd import workloads as W

from test_oracle_balanced import cell_axis, cells_of_tags, see_only

F = np.float32


def run_split(oracle, models, model, recs, lanes, steps=1, walk=64):
    with W.Simulation(oracle, model, recs, api.TAY_GPU_Z_GRID, funcs=models) as s:
        assert oracle.tay_set_lane_walk(s.state, walk)
        for k, p in enumerate(model.passes):
            if p.kind == "see" and lanes:
                s.see_split(k, lanes)
        s.run(steps)
        return s.agents(), oracle.grid_info(s.state, s.group), oracle.pair_counts(s.state, 0)[1]


def split_plain_sums(agents, info, radius, K, walk=64, assign="mod", merge="tree", shift=0, walk_rows=True):
    """Neighbour sums and counts of the plain model (a pair counts when the
    squared distance is at most r^2) in the split order of K lanes."""
    origin, size, counts, _ = info
    pos = agents["p"][:, :3].astype(F)
    n = len(agents)
    cell = cells_of_tags(agents["part_i"])
    ranges = {}
    for i in range(n):
        key = tuple(cell[i])
        b, e = ranges.get(key, (i, i))
        ranges[key] = (b, i + 1)
    r = F(radius)
    r2 = F(r * r)
    limit = 1
    for k in range(3):
        limit *= int(np.ceil(F(F(2.0) * r) / F(size[k]))) + 3
    seers = 64 // K
    order = [[[] for _ in range(K)] for _ in range(n)]     # per seer and partial: candidates met

    def box_ok(i, j):
        d = pos[i] - pos[j]
        return i != j and not np.any(np.abs(d) > r)

    def box_of(members, k):
        tmin = F(pos[list(members), k].min())
        tmax = F(pos[list(members), k].max())
        mlo = F(F(abs(tmin) + r) * F(2.0 ** -20))
        mhi = F(F(abs(tmax) + r) * F(2.0 ** -20))
        return (cell_axis(F(F(tmin - r) - mlo), origin[k], size[k], counts[k]),
                cell_axis(F(F(tmax + r) + mhi), origin[k], size[k], counts[k]))

    def cells_in(lo, hi):
        return [(x, y, z) for x in range(lo[0], hi[0] + 1) for y in range(lo[1], hi[1] + 1)
                for z in range(lo[2], hi[2] + 1)]

    for i0 in range(0, n, seers):
        n_active = min(seers, n - i0)
        width = w0 = seers
        beg = 0
        while beg < n_active:
            members = range(i0 + beg, i0 + min(beg + width, n_active))
            lo, hi = zip(*[box_of(members, k) for k in range(3)])
            ncells = (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1)
            if ncells > limit and width > 1 and width > walk:
                width //= 2
                continue
            if ncells > limit and width > 1:
                for i in members:
                    olo, ohi = zip(*[box_of([i], k) for k in range(3)])
                    ny = ohi[1] - olo[1] + 1
                    for c in cells_in(olo, ohi):
                        row = (c[0] - olo[0]) * ny + (c[1] - olo[1])
                        b, e = ranges.get(c, (0, 0))
                        order[i][row % K if walk_rows else 0] += [j for j in range(b, e) if box_ok(i, j)]
                beg += width
                while width < w0 and beg % (2 * width) == 0:
                    width *= 2
                continue
            seq = []
            for c in cells_in(lo, hi):
                b, e = ranges.get(c, (0, 0))
                seq += range(b, e)
            block = max(1, -(-len(seq) // K))
            for p, j in enumerate(seq):
                part = (p + shift) % K if assign == "mod" else p // block if assign == "block" else 0
                for i in members:
                    if box_ok(i, j):
                        order[i][part].append(j)
            beg += width
            while width < w0 and beg % (2 * width) == 0:
                width *= 2
    sums = np.zeros((n, 3), F)
    cnt = np.zeros(n, np.uint32)
    for i in range(n):
        partial = []
        for js in order[i]:
            js = [j for j in js if F(F(F(F(pos[j, 0] - pos[i, 0]) ** 2) + F(F(pos[j, 1] - pos[i, 1]) ** 2))
                                     + F(F(pos[j, 2] - pos[i, 2]) ** 2)) <= r2]
            cnt[i] += len(js)
            partial.append(np.add.accumulate(pos[js], axis=0, dtype=F)[-1] if js else np.zeros(3, F))
        if merge == "tree":
            step = 1
            while step < K:
                for k in range(0, K, 2 * step):
                    partial[k] = (partial[k] + partial[k + step]).astype(F)
                step *= 2
            sums[i] = partial[0]
        else:
            acc = partial[0]
            for k in range(1, K):
                acc = (acc + partial[k]).astype(F)
            sums[i] = acc
    return sums, cnt


@pytest.mark.parametrize("K,walk", [(2, 64), (4, 64), (8, 64), (4, 0), (4, 8)])
def test_numpy_restatement_matches_oracle(oracle, models, K, walk):
    # twelve clusters: tiles inside a cluster sweep together, tiles across
    # two clusters walk per seer
    g = np.random.default_rng(7)
    recs = W.plain_agents(1500, seed=41, box=1000.0)
    centers = g.uniform(100.0, 900.0, (12, 3))
    recs["p"][:, :3] = np.clip(centers[g.integers(0, 12, 1500)] + g.normal(0.0, 25.0, (1500, 3)), 0.0,
                               999.0).astype(F)
    assert not recs["nsum"].any() and not recs["ncount"].any()
    model = see_only(W.plain_model(30.0, box=1000.0))
    agents, info, _ = run_split(oracle, models, model, recs, K, walk=walk)
    sums, cnt = split_plain_sums(agents, info, 30.0, K, walk)
    assert np.array_equal(agents["ncount"], cnt)
    assert np.array_equal(agents["nsum"][:, :3].view(np.uint32), sums.view(np.uint32))
    # the order matters: planted variants of it do not reproduce the oracle
    # (with two partials a shift or another merge only swaps the operands of
    # one commutative addition, so those plants apply to four and eight)
    variants = [{"assign": "block"}, {"assign": "one"}, {"walk_rows": False}]
    if K > 2:
        variants += [{"shift": 1}, {"merge": "sequential"}]
    for kw in variants:
        if walk == 0 and "walk_rows" in kw:
            continue        # no group walks
        s2, c2 = split_plain_sums(agents, info, 30.0, K, walk, **kw)
        assert np.array_equal(c2, cnt)
        assert not np.array_equal(s2.view(np.uint32), sums.view(np.uint32)), kw


def test_one_lane_is_the_reference_order(oracle, models):
    recs = W.flock_agents(8000, seed=44, clusters=4, sigma=30.0, box=400.0)
    model = W.flock_model(25.0)
    ref, _, rp = run_split(oracle, models, model, recs, 0, steps=2)
    one, _, op = run_split(oracle, models, model, recs, 1, steps=2)
    for f in ref.dtype.names:
        assert np.ascontiguousarray(ref[f]).tobytes() == np.ascontiguousarray(one[f]).tobytes(), f
    assert rp == op


@pytest.mark.parametrize("K", [2, 4, 8])
def test_dense_state_same_pairs_other_rounding(oracle, models, K):
    recs = W.flock_agents(20000, seed=43, clusters=4, sigma=30.0, box=400.0)
    model = see_only(W.flock_model(25.0))
    ref, _, rp = run_split(oracle, models, model, recs, 0)
    spl, _, sp = run_split(oracle, models, model, recs, K)
    assert rp == sp
    assert np.array_equal(ref["cohesion_count"], spl["cohesion_count"])
    assert np.array_equal(ref["separation_count"], spl["separation_count"])
    n = np.maximum(ref["cohesion_count"], 1).astype(np.float64)
    for f in ("alignment", "cohesion", "separation"):
        d = np.abs(ref[f][:, :3].astype(np.float64) - spl[f][:, :3]).max(axis=1)
        assert np.all(d <= 1e-5 * n * 25.0), f
    assert not np.array_equal(ref["cohesion"], spl["cohesion"])


def test_bad_split_is_refused(oracle, models):
    recs = W.plain_agents(100, seed=45)
    with W.Simulation(oracle, W.plain_model(40.0), recs, api.TAY_GPU_Z_GRID, funcs=models) as s:
        clear = C.cast(models.plain_see_clear, C.c_void_p)
        combine = C.cast(models.plain_see_combine, C.c_void_p)
        assert not oracle.tay_set_see_split(s.state, 1, 2, clear, b"plain_see_clear", combine,
                                            b"plain_see_combine")      # pass 1 acts
        assert not oracle.tay_set_see_split(s.state, 0, 3, clear, b"plain_see_clear", combine,
                                            b"plain_see_combine")
        assert not oracle.tay_set_see_split(s.state, 0, 2, None, b"", None, b"")
        assert oracle.tay_set_see_split(s.state, 0, 4, clear, b"plain_see_clear", combine, b"plain_see_combine")


def test_test_model_clear_and_combine(oracle, models):
    """agent_see, box_agent_see, point_box_see and box_point_see split over
    four lanes (their clear and combine functions, tay_amd/agent_src/
    test_agent.c) meet the pairs of the unsplit pass: equal counts, buffer
    sums within 1e-5 of count x the largest offset."""
    from combo_harness import ALL_SEES, SEE, is_point, run_groups
    a = W.test_agents(2000, seed=61, clump=0.2)
    x = W.box_agents(1500, seed=62)
    groups = [(a, api.TAY_GPU_Z_GRID, 20.0), (x, api.TAY_GPU_Z_GRID, 20.0)]

    def split(lib, state, handles):
        for k, (s, t) in enumerate(ALL_SEES):
            clear, combine = W.SPLIT_FUNCS[SEE[(is_point(groups[s][0]), is_point(groups[t][0]))]]
            assert lib.tay_set_see_split(state, k, 4, C.cast(getattr(models, clear), C.c_void_p), clear.encode(),
                                         C.cast(getattr(models, combine), C.c_void_p), combine.encode())
    ref = run_groups(oracle, groups, ALL_SEES, 20.0, 1, funcs=models, acts=False)
    spl = run_groups(oracle, groups, ALL_SEES, 20.0, 1, funcs=models, acts=False, setup=split)
    for r, s in zip(ref, spl):
        assert np.array_equal(r["b_buffer_count"], s["b_buffer_count"])
        assert r["b_buffer_count"].sum() > 1000
        bound = 1e-5 * r["b_buffer_count"].astype(np.float64) * (20.0 + 50.0) + 1e-30
        d = np.abs(r["b_buffer"][:, :3].astype(np.float64) - s["b_buffer"][:, :3]).max(axis=1)
        assert np.all(d <= bound)


def test_sph_clear_and_combine(oracle, models):
    """sph_density and sph_forces split over four lanes (sph_*_clear and
    sph_*_combine, tay_amd/agent_src/sph.c): densities, sums of positive
    terms, within 1e-5 relative of the unsplit pass; the accumulators within
    1e-4 of the largest accumulator (a partial left out of the merge would
    be off by about a quarter), and their w lanes untouched."""
    recs, lo, hi = W.sph_particles(6000, h=1.0)
    ctx = W.sph_context(h=1.0, dt=1e-3, box_min=lo, box_max=hi)
    m = W.sph_model(ctx, 1.0)
    three = W.Model(m.dtype, m.passes[:3], m.part, m.dims)      # density, pressure, forces: no integration
    out = []
    for lanes in (0, 4):
        with W.Simulation(oracle, three, recs, api.TAY_GPU_Z_GRID, funcs=models) as s:
            if lanes:
                s.see_split(0, lanes)
                s.see_split(2, lanes)
            s.run(1)
            a = s.agents()
            out.append(a[np.lexsort((a["p"][:, 2], a["p"][:, 1], a["p"][:, 0]))])
    ref, spl = out
    rel = np.abs(ref["density"].astype(np.float64) - spl["density"]) / ref["density"]
    assert rel.max() <= 1e-5
    assert not np.array_equal(ref["density"], spl["density"])
    for f in ("pressure_accum", "viscosity_accum"):
        scale = np.abs(ref[f][:, :3]).max()
        assert np.abs(ref[f][:, :3].astype(np.float64) - spl[f][:, :3]).max() <= 1e-4 * scale, f
        assert np.array_equal(ref[f][:, 3].view(np.uint32), spl[f][:, 3].view(np.uint32))
