"""Pins of the oracle's balanced see order (oracle/oracle.c,
gpu_z_grid_see_tiles), on the CPU.

* A numpy restatement of the order, written from its definition (tiles of S
  consecutive seers; lane groups that halve while their box spans more than
  `limit` cells and they are wider than the lane walk, walk per seer when
  they may not halve, else sweep their cells 64 at a time in chunks: the
  segment's M candidates in sweep order cut into P = 64 / RUN parts of
  L = ceil(M / P), chunk k holding the k-th run of RUN consecutive
  candidates of every part, RUN = 4), accumulates the plain model's
  neighbour sums in float32 in that order: bit-exact against the oracle for
  several tile layouts.  A planted change of the order (the slots of a chunk
  taken in descending order, runs of another length) fails it.
* Where no segment holds more than 64 candidates the order is the reference
  sweep order, so the two orders agree bit for bit.
* In general the pairs and neighbour counts are those of the reference
  order and the sums agree within rounding.
"""
import numpy as np
import pytest

from tay_amd import api
from tay_amd import workloads as W

F = np.float32


def run_oracle(oracle, models, model, recs, order, seers=64, width=64, walk=64, steps=1):
    with W.Simulation(oracle, model, recs, api.TAY_GPU_Z_GRID, funcs=models) as s:
        assert oracle.tay_set_see_order(s.state, order)
        assert oracle.tay_set_tile_seers(s.state, seers)
        assert oracle.tay_set_tile_width(s.state, width)
        assert oracle.tay_set_lane_walk(s.state, walk)
        s.run(steps)
        return s.agents(), oracle.grid_info(s.state, s.group), oracle.pair_counts(s.state, 0)[1]


def see_only(model):
    return W.Model(model.dtype, model.passes[:1], model.part, model.dims)


def cells_of_tags(part_i):
    """(x, y, z) of 3D Morton cell keys: bit 3b + a is bit b of axis a."""
    out = np.zeros((len(part_i), 3), np.int64)
    for a in range(3):
        for b in range(10):
            out[:, a] |= ((part_i.astype(np.int64) >> (3 * b + a)) & 1) << b
    return out


def cell_axis(p, origin, size, count):
    c = int(np.floor(F(F(p) - F(origin)) / F(size)))
    return min(max(c, 0), count - 1)


RUN = 4     # candidates per run of a chunk (TAY_BALANCE_RUN)


def balanced_plain_sums(agents, info, radius, seers, width0, walk, run=RUN, ascending=True):
    """Neighbour sums and counts of the plain model (a pair counts when the
    squared distance is at most r^2) accumulated in the balanced order."""
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
    order = [[] for _ in range(n)]      # per seer: candidates in the order met

    def box_ok(i, j):
        d = pos[i] - pos[j]
        return i != j and not np.any(np.abs(d) > r)

    def own_range(i):
        lo, hi = [], []
        for k in range(3):
            p = pos[i, k]
            m = F(F(abs(p) + r) * F(2.0 ** -20))
            lo.append(cell_axis(F(F(p - r) - m), origin[k], size[k], counts[k]))
            hi.append(cell_axis(F(F(p + r) + m), origin[k], size[k], counts[k]))
        return lo, hi

    def cells_in(lo, hi):
        return [(x, y, z) for x in range(lo[0], hi[0] + 1) for y in range(lo[1], hi[1] + 1)
                for z in range(lo[2], hi[2] + 1)]

    for i0 in range(0, n, seers):
        n_active = min(seers, n - i0)
        width = min(width0, seers)
        w0 = width
        beg = 0
        while beg < n_active:
            members = range(i0 + beg, i0 + min(beg + width, n_active))
            lo, hi = [], []
            for k in range(3):
                tmin = F(pos[list(members), k].min())
                tmax = F(pos[list(members), k].max())
                mlo = F(F(abs(tmin) + r) * F(2.0 ** -20))
                mhi = F(F(abs(tmax) + r) * F(2.0 ** -20))
                lo.append(cell_axis(F(F(tmin - r) - mlo), origin[k], size[k], counts[k]))
                hi.append(cell_axis(F(F(tmax + r) + mhi), origin[k], size[k], counts[k]))
            ncells = (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1)
            if ncells > limit and width > 1 and width > walk:
                width //= 2
                continue
            if ncells > limit and width > 1:
                for i in members:
                    olo, ohi = own_range(i)
                    for c in cells_in(olo, ohi):
                        b, e = ranges.get(c, (0, 0))
                        order[i] += [j for j in range(b, e) if box_ok(i, j)]
            else:
                cl = cells_in(lo, hi)
                for c0 in range(0, len(cl), 64):
                    seq = []
                    for c in cl[c0:c0 + 64]:
                        b, e = ranges.get(c, (0, 0))
                        seq += range(b, e)
                    m = len(seq)
                    if not m:
                        continue
                    parts = 64 // run
                    plen = -(-m // parts)
                    chunks = []
                    for k in range(-(-plen // run)):
                        chunks.append([seq[j * plen + o] for j in range(parts) for o in range(k * run, (k + 1) * run)
                                       if o < plen and j * plen + o < m])
                    for ch in chunks:
                        for j in (ch if ascending else ch[::-1]):
                            for i in members:
                                if box_ok(i, j):
                                    order[i].append(j)
            beg += width
            while width < w0 and beg % (2 * width) == 0:
                width *= 2
    sums = np.zeros((n, 3), F)
    cnt = np.zeros(n, np.uint32)
    for i in range(n):
        js = [j for j in order[i] if F(F(F(F(pos[j, 0] - pos[i, 0]) ** 2) + F(F(pos[j, 1] - pos[i, 1]) ** 2))
                                       + F(F(pos[j, 2] - pos[i, 2]) ** 2)) <= r2]
        cnt[i] = len(js)
        if js:
            sums[i] = np.add.accumulate(pos[js], axis=0, dtype=F)[-1]
    return sums, cnt


@pytest.mark.parametrize("seers,width,walk", [(64, 64, 64), (32, 16, 0), (16, 8, 8)])
def test_numpy_restatement_matches_oracle(oracle, models, seers, width, walk):
    recs = W.plain_agents(1500, seed=41, box=300.0)
    recs["p"][:500, :3] = (recs["p"][:500, :3] * F(0.2)).astype(F)     # a dense corner: many chunks
    model = see_only(W.plain_model(30.0, box=300.0))
    agents, info, _ = run_oracle(oracle, models, model, recs, api.TAY_SEE_ORDER_BALANCED, seers, width, walk)
    sums, cnt = balanced_plain_sums(agents, info, 30.0, seers, width, walk)
    assert np.array_equal(agents["ncount"], cnt)
    assert np.array_equal(agents["nsum"][:, :3].view(np.uint32), sums.view(np.uint32))
    # the order matters: planted variants of it do not reproduce the oracle
    for kw in ({"ascending": False}, {"run": 1}, {"run": 16}, {"run": 64}):
        s2, c2 = balanced_plain_sums(agents, info, 30.0, seers, width, walk, **kw)
        assert np.array_equal(c2, cnt)
        assert not np.array_equal(s2.view(np.uint32), sums.view(np.uint32)), kw


def test_sparse_segments_give_the_reference_order(oracle, models):
    """At most 64 candidates per 64-cell segment: one chunk per segment in
    sweep order, so the balanced order is the reference order."""
    recs = W.plain_agents(3000, seed=42, box=2000.0)
    model = W.plain_model(40.0, box=2000.0)
    ref, _, rp = run_oracle(oracle, models, model, recs, api.TAY_SEE_ORDER_REFERENCE, steps=3)
    bal, _, bp = run_oracle(oracle, models, model, recs, api.TAY_SEE_ORDER_BALANCED, steps=3)
    for f in ref.dtype.names:
        assert np.ascontiguousarray(ref[f]).tobytes() == np.ascontiguousarray(bal[f]).tobytes(), f
    assert rp == bp


def test_dense_state_same_pairs_other_rounding(oracle, models):
    recs = W.flock_agents(20000, seed=43, clusters=4, sigma=30.0, box=400.0)
    model = see_only(W.flock_model(25.0))
    ref, _, rp = run_oracle(oracle, models, model, recs, api.TAY_SEE_ORDER_REFERENCE)
    bal, _, bp = run_oracle(oracle, models, model, recs, api.TAY_SEE_ORDER_BALANCED)
    assert rp == bp
    assert np.array_equal(ref["cohesion_count"], bal["cohesion_count"])
    assert np.array_equal(ref["separation_count"], bal["separation_count"])
    n = np.maximum(ref["cohesion_count"], 1).astype(np.float64)
    d = np.abs(ref["cohesion"][:, :3].astype(np.float64) - bal["cohesion"][:, :3]).max(axis=1)
    assert np.all(d <= 1e-5 * n * 25.0)
    assert not np.array_equal(ref["cohesion"], bal["cohesion"])


def test_unknown_order_is_refused(oracle):
    state = oracle.tay_create_state()
    try:
        assert not oracle.tay_set_see_order(state, 2)
        assert not oracle.tay_set_tile_seers(state, 48)
    finally:
        oracle.tay_destroy_state(state)
