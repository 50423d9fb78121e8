"""GPU Z-grid parity against the CPU oracle on identical seeded inputs.

The oracle's TAY_GPU_Z_GRID space is the CPU statement of what the GPU grid
computes (geometry of the reference's OpenCL grid, Morton keys, the
sub-cell storage order inside cells, x->y->z cell sweep), so with identical
inputs every integer output
(cell keys, ranks, storage order, neighbour counts) and every float output
(sums accumulated in the same pair order, positions after the act pass) must
match bit for bit.  Neighbour counts are also checked against CpuSimple, the
all-pairs definition of the pair set.
"""
import numpy as np
import pytest

from tay_amd import agents as A
from tay_amd import api
from tay_amd import workloads as W

pytestmark = pytest.mark.gpu


def run_pair(gpu, oracle, models, model, records, steps, space=api.TAY_GPU_Z_GRID, oracle_space=None, **kw):
    """Run the same scenario on the GPU library and the oracle; return both
    agent arrays (storage order) and the simulations' pair counts."""
    print("case: %s n=%d steps=%d" % (model.passes[0].func, len(records), steps), flush=True)
    with W.Simulation(gpu, model, records, space, **kw) as g:
        gpu.tay_set_pair_counting(g.state, 1)
        g.run(steps)
        ga = g.agents()
        gpairs = [gpu.pair_counts(g.state, i)[1] for i, p in enumerate(model.passes) if p.kind == "see"]
    with W.Simulation(oracle, model, records, oracle_space or api.TAY_GPU_Z_GRID, funcs=models, **kw) as o:
        o.run(steps)
        oa = o.agents()
        opairs = [oracle.pair_counts(o.state, i)[1] for i, p in enumerate(model.passes) if p.kind == "see"]
    return ga, oa, gpairs, opairs


def assert_bitwise(a, b, fields=None):
    assert a.dtype == b.dtype and len(a) == len(b)
    for f in fields or a.dtype.names:
        x, y = a[f], b[f]
        if x.dtype.kind == "f":
            x, y = x.view(np.uint32), y.view(np.uint32)
        bad = np.nonzero((x != y).reshape(len(a), -1).any(axis=1))[0]
        assert len(bad) == 0, "field %s differs at %d agents, first %s: gpu %s oracle %s" % (
            f, len(bad), bad[:5], a[f][bad[:3]], b[f][bad[:3]])


def xyz_fields(dtype):
    """Fields to compare bitwise; the unused .w lane of float4 fields is
    compared too: the model helpers define it (a result keeps its first
    operand's w), so the GPU and the oracle must agree on it as well."""
    return list(dtype.names)


def test_plain_one_step_bit_exact(gpu, oracle, models):
    recs = W.plain_agents(10000, seed=11)
    ga, oa, gp, op = run_pair(gpu, oracle, models, W.plain_model(40.0), recs, 1)
    assert_bitwise(ga, oa)
    assert gp == op


def test_plain_trajectory_20_steps_bit_exact(gpu, oracle, models):
    recs = W.plain_agents(10000, seed=12)
    ga, oa, gp, op = run_pair(gpu, oracle, models, W.plain_model(40.0), recs, 20)
    assert_bitwise(ga, oa)
    assert gp == op


def test_plain_counts_match_all_pairs(gpu, oracle, models):
    recs = W.plain_agents(4000, seed=13)
    ga, oa, gp, op = run_pair(gpu, oracle, models, W.plain_model(60.0), recs, 1,
                              oracle_space=api.TAY_CPU_SIMPLE)
    ga = ga[np.argsort(ga["result_index"])]
    oa = oa[np.argsort(oa["result_index"])]
    assert np.array_equal(ga["last_ncount"], oa["last_ncount"])
    assert gp == op
    # a symmetric test counts every unordered pair twice
    assert ga["last_ncount"].sum() % 2 == 0


def test_reference_test_model_bit_exact(gpu, oracle, models):
    recs = W.test_agents(10000, seed=1)
    ga, oa, gp, op = run_pair(gpu, oracle, models, W.test_model(50.0), recs, 5)
    assert_bitwise(ga, oa)
    assert gp == op


def test_flocking_clustered_bit_exact(gpu, oracle, models):
    recs = W.flock_agents(30000, seed=5, clusters=8, sigma=30.0, box=600.0)
    ga, oa, gp, op = run_pair(gpu, oracle, models, W.flock_model(25.0), recs, 3)
    assert_bitwise(ga, oa)
    assert gp == op


def test_flocking_70k_bit_exact(gpu, oracle, models):
    """70,000 clustered flocking agents: the seer count picks 32-seer tiles
    (64-seer tiles would give the SIMDs fewer than two wavefronts each), the
    radix sort runs its multi-kernel path; bit-exact over two steps."""
    recs = W.flock_agents(70000, seed=15, clusters=16, sigma=30.0, box=1000.0)
    ga, oa, gp, op = run_pair(gpu, oracle, models, W.flock_model(25.0), recs, 2)
    assert_bitwise(ga, oa)
    assert gp == op


def test_grid_geometry_bit_exact(gpu, oracle, models):
    recs = W.flock_agents(5000, seed=6, clusters=4, sigma=40.0, box=800.0)
    model = W.flock_model(25.0)
    out = []
    for lib, funcs in ((gpu, None), (oracle, models)):
        with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID, funcs=funcs) as s:
            s.run(1)
            out.append(lib.grid_info(s.state, s.group))
    (o1, s1, c1, n1), (o2, s2, c2, n2) = out
    assert np.array_equal(o1[:3].view(np.uint32), o2[:3].view(np.uint32))
    assert np.array_equal(s1[:3].view(np.uint32), s2[:3].view(np.uint32))
    assert np.array_equal(c1[:3], c2[:3]) and n1 == n2


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1000, 4097, 15360, 15361, 16384, 16385])
def test_ragged_sizes(gpu, oracle, models, n):
    recs = W.plain_agents(n, seed=100 + n, box=200.0)
    ga, oa, gp, op = run_pair(gpu, oracle, models, W.plain_model(30.0), recs, 2)
    assert_bitwise(ga, oa)
    assert gp == op


def test_all_agents_in_one_cell(gpu, oracle, models):
    g = W.rng(7)
    recs = W.plain_agents(3000, seed=7)
    recs["p"][:, :3] = g.uniform(0.0, 5.0, size=(3000, 3)).astype(np.float32)
    ga, oa, gp, op = run_pair(gpu, oracle, models, W.plain_model(40.0), recs, 2)
    assert_bitwise(ga, oa)
    assert gp == op


def test_agents_on_cell_boundaries(gpu, oracle, models):
    # positions on a lattice of the cell size: many land exactly on faces
    g = W.rng(8)
    recs = W.plain_agents(5000, seed=8)
    recs["p"][:, :3] = (g.integers(0, 13, size=(5000, 3)) * 80.0).astype(np.float32)
    recs["v"][:] = 0.0
    ga, oa, gp, op = run_pair(gpu, oracle, models, W.plain_model(40.0), recs, 2)
    assert_bitwise(ga, oa)
    assert gp == op


def test_sph_two_see_passes_bit_exact(gpu, oracle, models):
    recs, lo, hi = W.sph_particles(8000, h=1.0)
    ctx = W.sph_context(h=1.0, dt=1e-3, box_min=lo, box_max=hi)
    ga, oa, gp, op = run_pair(gpu, oracle, models, W.sph_model(ctx, 1.0), recs, 3)
    assert_bitwise(ga, oa)
    assert gp == op


def test_gpu_simple_space_matches_oracle_simple(gpu, oracle, models):
    recs = W.plain_agents(3000, seed=9)
    ga, oa, gp, op = run_pair(gpu, oracle, models, W.plain_model(50.0), recs, 2,
                              space=api.TAY_GPU_SIMPLE, oracle_space=api.TAY_CPU_SIMPLE)
    assert_bitwise(ga, oa)
    assert gp == op


def test_empty_group_runs(gpu):
    with W.Simulation(gpu, W.plain_model(40.0), W.plain_agents(0), api.TAY_GPU_Z_GRID, capacity=16) as s:
        s.run(3)
        assert len(s.agents()) == 0


def test_sorted_storage_properties_large(gpu):
    """1M clustered agents: the storage is a permutation of the input,
    ordered by Morton key, ranks count up inside each cell."""
    recs = W.flock_agents(1_000_000, seed=21)
    with W.Simulation(gpu, W.flock_model(25.0), recs, api.TAY_GPU_Z_GRID) as s:
        s.run(1)
        a = s.agents()
    assert np.array_equal(np.sort(a["result_index"]), np.arange(len(recs)))
    keys = a["part_i"].astype(np.int64)
    assert np.all(np.diff(keys) >= 0)
    starts = np.r_[0, np.nonzero(np.diff(keys))[0] + 1]
    expect = np.arange(len(a)) - np.repeat(starts, np.diff(np.r_[starts, len(a)]))
    assert np.array_equal(a["cell_agent_i"], expect)


@pytest.mark.parametrize("bits", [0, 1, 3])
def test_order_bits_bit_exact(gpu, oracle, models, bits):
    """Every sub-cell storage order (2, the default, is covered above) gives
    the oracle's storage order and sums bit for bit."""
    recs = W.flock_agents(20000, seed=63, clusters=8, sigma=30.0, box=600.0)
    model = W.flock_model(25.0)
    print("case: flock order_bits=%d" % bits, flush=True)
    out = []
    for lib, funcs in ((gpu, None), (oracle, models)):
        with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID, funcs=funcs) as s:
            assert lib.tay_set_order_bits(s.state, s.group, bits)
            s.run(4)
            out.append(s.agents())
    assert_bitwise(out[0], out[1])


@pytest.mark.parametrize("dims", [1, 2])
def test_lower_dimensional_grids_bit_exact(gpu, oracle, models, dims):
    """1D and 2D Z-grids (Morton keys of 16 bits per axis in 2D, the index
    itself in 1D): the box test uses the first `dims` axes only."""
    m = W.test_model(30.0)
    model = W.Model(m.dtype, m.passes, m.part, dims)
    recs = W.test_agents(6000 if dims == 2 else 3000, seed=70 + dims, clump=0.1)
    if dims == 1:
        recs["p"][:, 1:3] = 0.0
        recs["v"][:, 1:3] = 0.0
    print("case: test model dims=%d" % dims, flush=True)
    ga, oa, gp, op = run_pair(gpu, oracle, models, model, recs, 3)
    assert_bitwise(ga, oa)
    assert gp == op
    _, sa, _, sp = run_pair(gpu, oracle, models, model, recs, 1, oracle_space=api.TAY_CPU_SIMPLE)
    ga1, _, gp1, _ = run_pair(gpu, oracle, models, model, recs, 1)
    assert gp1 == sp


def test_2d_key_of_the_top_corner_sub_cell(gpu, oracle, models):
    """A 2D grid of 8192 cells per axis with 3 sub-cell bits: the agent in
    the top corner sub-cell has storage key 0xFFFFFFFF (Morton of 65535,
    65535).  Its cell is (8191, 8191) like any other key's; bit-exact with the
    oracle, and the corner agent's pairs are all there."""
    m = W.test_model(2.0, box=10000.0)
    model = W.Model(m.dtype, m.passes, 1.0, 2)
    g = W.rng(71)
    recs = W.test_agents(3000, seed=71)
    recs["p"][:, :2] = g.uniform(8150.0, 8191.5, size=(3000, 2)).astype(np.float32)
    recs["p"][:, 2] = 0.0
    recs["v"][:] = 0.0
    recs["p"][0, :2] = 0.0          # stretches the box to [0, 8191.5]
    recs["p"][1, :2] = 8191.5       # the top corner: q = 8191.99.., sub-cell 7 on both axes
    recs["p"][2, :2] = 8190.5       # a neighbour of the corner agent
    print("case: 2D top-corner key", flush=True)
    out = []
    for lib, funcs in ((gpu, None), (oracle, models)):
        with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID, funcs=funcs, shared_mb=512) as s:
            assert lib.tay_set_order_bits(s.state, s.group, 3)
            s.run(1)
            a = s.agents()
            _, _, counts, _ = lib.grid_info(s.state, s.group)
            out.append(a)
            assert counts[0] == 8192 and counts[1] == 8192
    assert_bitwise(out[0], out[1])
    corner = out[0][out[0]["result_index"] == 1][0]
    assert corner["part_i"] == 0xFFFFFFFF >> 6
    assert corner["b_buffer_count"] == 0    # the act pass cleared it; the sum moved into f_buffer
    assert np.all(corner["f_buffer"][:2] < 0)


def test_unequal_axis_radii_bit_exact(gpu, oracle, models):
    """Per-axis see radii that differ take the per-axis box test instead of
    the single max-of-axes compare used when all radii are equal."""
    m = W.plain_model(40.0)
    see = m.passes[0]
    model = W.Model(m.dtype, [W.Pass("see", see.func, see.context, 40.0, radii=(40.0, 25.0, 60.0)), m.passes[1]],
                    m.part)
    recs = W.plain_agents(8000, seed=65)
    ga, oa, gp, op = run_pair(gpu, oracle, models, model, recs, 3)
    assert_bitwise(ga, oa)
    assert gp == op
    _, _, sp, _ = run_pair(gpu, oracle, models, model, recs, 1)
    _, _, _, cp = run_pair(gpu, oracle, models, model, recs, 1, oracle_space=api.TAY_CPU_SIMPLE)
    assert sp == cp


@pytest.mark.parametrize("walk", [0, 8, 64])
def test_lane_walk_bit_exact(gpu, oracle, models, walk):
    """Sparse lane groups walking per lane meet every seer's pairs in the
    same order as groups that keep splitting (walk 0): bit-exact with the
    oracle (sparse clusters make many tiles span far-apart seers)."""
    recs = W.flock_agents(20000, seed=64, clusters=40, sigma=15.0, box=2000.0)
    model = W.flock_model(25.0)
    print("case: flock lane_walk=%d" % walk, flush=True)
    out = []
    for lib, funcs in ((gpu, None), (oracle, models)):
        with W.Simulation(lib, model, recs, api.TAY_GPU_Z_GRID, funcs=funcs) as s:
            if funcs is None:
                assert lib.tay_set_lane_walk(s.state, walk)
            s.run(3)
            out.append(s.agents())
    assert_bitwise(out[0], out[1])


def test_tile_seers_bit_exact(gpu, oracle, models):
    """Tiles of 64, 32 or 16 seers (and the choice by seer count), with
    narrower starting lane groups and without the lane walk: every seer
    still meets its pairs in the reference's order, bit-exact with the
    oracle, on a clustered flocking state (dense flocks and sparse
    outskirts) and on the uniform 10k-agent plain workload."""
    cases = [(W.flock_model(25.0), W.flock_agents(20000, seed=67, clusters=30, sigma=20.0, box=1500.0), 3),
             (W.plain_model(40.0), W.plain_agents(10000, seed=68), 2)]
    for model, recs, steps in cases:
        with W.Simulation(oracle, model, recs, api.TAY_GPU_Z_GRID, funcs=models) as s:
            s.run(steps)
            want = s.agents()
        for seers, width, walk in ((0, 64, 64), (64, 64, 64), (32, 64, 64), (16, 64, 64), (16, 8, 64), (32, 16, 0)):
            print("case: %s seers=%d width=%d walk=%d" % (model.passes[0].func, seers, width, walk), flush=True)
            with W.Simulation(gpu, model, recs, api.TAY_GPU_Z_GRID) as s:
                assert gpu.tay_set_tile_seers(s.state, seers)
                assert gpu.tay_set_tile_width(s.state, width)
                assert gpu.tay_set_lane_walk(s.state, walk)
                s.run(steps)
                assert_bitwise(s.agents(), want)
                assert not gpu.tay_set_tile_seers(s.state, 48)


def test_device_sqrt_is_ieee_for_every_float(gpu):
    """The square root model code gets on the device (sqrt/sqrtf of a float)
    equals the IEEE square root bit for bit on all 2^32 float inputs, both
    the compiler's float square root and the double one rounded to float."""
    import ctypes as C
    bad = (C.c_ulonglong * 2)()
    first = (C.c_uint * 2)()
    print("case: sqrt self-test over 2^32 inputs", flush=True)
    assert gpu.tay_gpu_sqrt_selftest(0, 1 << 32, bad, first), gpu.tay_gpu_last_error().decode()
    assert bad[0] == 0, "differs from the float square root at %d inputs, first 0x%08x" % (bad[0], first[0])
    assert bad[1] == 0, "differs from the rounded double square root at %d inputs, first 0x%08x" % (bad[1], first[1])


def test_flocking_coincident_agents_bit_exact(gpu, oracle, models):
    """Pairs of agents on the same spot: the offset's length is the square
    root of 0 (the dialect's full IEEE path on those lanes) and
    float4_normalize_to takes its near-zero branch; bit-exact with the
    oracle, whose host sqrt is the IEEE one."""
    recs = W.flock_agents(20000, seed=66, clusters=8, sigma=20.0, box=600.0)
    recs["p"][1::2] = recs["p"][0::2]
    print("case: flock with coincident pairs", flush=True)
    ga, oa, gp, op = run_pair(gpu, oracle, models, W.flock_model(25.0), recs, 3)
    assert_bitwise(ga, oa)
    assert gp == op


def test_page_locked_host_storage(gpu):
    """Group storage comes from tay_gpu_host_alloc on a GPU host: zeroed,
    at most 4 GiB per allocation, and agents fetched into it after a run
    equal the oracle's (the other tests of this file cover the values)."""
    import ctypes as C
    p = gpu.tay_gpu_host_alloc(1 << 20)
    assert p
    assert bytes((C.c_char * (1 << 20)).from_address(p)) == bytes(1 << 20)
    gpu.tay_gpu_host_free(p)
    assert not gpu.tay_gpu_host_alloc((4 << 30) + 1)
    assert not gpu.tay_gpu_host_alloc(0)
