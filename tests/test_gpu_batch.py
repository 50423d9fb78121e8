"""Batched ensembles on the GPU: every instance of a batch must compute,
bit for bit, what a state of its own computes with TAY_GPU_Z_GRID and the
same cell size and cell table, which in turn is the oracle's result.  The
captured-graph path must match the launch-by-launch path on every replay.
"""
import numpy as np
import pytest

from tay_amd import api
from tay_amd import workloads as W

from test_gpu_parity import assert_bitwise

pytestmark = pytest.mark.gpu


def small_ensemble(instances=12, seed=3):
    return W.ensemble(instances, seed=seed, counts=(40, 3000), boxes=(60.0, 400.0), radii=(10.0, 40.0))


def parts_and_tables(ens, factor=1.0):
    part = [factor * m.passes[0].radius for _, m, _ in ens]
    tables = [W.cell_table_mb(side, p) for (_, _, side), p in zip(ens, part)]
    return part, tables


def per_instance(lib, ens, part, tables, steps, funcs=None, space=api.TAY_GPU_Z_GRID):
    out, pairs = [], []
    for (rec, model, _), p, mb in zip(ens, part, tables):
        with W.Simulation(lib, model, rec, space, funcs=funcs, shared_mb=mb, part=p) as s:
            if funcs is None:
                lib.tay_set_pair_counting(s.state, 1)
            s.run(steps)
            out.append(s.agents())
            pairs.append(lib.pair_counts(s.state, 0)[1])
    return out, pairs


def test_batch_matches_per_instance_states(gpu):
    ens = small_ensemble()
    part, tables = parts_and_tables(ens)
    print("case: batch 12 instances, 4 steps, eager", flush=True)
    with W.Batch(gpu, ens, part, tables, count_pairs=True) as b:
        b.run(4, graph=False)
        got = b.agents()
        bpairs = b.pair_count(0)
    want, pairs = per_instance(gpu, ens, part, tables, 4)
    for g, w in zip(got, want):
        assert_bitwise(g, w)
    assert bpairs == sum(pairs)


def test_batch_matches_oracle(gpu, oracle, models):
    ens = small_ensemble(5, seed=4)
    part, tables = parts_and_tables(ens, 1.5)
    print("case: batch 5 instances, 3 steps, graph, vs oracle", flush=True)
    with W.Batch(gpu, ens, part, tables) as b:
        b.run(3, graph=True)
        got = b.agents()
    want, _ = per_instance(oracle, ens, part, tables, 3, funcs=models)
    for g, w in zip(got, want):
        assert_bitwise(g, w)


def test_batch_graph_replays_match_eager(gpu):
    ens = small_ensemble(9, seed=5)
    part, tables = parts_and_tables(ens)
    print("case: batch graph vs eager, 3 runs of 5 steps", flush=True)
    with W.Batch(gpu, ens, part, tables, count_pairs=True) as g, W.Batch(gpu, ens, part, tables,
                                                                        count_pairs=True) as e:
        for run in range(3):
            g.run(5, graph=True)
            e.run(5, graph=False)
            gpu.tay_gpu_synchronize()
            for x, y in zip(g.agents(), e.agents()):
                assert_bitwise(x, y)
        assert g.pair_count(0) == e.pair_count(0)


def test_batch_tile_width_does_not_change_results(gpu):
    ens = small_ensemble(6, seed=6)
    part, tables = parts_and_tables(ens)
    print("case: batch tile widths 64/32/8, lane walks 0/8/64, tile orders", flush=True)
    results = []
    for width, walk, order in ((64, 0, 0), (32, 0, 1), (8, 0, 0), (64, 8, 0), (64, 64, 1), (16, 64, 0)):
        with W.Batch(gpu, ens, part, tables) as b:
            assert gpu.tay_gpu_batch_set_tile_width(b.batch, width)
            assert gpu.tay_gpu_batch_set_lane_walk(b.batch, walk)
            assert gpu.tay_gpu_batch_set_tile_order(b.batch, order)
            b.run(3, graph=True)
            results.append(b.agents())
    for other in results[1:]:
        for x, y in zip(results[0], other):
            assert_bitwise(x, y)


def test_batch_autotune_restores_agents(gpu):
    import ctypes as C
    ens = small_ensemble(6, seed=8)
    part, _ = parts_and_tables(ens)
    tables = [W.cell_table_mb(side, 0.5 * m.passes[0].radius) for _, m, side in ens]
    print("case: batch autotune 3 factors x 2 widths", flush=True)
    with W.Batch(gpu, ens, part, tables) as b:
        before = b.agents()
        factors = (C.c_float * 3)(1.0, 1.5, 2.0)
        widths = (C.c_uint * 2)(64, 16)
        walks = (C.c_uint * 2)(0, 64)
        table = (C.c_double * 12)()
        bf, bw, bl = C.c_float(), C.c_uint(), C.c_uint()
        assert gpu.tay_gpu_batch_autotune(b.batch, factors, 3, widths, 2, walks, 2, 2, table, C.byref(bf),
                                          C.byref(bw), C.byref(bl)), gpu.tay_gpu_last_error().decode()
        assert bf.value in (1.0, 1.5, 2.0) and bw.value in (64, 16) and bl.value in (0, 64)
        assert all(t > 0 for t in table)
        for x, y in zip(before, b.agents()):
            assert_bitwise(x, y)
        b.run(2, graph=True)
        got = b.agents()
    chosen = [bf.value * m.passes[0].radius for _, m, _ in ens]
    want, _ = per_instance(gpu, ens, chosen, tables, 2)
    for x, y in zip(got, want):
        assert_bitwise(x, y)


def test_batch_density_part_sizes_and_autotune(gpu):
    """Part sizes by agent density follow their rule (the host restatement
    W.density_part, both clamps, equal bounds), and the density autotuner's
    choice runs bit for bit like per-instance states with those sizes."""
    import ctypes as C
    ens = small_ensemble(6, seed=9)
    radius = [m.passes[0].radius for _, m, _ in ens]
    tables = [W.cell_table_mb(side, 0.5 * r) for (_, _, side), r in zip(ens, radius)]
    fptr = C.POINTER(C.c_float)
    print("case: density part sizes, 6 instances", flush=True)
    with W.Batch(gpu, ens, radius, tables) as b:
        for target, lo, hi in ((4.0, 0.5, 4.0), (1e-6, 0.5, 4.0), (1e9, 0.5, 4.0), (8.0, 1.25, 1.25)):
            parts = np.zeros((len(ens), 4), np.float32)
            assert gpu.tay_gpu_batch_density_part_sizes(b.batch, target, lo, hi, parts.ctypes.data_as(fptr)), \
                gpu.tay_gpu_last_error().decode()
            want = np.array([W.density_part(rec, r, target, lo, hi) for (rec, _, _), r in zip(ens, radius)])
            np.testing.assert_allclose(parts[:, 0], want, rtol=1e-6)
            assert (parts[:, 1] == parts[:, 0]).all() and (parts[:, 2] == parts[:, 0]).all()
            if target < 1e-3 or lo == hi:
                np.testing.assert_allclose(parts[:, 0], lo * np.array(radius), rtol=1e-6)
            if target > 1e3:
                np.testing.assert_allclose(parts[:, 0], hi * np.array(radius), rtol=1e-6)
        assert not gpu.tay_gpu_batch_density_part_sizes(b.batch, 0.0, 0.5, 4.0, parts.ctypes.data_as(fptr))
        assert b"agents per cell" in gpu.tay_gpu_last_error()
        assert not gpu.tay_gpu_batch_density_part_sizes(b.batch, 4.0, 2.0, 1.0, parts.ctypes.data_as(fptr))
        assert b"cell factor bounds" in gpu.tay_gpu_last_error()

        print("case: density autotune 3 targets x 2 lane walks", flush=True)
        before = b.agents()
        targets = (C.c_float * 3)(1.0, 4.0, 16.0)
        widths = (C.c_uint * 1)(64)
        walks = (C.c_uint * 2)(16, 64)
        table = (C.c_double * 6)()
        bt, bw, bl = C.c_float(), C.c_uint(), C.c_uint()
        chosen = np.zeros((len(ens), 4), np.float32)
        assert gpu.tay_gpu_batch_autotune_density(b.batch, targets, 3, 0.5, 4.0, widths, 1, walks, 2, 2, table,
                                                  C.byref(bt), chosen.ctypes.data_as(fptr), C.byref(bw),
                                                  C.byref(bl)), gpu.tay_gpu_last_error().decode()
        assert bt.value in (1.0, 4.0, 16.0) and bw.value == 64 and bl.value in (16, 64)
        assert all(x > 0 for x in table)
        applied = np.zeros((len(ens), 4), np.float32)
        assert gpu.tay_gpu_batch_get_part_sizes(b.batch, applied.ctypes.data_as(fptr))
        np.testing.assert_array_equal(applied, chosen)
        want = np.array([W.density_part(rec, r, bt.value, 0.5, 4.0) for (rec, _, _), r in zip(ens, radius)])
        np.testing.assert_allclose(chosen[:, 0], want, rtol=1e-6)
        for x, y in zip(before, b.agents()):
            assert_bitwise(x, y)
        b.run(2, graph=True)
        got = b.agents()
    want, _ = per_instance(gpu, ens, [float(x) for x in chosen[:, 0]], tables, 2)
    for x, y in zip(got, want):
        assert_bitwise(x, y)


def test_batch_of_more_than_256_instances(gpu):
    """Instances beyond 256 (the batch sort works per instance, whatever
    their number): still identical to per-instance states."""
    ens = W.ensemble(260, seed=9, counts=(20, 200), boxes=(60.0, 200.0), radii=(10.0, 30.0))
    part, tables = parts_and_tables(ens)
    print("case: batch 260 instances, 2 steps, graph", flush=True)
    with W.Batch(gpu, ens, part, tables) as b:
        b.run(2, graph=True)
        got = b.agents()
    want, _ = per_instance(gpu, ens, part, tables, 2)
    for g, w in zip(got, want):
        assert_bitwise(g, w)


def test_batch_rejects_bad_arguments(gpu):
    ens = small_ensemble(2)
    part, tables = parts_and_tables(ens)
    with W.Batch(gpu, ens, part, tables) as b:
        assert not gpu.tay_gpu_batch_set_tile_width(b.batch, 48)
        assert b"tile width" in gpu.tay_gpu_last_error()
        assert not gpu.tay_gpu_batch_set_radii(b.batch, 1, None)
        assert b"not a see pass" in gpu.tay_gpu_last_error()


def test_full_size_ensemble_matches_per_instance_states_and_oracle(gpu, oracle, models):
    """The benchmark's ensemble at its real shape: 256 instances of 1k-64k
    agents (3.9M in all), box sides 200-1500, radii 10-40, cells of about 8
    agents by density (0.5 to 4 radii); two graph-replayed steps, every
    instance bit for bit equal to its own TAY_GPU_Z_GRID state, and four
    instances (the smallest, the largest and two seeded others) equal to the
    oracle's."""
    ens = W.ensemble(256, seed=7)
    part = [W.density_part(rec, m.passes[0].radius, 8.0, 0.5, 4.0) for rec, m, _ in ens]
    tables = [W.cell_table_mb(side, p) for (_, _, side), p in zip(ens, part)]
    print("case: full-size ensemble, 256 instances, 2 steps, graph", flush=True)
    with W.Batch(gpu, ens, part, tables, count_pairs=True) as b:
        b.run(2, graph=True)
        got = b.agents()
        bpairs = b.pair_count(0)
    want, pairs = per_instance(gpu, ens, part, tables, 2)
    for g, w in zip(got, want):
        assert_bitwise(g, w)
    assert bpairs == sum(pairs)
    sizes = np.array([len(r) for r, _, _ in ens])
    pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes))} | set(W.rng(11).choice(256, 2, replace=False).tolist()))
    print("case: full-size ensemble, oracle instances %s" % pick, flush=True)
    sub = [ens[i] for i in pick]
    owant, _ = per_instance(oracle, sub, [part[i] for i in pick], [tables[i] for i in pick], 2, funcs=models)
    for i, w in zip(pick, owant):
        assert_bitwise(got[i], w)
