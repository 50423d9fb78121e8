"""FP32 see outputs of the GPU Z-grid against an FP64 recomputation over the
same pairs.

One see pass (no act) from seeded inputs; the pair set is decided with the
model's own float32 predicate, evaluated here in numpy float32 (IEEE, no
contraction, correctly rounded sqrt), so integer outputs must match exactly
and every float sum must satisfy the condition-aware bound
    |x32 - x64| <= 1e-5 * sum |terms|
(a plain relative bound is ill-conditioned where terms cancel, e.g. sums of
offsets).  Sums of positive terms (SPH density) also meet 1e-5 relative.
Each check runs the pass unsplit and split over 4 lanes per seer
(tay_set_see_split), whose partial sums merge in a fixed tree.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from tay_amd import api
from tay_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-5


def see_only(model):
    return W.Model(model.dtype, [model.passes[0]], model.part, model.dims)


def run_see(gpu, model, recs, lanes=0):
    """Agents after one see pass (split over `lanes` lanes per seer when
    lanes > 1), in storage order; a see pass changes only the seers' sums,
    so positions and velocities are the inputs'."""
    with W.Simulation(gpu, see_only(model), recs, api.TAY_GPU_Z_GRID) as s:
        if lanes > 1:
            s.see_split(0, lanes)
        s.run(1)
        return s.agents()


def candidate_pairs(p, r):
    """Every (i, j), i != j, within Chebyshev distance r (a superset of the
    box test's pairs, padded for rounding)."""
    tree = cKDTree(p.astype(np.float64))
    pairs = tree.query_pairs(r * (1 + 1e-5), p=np.inf, output_type="ndarray")
    both = np.concatenate([pairs, pairs[:, ::-1]])
    return both[:, 0], both[:, 1]


def box_ok(p, i, j, r):
    d = p[i] - p[j]
    return (np.abs(d) <= np.float32(r)).all(axis=1)


def check_sum(x32, terms64, seer, n, name):
    x64 = np.zeros((n,) + terms64.shape[1:])
    mag = np.zeros_like(x64)
    np.add.at(x64, seer, terms64)
    np.add.at(mag, seer, np.abs(terms64))
    err = np.abs(x32.astype(np.float64) - x64)
    bad = err > TOL * mag + 1e-30
    assert not bad.any(), "%s: %d components beyond the bound, worst %g of %g" % (
        name, bad.sum(), err[bad].max(), (TOL * mag)[bad][np.argmax(err[bad])])
    return float((err / np.maximum(mag, 1e-30)).max())


@pytest.mark.parametrize("lanes", [0, 4])
def test_plain_sums_within_fp64_bound(gpu, lanes):
    recs = W.plain_agents(20000, seed=41)
    r = 40.0
    out = run_see(gpu, W.plain_model(r), recs, lanes)
    p = out["p"][:, :3].astype(np.float32)
    i, j = candidate_pairs(p, r)
    i, j = i[box_ok(p, i, j, r)], j[box_ok(p, i, j, r)]
    d = p[j] - p[i]
    keep = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]) <= np.float32(r) * np.float32(r)
    i, j = i[keep], j[keep]
    assert np.array_equal(out["ncount"], np.bincount(i, minlength=len(p)))
    worst = check_sum(out["nsum"][:, :3], p[j].astype(np.float64), i, len(p), "nsum")
    print("plain nsum worst relative to sum|terms|: %.3g" % worst)


@pytest.mark.parametrize("lanes", [0, 4])
def test_reference_test_model_offsets_within_fp64_bound(gpu, lanes):
    recs = W.test_agents(20000, seed=42, clump=0.1)
    r = 50.0
    out = run_see(gpu, W.test_model(r), recs, lanes)
    p = out["p"][:, :3].astype(np.float32)
    i, j = candidate_pairs(p, r)
    ok = box_ok(p, i, j, r)
    i, j = i[ok], j[ok]
    assert np.array_equal(out["b_buffer_count"], np.bincount(i, minlength=len(p)))
    worst = check_sum(out["b_buffer"][:, :3], p[j].astype(np.float64) - p[i], i, len(p), "b_buffer")
    print("test-model offsets worst relative to sum|terms|: %.3g" % worst)


@pytest.mark.parametrize("lanes", [0, 4])
def test_flocking_sums_within_fp64_bound(gpu, lanes):
    recs = W.flock_agents(50000, seed=43, clusters=16, sigma=30.0, box=1000.0)
    r = 25.0
    out = run_see(gpu, W.flock_model(r), recs, lanes)
    p = out["p"][:, :3].astype(np.float32)
    v = out["v"][:, :3].astype(np.float32)
    i, j = candidate_pairs(p, r)
    ok = box_ok(p, i, j, r)
    i, j = i[ok], j[ok]
    d = p[j] - p[i]
    dl = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    front = (v[i, 0] * d[:, 0] + v[i, 1] * d[:, 1] + v[i, 2] * d[:, 2]) > np.float32(0)
    keep = (dl < np.float32(r)) & front
    i, j, d = i[keep], j[keep], d[keep]
    assert np.array_equal(out["cohesion_count"], np.bincount(i, minlength=len(p)))
    w1 = check_sum(out["alignment"][:, :3], v[j].astype(np.float64), i, len(p), "alignment")
    w2 = check_sum(out["cohesion"][:, :3], d.astype(np.float64), i, len(p), "cohesion")
    print("flocking worst relative to sum|terms|: alignment %.3g cohesion %.3g" % (w1, w2))


@pytest.mark.parametrize("lanes", [0, 4])
def test_sph_density_within_fp64_relative(gpu, lanes):
    recs, lo, hi = W.sph_particles(100000, h=1.0)
    ctx = W.sph_context(h=1.0, dt=1e-3, box_min=lo, box_max=hi)
    out = run_see(gpu, W.sph_model(ctx, 1.0), recs, lanes)
    p = out["p"][:, :3].astype(np.float32)
    h = 1.0
    i, j = candidate_pairs(p, h)
    ok = box_ok(p, i, j, h)
    i, j = i[ok], j[ok]
    # self included (self_see = 1)
    i = np.concatenate([i, np.arange(len(p))])
    j = np.concatenate([j, np.arange(len(p))])
    d = p[j] - p[i]
    q32 = np.float32(ctx["h2"][0]) - (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    keep = q32 > np.float32(0)
    q = q32[keep].astype(np.float64)
    terms = float(ctx["poly6"][0]) * q * q * q
    rho64 = np.zeros(len(p))
    np.add.at(rho64, i[keep], terms)
    err = np.abs(out["density"].astype(np.float64) - rho64) / rho64
    print("SPH density worst relative error: %.3g" % err.max())
    assert err.max() <= TOL


@pytest.mark.parametrize("lanes", [0, 4])
def test_full_size_flocking_positions_counts_and_offsets(gpu, lanes):
    """At the full 1M-agent flocking configuration (64 Gaussian clusters,
    sigma 30, box 2000, radius 25, cell size 25) with the reference's test
    model (b_buffer += p_b - p_a, count += 1; test/agent.c:2-8), one see
    pass: the pair counter equals the sum of the per-seer counts and is even
    (the box test is symmetric); for 2,000 seers drawn at random every count
    equals the k-d tree's float32 box-test count and every offset sum meets
    the FP64 bound."""
    from tay_amd import agents as A
    fl = W.flock_agents(1_000_000, seed=1)
    recs = np.zeros(len(fl), A.AGENT)
    recs["p"], recs["v"], recs["result_index"] = fl["p"], fl["v"], np.arange(len(fl))
    r = 25.0
    with W.Simulation(gpu, see_only(W.test_model(r, part=25.0)), recs, api.TAY_GPU_Z_GRID) as s:
        gpu.tay_set_pair_counting(s.state, 1)
        if lanes > 1:
            s.see_split(0, lanes)
            gpu.tay_set_lane_walk(s.state, 0)
        s.run(1)
        out = s.agents()
        pairs = gpu.pair_counts(s.state, 0)[1]
    counts = out["b_buffer_count"].astype(np.int64)
    assert int(counts.sum()) == pairs and pairs % 2 == 0
    p = out["p"][:, :3].astype(np.float32)
    sample = np.random.default_rng(5).choice(len(p), 2000, replace=False)
    near = cKDTree(p.astype(np.float64)).query_ball_point(p[sample].astype(np.float64), r * (1 + 1e-5), p=np.inf,
                                                          return_sorted=False)
    i = np.repeat(sample, [len(x) for x in near])
    j = np.concatenate([np.asarray(x, dtype=np.int64) for x in near])
    keep = (i != j) & box_ok(p, i, j, r)
    i, j = i[keep], j[keep]
    assert np.array_equal(counts[sample], np.bincount(i, minlength=len(p))[sample])
    rows = np.searchsorted(np.sort(sample), i)
    order = np.argsort(sample)
    worst = check_sum(out["b_buffer"][sample[order], :3], p[j].astype(np.float64) - p[i], rows, len(sample),
                      "b_buffer")
    print("1M flocking positions: %d pairs, sampled offsets worst relative to sum|terms|: %.3g" % (pairs, worst))
