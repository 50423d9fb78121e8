"""Model functions on the GPU, one call per pair or agent (the harness of
tests/model_harness.py), bit for bit:

* the reference plug-ins' functions against the reference's own outputs
  (ref_models.npz, written by the reference code compiled from its sources):
  boid_see / boid_act (m_flocking/agent.c:2-79), sph_particle_density,
  sph_force_terms, sph_particle_leapfrog (m_sph/agent.c:40-111) and their
  dam-break equivalents sph_density / sph_forces, agent_see / agent_act /
  box_agent_see / box_agent_act (test/agent.c:2-119);
  sph_particle_pressure takes the device's double pow, so its result may
  differ from the host's by one unit in the last place of the float power;
* the workload variants (flock_*, plain_*, sph_pressure, sph_integrate,
  smooth_*, point_box_see, box_point_see) against the same functions built
  for the host, which tests/test_model_fixtures.py checks against float64
  restatements.

Seen groups run in a Z-grid (per-lane walk of sparse tiles) and in a Simple
space (the staged-chunk filter and compact loop)."""
import os

import numpy as np
import pytest

from tay_amd import agents as A
from tay_amd import api
from tay_amd import workloads as W

from conftest import GOLDEN
from model_harness import gpu_act, gpu_see_pairs, host_act, host_models, host_see, mismatches, site_pairs

pytestmark = pytest.mark.gpu

SEEN = {"zgrid": api.TAY_GPU_Z_GRID, "simple": api.TAY_GPU_SIMPLE}


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLDEN, "ref_models.npz"))


@pytest.fixture(scope="module")
def host():
    return host_models()


@pytest.mark.parametrize("seen", list(SEEN))
def test_boid_see_matches_m_flocking(gpu, ref, seen):
    print("case: boid_see per pair, seen %s" % seen, flush=True)
    got = gpu_see_pairs(gpu, "boid_see", ref["boid_see_seer"], ref["boid_see_seen"], ref["boid_see_ctx"], 25.0,
                        SEEN[seen])
    assert mismatches(got, ref["boid_see_out"]) == []


def test_boid_act_matches_m_flocking(gpu, ref):
    print("case: boid_act per agent", flush=True)
    got = gpu_act(gpu, "boid_act", ref["boid_act_in"], ref["boid_act_ctx"])
    assert mismatches(got, ref["boid_act_out"]) == []


@pytest.mark.parametrize("seen", list(SEEN))
@pytest.mark.parametrize("k", [0, 1], ids=["h1", "h_reference"])
def test_sph_see_functions_match_m_sph(gpu, ref, k, seen):
    p = "sph%d_" % k
    ctx, seer, s = ref[p + "ctx"], ref[p + "seer"], ref[p + "seen"]
    h = float(ctx["h"][0])
    for name, want in (("sph_particle_density", "density_out"), ("sph_density", "density_out"),
                       ("sph_force_terms", "forces_out"), ("sph_forces", "forces_out")):
        print("case: %s per pair, %s, seen %s" % (name, p, seen), flush=True)
        got = gpu_see_pairs(gpu, name, seer, s, ctx, h, SEEN[seen])
        assert mismatches(got, ref[p + want]) == [], name


@pytest.mark.parametrize("k", [0, 1], ids=["h1", "h_reference"])
def test_sph_act_functions_match_m_sph(gpu, ref, host, k):
    p = "sph%d_" % k
    ctx, act = ref[p + "ctx"], ref[p + "act_in"]
    print("case: sph act functions per agent, %s" % p, flush=True)
    assert mismatches(gpu_act(gpu, "sph_particle_leapfrog", act, ctx), ref[p + "leapfrog_out"]) == []
    got = gpu_act(gpu, "sph_particle_pressure", act, ctx)
    want = ref[p + "pressure_out"]
    assert mismatches(got, want, skip=("part_i", "cell_agent_i", "pressure")) == []
    # one unit in the last place of the float power, scaled by K
    x7 = np.float32(act["density"] / ctx["density"][0]).astype(np.float64) ** 7
    tol = float(ctx["K"][0]) * np.spacing(np.float32(x7)).astype(np.float64) * 1.01
    assert np.all(np.abs(got["pressure"].astype(np.float64) - want["pressure"]) <= tol + np.spacing(
        np.abs(want["pressure"])))
    for name in ("sph_pressure", "sph_integrate"):
        assert mismatches(gpu_act(gpu, name, act, ctx), host_act(host, name, act, ctx)) == [], name


@pytest.mark.parametrize("seen", list(SEEN))
def test_test_model_functions_match_reference(gpu, ref, seen):
    print("case: test model per pair, seen %s" % seen, flush=True)
    got = gpu_see_pairs(gpu, "agent_see", ref["agent_see_seer"], ref["agent_see_seen"], ref["see_ctx"], 25.0,
                        SEEN[seen])
    assert mismatches(got, ref["agent_see_out"]) == []
    if seen == "simple":
        got = gpu_see_pairs(gpu, "box_agent_see", ref["box_see_seer"], ref["box_see_seen"], ref["see_ctx"], 25.0,
                            SEEN[seen], seer_point=False, seen_point=False)
        assert mismatches(got, ref["box_see_out"]) == []
    assert mismatches(gpu_act(gpu, "agent_act", ref["agent_act_in"], ref["act_ctx"]), ref["agent_act_out"]) == []
    assert mismatches(gpu_act(gpu, "box_agent_act", ref["box_act_in"], ref["act_ctx"]), ref["box_act_out"]) == []


def _variant_cases():
    flock_see = np.zeros(1, A.FLOCK_SEE_CONTEXT)
    flock_see["r"], flock_see["separation_r"] = 25.0, 12.5
    flock_act = np.zeros(1, A.FLOCK_ACT_CONTEXT)
    flock_act["separation_w"], flock_act["alignment_w"], flock_act["cohesion_w"] = 1.5, 0.7, 1.3
    flock_act["dt"], flock_act["max_speed"] = 0.1, 2.0
    plain_see = np.zeros(1, A.PLAIN_SEE_CONTEXT)
    plain_see["r2"] = 1600.0
    plain_act = np.zeros(1, A.ACT_CONTEXT)
    plain_act["min"] = [-1000.0, -1000.0, -1000.0, 0.0]
    plain_act["max"] = [1000.0, 1000.0, 1000.0, 0.0]
    smooth = W.smooth_model(radius=20.0, weights=(1.2, 0.8, 0.9))
    return [("flock", A.FLOCK_AGENT, "flock_see", flock_see, "flock_act", flock_act, 25.0),
            ("plain", A.PLAIN_AGENT, "plain_see", plain_see, "plain_act", plain_act, 40.0),
            ("smooth", A.SMOOTH_AGENT, "smooth_see", smooth.passes[0].context, "smooth_act",
             smooth.passes[1].context, 20.0)]


@pytest.mark.parametrize("seen", list(SEEN))
@pytest.mark.parametrize("case", _variant_cases(), ids=lambda c: c[0])
def test_workload_variants_match_host_build(gpu, host, case, seen):
    name, dtype, see, see_ctx, act, act_ctx, radius = case
    a, b, g = site_pairs(dtype, 90, radius, radius)
    for f in dtype.names:
        if dtype[f].shape == (4,) and f not in ("p", "v"):
            a[f][:, :3] = g.normal(scale=3.0, size=(len(a), 3))
    print("case: %s per pair, seen %s" % (see, seen), flush=True)
    got = gpu_see_pairs(gpu, see, a, b, see_ctx, radius, SEEN[seen])
    want = host_see(host, see, a, b, see_ctx)
    assert mismatches(got, want) == []
    assert mismatches(gpu_act(gpu, act, got, act_ctx), host_act(host, act, want, act_ctx)) == []


@pytest.mark.parametrize("pair", ["point_box", "box_point"])
def test_mixed_point_box_pairs_match_host_build(gpu, host, pair):
    """A point group seeing a box group and the reverse (the PN and NP
    narrow phases) against the host build of the same models."""
    pa, _, g = site_pairs(A.AGENT, 91, 25.0, 25.0)
    bx, _, _ = site_pairs(A.BOX_AGENT, 92, 25.0, 25.0, pos="min")
    bx["min"][:, :3] = pa["p"][:, :3] + g.uniform(-40, 30, size=(len(pa), 3)).astype(np.float32)
    bx["max"][:, :3] = bx["min"][:, :3] + g.uniform(0.5, 10.0, size=(len(pa), 3)).astype(np.float32)
    ctx = np.zeros(1, A.SEE_CONTEXT)
    print("case: %s_see per pair" % pair, flush=True)
    if pair == "point_box":
        got = gpu_see_pairs(gpu, "point_box_see", pa, bx, ctx, 25.0, api.TAY_GPU_SIMPLE, seen_point=False)
        want = host_see(host, "point_box_see", pa, bx, ctx)
        # the pairs the PN box test passes: every axis within 25 of the box
        d_lo = pa["p"][:, :3] - bx["max"][:, :3]
        d_hi = bx["min"][:, :3] - pa["p"][:, :3]
    else:
        got = gpu_see_pairs(gpu, "box_point_see", bx, pa, ctx, 25.0, api.TAY_GPU_SIMPLE, seer_point=False)
        want = host_see(host, "box_point_see", bx, pa, ctx)
        d_lo = bx["min"][:, :3] - pa["p"][:, :3]
        d_hi = pa["p"][:, :3] - bx["max"][:, :3]
    passes = ~((d_lo > 25.0) | (d_hi > 25.0)).any(axis=1)
    assert passes.sum() > 100 and (~passes).sum() > 100
    # the host call runs every pair; only passing pairs may change on the GPU
    src = pa if pair == "point_box" else bx
    want = np.where(passes, want, src)
    assert mismatches(got, want) == []
