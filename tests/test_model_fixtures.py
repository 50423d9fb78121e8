"""Model functions against independent checks (CPU).

1. The reference plug-ins' own model functions, compiled from the reference
   sources (tests/golden/make_golden.py, ref_models.npz), against this
   repository's re-expressions of them, built as host callbacks: bit for bit
   per pair and per agent.  m_flocking/agent.c:2-79 (boid_see, boid_act),
   m_sph/agent.c:40-101 (sph_particle_density = sph_density, sph_force_terms
   = sph_forces, sph_particle_pressure, sph_particle_leapfrog,
   sph_particle_reset), test/agent.c:2-119 (agent_see, agent_act,
   box_agent_see, box_agent_act).  The inputs reach every branch; the tests
   assert that they do.

2. The workload variants that have no reference counterpart (flock_see /
   flock_act of the 1M flocking workload, plain_see / plain_act of the 10k
   plain workload, sph_pressure / sph_integrate of the dam break,
   smooth_see / smooth_act of the differentiable flocking, point_box_see /
   box_point_see) against float64 restatements written from the workloads'
   definitions (the behaviour stated in the models' headers), not from the C
   sources: every output within a few float32 units of the last place of the
   float64 value, so a wrong sign, weight, branch or clamp fails.
"""
import os

import numpy as np
import pytest

from tay_amd import agents as A
from tay_amd import workloads as W

from conftest import GOLDEN
from model_harness import host_act, host_models, host_see, mismatches


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLDEN, "ref_models.npz"))


@pytest.fixture(scope="module")
def lib():
    return host_models()


# ----------------------------------------------------- reference fixtures

def test_boid_see_matches_m_flocking(ref, lib):
    seer, seen, ctx = ref["boid_see_seer"], ref["boid_see_seen"], ref["boid_see_ctx"]
    want = ref["boid_see_out"]
    assert mismatches(host_see(lib, "boid_see", seer, seen, ctx), want) == []
    # every branch: outside r, behind the boid, in front (cohesion), closer
    # than separation_r, and normalize_to's near-zero offset
    dl = np.linalg.norm(seen["p"][:, :3].astype(np.float64) - seer["p"][:, :3], axis=1)
    front = (want["cohesion_count"] != seer["cohesion_count"])
    sep = (want["separation_count"] != seer["separation_count"])
    assert (dl >= ctx["r"][0]).sum() > 50 and front.sum() > 100 and sep.sum() > 50
    assert ((dl < ctx["r"][0]) & ~front).sum() > 50
    assert (sep & (dl < 1e-6)).sum() >= 1


def test_boid_act_matches_m_flocking(ref, lib):
    a, ctx, want = ref["boid_act_in"], ref["boid_act_ctx"], ref["boid_act_out"]
    assert mismatches(host_act(lib, "boid_act", a, ctx), want) == []
    far = np.linalg.norm(a["p"][:, :3].astype(np.float64), axis=1) > 400
    assert far.sum() > 50 and (~far).sum() > 50
    assert (a["cohesion_count"] == 0).sum() > 50 and (a["separation_count"] == 0).sum() > 50
    assert (want["dir"][:, 0] == 1.0).sum() >= 16     # normalize's near-zero branch


@pytest.mark.parametrize("k", [0, 1], ids=["h1", "h_reference"])
def test_sph_functions_match_m_sph(ref, lib, k):
    p = "sph%d_" % k
    ctx, seer, seen, act = ref[p + "ctx"], ref[p + "seer"], ref[p + "seen"], ref[p + "act_in"]
    for name in ("sph_particle_density", "sph_density"):
        assert mismatches(host_see(lib, name, seer, seen, ctx), ref[p + "density_out"]) == [], name
    for name in ("sph_force_terms", "sph_forces"):
        assert mismatches(host_see(lib, name, seer, seen, ctx), ref[p + "forces_out"]) == [], name
    assert mismatches(host_act(lib, "sph_particle_pressure", act, ctx), ref[p + "pressure_out"]) == []
    assert mismatches(host_act(lib, "sph_particle_leapfrog", act, ctx), ref[p + "leapfrog_out"]) == []
    assert mismatches(host_act(lib, "sph_particle_reset", act), ref[p + "reset_out"]) == []
    # branches: inside and outside h, coincident pairs (r < 1e-5), walls
    h = float(ctx["h"][0])
    r = np.linalg.norm(seen["p"][:, :3].astype(np.float64) - seer["p"][:, :3], axis=1)
    assert (r < h).sum() > 100 and (r >= h).sum() > 50 and (r < 1e-5).sum() >= 32
    lo, hi = ctx["min"][0][:3], ctx["max"][0][:3]
    moved = ref[p + "leapfrog_out"]["p"][:, :3]
    assert (moved == lo).any(axis=1).sum() > 20 and (moved == hi).any(axis=1).sum() > 20


@pytest.mark.parametrize("k", [0, 1], ids=["h1", "h_reference"])
def test_sph_pressure_power_within_pow(ref, lib, k):
    """The dam-break pressure takes (rho / rho0)^7 by three multiplies
    instead of libm pow; against the reference's pow it may differ only by
    the rounding of those multiplies: |p - p_ref| <= 1e-6 K max(x^7, 1)."""
    p = "sph%d_" % k
    ctx, act = ref[p + "ctx"], ref[p + "act_in"]
    got = host_act(lib, "sph_pressure", act, ctx)["pressure"].astype(np.float64)
    want = ref[p + "pressure_out"]["pressure"].astype(np.float64)
    x7 = (act["density"].astype(np.float64) / float(ctx["density"][0])) ** 7
    assert np.all(np.abs(got - want) <= 1e-6 * float(ctx["K"][0]) * np.maximum(x7, 1.0))


def test_test_model_functions_match_reference(ref, lib):
    assert mismatches(host_see(lib, "agent_see", ref["agent_see_seer"], ref["agent_see_seen"], ref["see_ctx"]),
                      ref["agent_see_out"]) == []
    assert mismatches(host_see(lib, "box_agent_see", ref["box_see_seer"], ref["box_see_seen"], ref["see_ctx"]),
                      ref["box_see_out"]) == []
    assert mismatches(host_act(lib, "agent_act", ref["agent_act_in"], ref["act_ctx"]), ref["agent_act_out"]) == []
    assert mismatches(host_act(lib, "box_agent_act", ref["box_act_in"], ref["act_ctx"]), ref["box_act_out"]) == []
    # walls on both sides moved boxes and reflected velocities
    a, out = ref["box_act_in"], ref["box_act_out"]
    assert (np.sign(out["v"][:, :3]) != np.sign(a["v"][:, :3])).any(axis=1).sum() > 20


# ------------------------------------------------- float64 restatements

def close(got, want, scale, ulps=8):
    """|got - want| within `ulps` float32 units of the last place of the
    magnitude `scale` (elementwise)."""
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    tol = ulps * np.spacing(np.abs(np.asarray(scale, np.float32))).astype(np.float64) + 1e-30
    return np.abs(got - want) <= tol


def pair_inputs(dtype, n, seed, radius, box):
    """Seer/seen records of `dtype` with position p and velocity v: seen at
    offsets inside the per-axis box, half of them inside the radius."""
    g = W.rng(seed)
    a = np.zeros(n, dtype)
    b = np.zeros(n, dtype)
    a["p"][:, :3] = g.uniform(-100, 100, size=(n, 3))
    off = np.where(g.uniform(size=(n, 1)) < 0.5, g.uniform(-box, box, size=(n, 3)),
                   g.normal(size=(n, 3)) * radius / 3.0)
    b["p"][:, :3] = (a["p"][:, :3] + off).astype(np.float32)
    a["v"][:, :3] = g.normal(size=(n, 3))
    b["v"][:, :3] = g.normal(size=(n, 3))
    return a, b, g


def f64(x):
    return np.asarray(x, np.float64)


def test_flock_see_restated(lib):
    """flock_see: within Euclidean distance r and in front of the agent
    (v . d > 0, d = p_b - p_a): alignment += v_b, cohesion += d, count += 1;
    closer than separation_r also separation -= d / |d| (separation_r - |d|),
    separation count += 1."""
    n, r, sep_r = 4000, 25.0, 12.5
    a, b, g = pair_inputs(A.FLOCK_AGENT, n, 81, r, 25.0)
    for f in ("separation", "alignment", "cohesion"):
        a[f][:, :3] = g.normal(scale=3.0, size=(n, 3))
    a["separation_count"] = g.integers(0, 9, n)
    a["cohesion_count"] = g.integers(0, 9, n)
    ctx = np.zeros(1, A.FLOCK_SEE_CONTEXT)
    ctx["r"], ctx["separation_r"] = r, sep_r
    got = host_see(lib, "flock_see", a, b, ctx)

    d = f64(b["p"][:, :3]) - f64(a["p"][:, :3])
    dl = np.linalg.norm(d, axis=1)
    dot = (f64(a["v"][:, :3]) * d).sum(1)
    # keep clear of the decision boundaries, where float32 and float64 may
    # round to different sides
    clear = (np.abs(dl - r) > 1e-3) & (np.abs(dl - sep_r) > 1e-3) & (np.abs(dot) > 1e-3)
    seen_ = (dl < r) & (dot > 0)
    sep = seen_ & (dl < sep_r)
    assert clear.sum() > 0.95 * n and seen_.sum() > 500 and sep.sum() > 200
    ali = f64(a["alignment"][:, :3]) + np.where(seen_[:, None], f64(b["v"][:, :3]), 0.0)
    coh = f64(a["cohesion"][:, :3]) + np.where(seen_[:, None], d, 0.0)
    push = d / np.maximum(dl, 1e-30)[:, None] * (sep_r - dl)[:, None]
    sepv = f64(a["separation"][:, :3]) - np.where(sep[:, None], push, 0.0)
    c = clear
    assert np.array_equal(got["cohesion_count"][c], (a["cohesion_count"] + seen_)[c])
    assert np.array_equal(got["separation_count"][c], (a["separation_count"] + sep)[c])
    scale = np.abs(f64(a["alignment"][:, :3])) + np.abs(f64(b["v"][:, :3]))
    assert close(got["alignment"][:, :3], ali, scale)[c].all()
    scale = np.abs(f64(a["cohesion"][:, :3])) + np.abs(d) + np.abs(f64(a["p"][:, :3]))
    assert close(got["cohesion"][:, :3], coh, scale)[c].all()
    scale = np.abs(f64(a["separation"][:, :3])) + sep_r
    assert close(got["separation"][:, :3], sepv, scale, ulps=16)[c].all()


def test_flock_act_restated(lib):
    """flock_act: acc = w_sep separation / n_sep (when n_sep > 0) +
    (w_ali alignment + w_coh cohesion) / n (when n > 0); v += acc dt;
    |v| clamped to max_speed; p += v dt; accumulators and counts cleared."""
    n = 4000
    g = W.rng(82)
    a = np.zeros(n, A.FLOCK_AGENT)
    a["p"][:, :3] = g.uniform(-500, 500, size=(n, 3))
    a["v"][:, :3] = g.normal(scale=1.2, size=(n, 3))
    for f in ("separation", "alignment", "cohesion"):
        a[f][:, :3] = g.normal(scale=20.0, size=(n, 3))
    a["separation_count"] = g.integers(0, 6, n)
    a["cohesion_count"] = g.integers(0, 20, n)
    ctx = np.zeros(1, A.FLOCK_ACT_CONTEXT)
    # three different weights (the workload's are 1.5 / 1.0 / 1.0), so a
    # weight applied to the wrong sum shows
    ws, wa, wc, dt, vmax = 1.5, 0.7, 1.3, 0.1, 2.0
    ctx["separation_w"], ctx["alignment_w"], ctx["cohesion_w"], ctx["dt"], ctx["max_speed"] = ws, wa, wc, dt, vmax
    got = host_act(lib, "flock_act", a, ctx)

    ns, nc = f64(a["separation_count"])[:, None], f64(a["cohesion_count"])[:, None]
    acc = np.where(ns > 0, ws * f64(a["separation"][:, :3]) / np.maximum(ns, 1), 0.0)
    acc += np.where(nc > 0, (wa * f64(a["alignment"][:, :3]) + wc * f64(a["cohesion"][:, :3])) / np.maximum(nc, 1),
                    0.0)
    v = f64(a["v"][:, :3]) + acc * dt
    speed = np.linalg.norm(v, axis=1)
    clamp = speed > vmax
    v = np.where(clamp[:, None], v * (vmax / speed)[:, None], v)
    p = f64(a["p"][:, :3]) + v * dt
    clear = np.abs(speed - vmax) > 1e-4
    assert clamp.sum() > 500 and (~clamp).sum() > 500
    vscale = np.abs(f64(a["v"][:, :3])) + np.abs(acc * dt) + vmax
    assert close(got["v"][:, :3], v, vscale, ulps=16)[clear].all()
    assert close(got["p"][:, :3], p, np.abs(f64(a["p"][:, :3])) + 1.0, ulps=8)[clear].all()
    for f in ("separation", "alignment", "cohesion"):
        assert not got[f][:, :3].any()
    assert not got["separation_count"].any() and not got["cohesion_count"].any()


def test_plain_functions_restated(lib):
    """plain_see: a neighbour within Euclidean distance r (d^2 <= r^2)
    adds its position to nsum and 1 to ncount.  plain_act: the finished
    sums move to last_nsum / last_ncount, the accumulators clear, p += v,
    and a coordinate beyond [min, max] is put on the wall with its velocity
    component reversed."""
    n, r = 4000, 40.0
    a, b, g = pair_inputs(A.PLAIN_AGENT, n, 83, r, 40.0)
    a["nsum"][:, :3] = g.normal(scale=100.0, size=(n, 3))
    a["ncount"] = g.integers(0, 50, n)
    ctx = np.zeros(1, A.PLAIN_SEE_CONTEXT)
    ctx["r2"] = np.float32(r) * np.float32(r)
    got = host_see(lib, "plain_see", a, b, ctx)
    d2 = ((f64(b["p"][:, :3]) - f64(a["p"][:, :3])) ** 2).sum(1)
    clear = np.abs(d2 - r * r) > 1e-2
    inside = d2 <= r * r
    assert inside.sum() > 500 and (~inside).sum() > 500
    assert np.array_equal(got["ncount"][clear], (a["ncount"] + inside)[clear])
    want = f64(a["nsum"][:, :3]) + np.where(inside[:, None], f64(b["p"][:, :3]), 0.0)
    scale = np.abs(f64(a["nsum"][:, :3])) + np.abs(f64(b["p"][:, :3]))
    assert close(got["nsum"][:, :3], want, scale, ulps=1)[clear].all()

    act = got.copy()
    act["p"][:, :3] = g.uniform(-10, 1010, size=(n, 3))
    act["v"][:, :3] = g.normal(scale=8.0, size=(n, 3))
    actx = np.zeros(1, A.ACT_CONTEXT)
    actx["max"] = [1000.0, 1000.0, 1000.0, 0.0]
    out = host_act(lib, "plain_act", act, actx)
    p = f64(act["p"][:, :3]) + f64(act["v"][:, :3])
    v = f64(act["v"][:, :3]).copy()
    lo_hit, hi_hit = p < 0.0, p > 1000.0
    p = np.clip(p, 0.0, 1000.0)
    v = np.where(lo_hit | hi_hit, -v, v)
    assert lo_hit.any(axis=1).sum() > 20 and hi_hit.any(axis=1).sum() > 20
    assert close(out["p"][:, :3], p, np.abs(p) + 1.0, ulps=1).all()
    assert np.array_equal(out["v"][:, :3], v.astype(np.float32))
    assert np.array_equal(out["last_ncount"], act["ncount"])
    assert np.array_equal(out["last_nsum"][:, :3], act["nsum"][:, :3])
    assert not out["ncount"].any() and not out["nsum"][:, :3].any()


def test_sph_integrate_restated(lib):
    """sph_integrate: f = pressure_accum + dynamic_viscosity viscosity_accum,
    a = f / density, a_z -= 9.81; symplectic Euler v += a dt, vh = v,
    p += v dt; a coordinate beyond [min, max] goes on the wall and that
    axis's v and vh are multiplied by -0.5; density and both accumulators
    are cleared."""
    n = 4000
    g = W.rng(84)
    lo, hi = np.array([0, 0, 0, 0], np.float32), np.array([10, 12, 14, 0], np.float32)
    ctx = W.sph_context(h=1.0, dt=1e-3, box_min=lo, box_max=hi)
    a = np.zeros(n, A.SPH_PARTICLE)
    a["p"][:, :3] = g.uniform(-0.05, 1.05, size=(n, 3)) * hi[:3]
    a["v"][:, :3] = g.normal(scale=30.0, size=(n, 3))
    a["vh"][:, :3] = g.normal(size=(n, 3))
    a["pressure_accum"][:, :3] = g.normal(scale=5e4, size=(n, 3))
    a["viscosity_accum"][:, :3] = g.normal(scale=5e3, size=(n, 3))
    a["density"] = g.uniform(500.0, 2000.0, n)
    got = host_act(lib, "sph_integrate", a, ctx)

    mu, dt = float(ctx["dynamic_viscosity"][0]), float(ctx["dt"][0])
    f = f64(a["pressure_accum"][:, :3]) + mu * f64(a["viscosity_accum"][:, :3])
    acc = f / f64(a["density"])[:, None]
    acc[:, 2] -= 9.81
    v = f64(a["v"][:, :3]) + acc * dt
    p = f64(a["p"][:, :3]) + v * dt
    out_lo, out_hi = p < lo[:3], p > hi[:3]
    wall = out_lo | out_hi
    p = np.clip(p, lo[:3], hi[:3])
    v = np.where(wall, -0.5 * v, v)
    assert out_lo.any(axis=1).sum() > 50 and out_hi.any(axis=1).sum() > 50
    clear = (np.abs(f64(a["p"][:, :3]) + (f64(a["v"][:, :3]) + acc * dt) * dt - np.clip(
        f64(a["p"][:, :3]), lo[:3], hi[:3])) > 1e-3).all(axis=1) | ~wall.any(axis=1)
    vscale = np.abs(f64(a["v"][:, :3])) + np.abs(acc * dt) + 1.0
    assert close(got["v"][:, :3], v, vscale, ulps=16)[clear].all()
    assert close(got["vh"][:, :3], v, vscale, ulps=16)[clear].all()
    assert close(got["p"][:, :3], p, np.abs(p) + np.abs(v * dt) + 1e-3, ulps=8)[clear].all()
    assert not got["density"].any()
    assert not got["pressure_accum"][:, :3].any() and not got["viscosity_accum"][:, :3].any()


def test_smooth_functions_restated(lib):
    """smooth_see: for r^2 = |d|^2 < R^2 (d = p_b - p_a), w = (1 - r^2/R^2)^3;
    weight_sum += w, alignment += w (v_b - v_a), cohesion += w d,
    separation -= w d / (r^2 + h2).  smooth_act: a = (w_s S + w_a A + w_c C)
    / (weight_sum + 1), v += dt (a - damping v), p += dt v, sums cleared."""
    n, R, h2 = 4000, 20.0, 1.0
    a, b, g = pair_inputs(A.SMOOTH_AGENT, n, 85, R, 20.0)
    for f in ("separation", "alignment", "cohesion"):
        a[f][:, :3] = g.normal(scale=2.0, size=(n, 3))
    a["weight_sum"] = g.uniform(0, 5, n)
    m = W.smooth_model(radius=R, h2=h2)
    see_ctx, act_ctx = m.passes[0].context, m.passes[1].context
    got = host_see(lib, "smooth_see", a, b, see_ctx)
    d = f64(b["p"][:, :3]) - f64(a["p"][:, :3])
    r2 = (d ** 2).sum(1)
    inside = r2 < R * R
    clear = np.abs(r2 - R * R) > 1e-2
    w = np.where(inside, (1.0 - r2 / (R * R)) ** 3, 0.0)
    assert inside.sum() > 500 and (~inside).sum() > 500
    want_w = f64(a["weight_sum"]) + w
    assert close(got["weight_sum"], want_w, np.abs(want_w) + 1.0, ulps=16)[clear].all()
    dv = f64(b["v"][:, :3]) - f64(a["v"][:, :3])
    for name, term, scale in (("alignment", w[:, None] * dv, np.abs(dv) + 2.0),
                              ("cohesion", w[:, None] * d, np.abs(d) + 2.0),
                              ("separation", -w[:, None] * d / (r2 + h2)[:, None], 2.0 + np.abs(d))):
        want = f64(a[name][:, :3]) + term
        assert close(got[name][:, :3], want, np.abs(f64(a[name][:, :3])) + scale, ulps=64)[clear].all(), name

    act = got.copy()
    out = host_act(lib, "smooth_act", act, act_ctx)
    ws, wa, wc = (float(act_ctx[k][0]) for k in ("separation_w", "alignment_w", "cohesion_w"))
    dt, damping = float(act_ctx["dt"][0]), float(act_ctx["damping"][0])
    acc = (ws * f64(act["separation"][:, :3]) + wa * f64(act["alignment"][:, :3]) +
           wc * f64(act["cohesion"][:, :3])) / (f64(act["weight_sum"]) + 1.0)[:, None]
    v = f64(act["v"][:, :3]) + dt * (acc - damping * f64(act["v"][:, :3]))
    p = f64(act["p"][:, :3]) + dt * v
    scale = np.abs(f64(act["v"][:, :3])) + dt * (np.abs(acc) + 1.0)
    assert close(out["v"][:, :3], v, scale, ulps=16).all()
    assert close(out["p"][:, :3], p, np.abs(p) + 1.0, ulps=8).all()
    assert not out["weight_sum"].any() and not out["separation"][:, :3].any()


def test_mixed_point_box_see_restated(lib):
    """point_box_see: b_buffer += seen min corner - seer position, count += 1;
    box_point_see: b_buffer += seen position - seer min corner, count += 1
    (one float32 subtraction and one addition per axis: exact in numpy)."""
    g = W.rng(86)
    n = 1000
    pa = W.test_agents(n, seed=87)
    pa["b_buffer"][:, :3] = g.normal(scale=10.0, size=(n, 3))
    pa["b_buffer_count"] = g.integers(0, 9, n)
    bx = W.box_agents(n, seed=88)
    bx["b_buffer"][:, :3] = g.normal(scale=10.0, size=(n, 3))
    bx["b_buffer_count"] = g.integers(0, 9, n)
    ctx = np.zeros(1, A.SEE_CONTEXT)
    got = host_see(lib, "point_box_see", pa, bx, ctx)
    want = pa["b_buffer"][:, :3] + (bx["min"][:, :3] - pa["p"][:, :3])
    assert np.array_equal(got["b_buffer"][:, :3], want)
    assert np.array_equal(got["b_buffer_count"], pa["b_buffer_count"] + 1)
    got = host_see(lib, "box_point_see", bx, pa, ctx)
    want = bx["b_buffer"][:, :3] + (pa["p"][:, :3] - bx["min"][:, :3])
    assert np.array_equal(got["b_buffer"][:, :3], want)
    assert np.array_equal(got["b_buffer_count"], bx["b_buffer_count"] + 1)
