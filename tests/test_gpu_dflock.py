"""Differentiable flocking on the GPU against the FP64 oracle (itself pinned
by FP64 finite differences in test_dflock_oracle.py).

Tolerances: the GPU computes in FP32 and walks neighbours in another order
than the oracle, so results agree to rounding, amplified over the steps of
the trajectory.  Forward positions: |p_gpu - p_64| <= 1e-4 x the distance
travelled.  Gradients: relative error of the weight gradient and norm-wise
relative error of the initial-velocity and initial-position gradients
<= 1e-4 when the FP64 adjoint runs along the GPU's own FP32 trajectory with the
same upstream gradient: that pins the adjoint kernels.  Against the FP64
model's own trajectory the weight gradient still agrees to 1e-4.  The
per-agent gradients differ from the FP64 model's by what the FP32 trajectory
itself implies: the FP64 adjoint run along the GPU's FP32 trajectory, with
the GPU's upstream gradient, differs from the FP64 model's gradient by the
same 2-3e-4 (50 steps), so the test asserts the GPU's end-to-end error
within that drift bound + 1e-4 (and below 1e-3).  A shorter trajectory does
not tighten this: the residual p_T - target shrinks with it faster than the
trajectories part (5 steps: 6e-4).  Central differences of the FP32
forward with the loss summed in float64: at eps 1e-3 every weight's
difference quotient is within 2e-4 of the gradient's norm (the separation
weight's own derivative is a hundred times smaller than the cohesion
weight's), at eps 3e-3 within 1e-3 of its own value."""
import numpy as np
import pytest
import torch

from tay_amd import api
from tay_amd import workloads as W
from tay_amd.dflock import DiffFlock

from adjoint_harness import simulate
from dflock_common import (DflockParams, backprop64, backprop_along64, clustered, forward64, gradient64,
                           load_oracle, mse)

pytestmark = pytest.mark.gpu

R, H2, DT, DAMP = 20.0, 1.0, 0.1, 0.05
PRM64 = DflockParams(R, H2, DT, DAMP)
HIDDEN = np.array([1.4, 0.7, 0.9])
GUESS = np.array([1.0, 1.0, 1.0], np.float32)


@pytest.fixture(scope="module")
def oracle64():
    return load_oracle()


def problem(n, steps, seed, clusters=16, box=500.0, sigma=30.0):
    p0, v0 = clustered(n, clusters, sigma, box, seed)
    return p0, v0


def target_of(lib64, p0, v0, steps):
    t, _ = forward64(lib64, p0, v0, HIDDEN, PRM64, steps)
    return t


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_forward_matches_fp64(gpu, oracle64):
    p0, v0 = problem(8192, 20, seed=31)
    print("case: dflock forward n=8192 steps=20", flush=True)
    with DiffFlock(gpu, len(p0), R, H2, DT, DAMP) as f:
        p, v = f.forward(p0, v0, GUESS, 20)
        assert f.max_neighbours() > 20
    p64, v64 = forward64(oracle64, p0, v0, GUESS, PRM64, 20)
    travelled = np.linalg.norm(p64 - p0, axis=1).mean()
    err = np.abs(p.astype(np.float64) - p64).max()
    assert err <= 1e-4 * travelled, (err, travelled)


def test_forward_matches_the_tay_model(gpu):
    """The dedicated kernels and smooth_see/smooth_act through tay.h compute
    the same model (to rounding: the Tay path sums neighbours in its storage
    order)."""
    p0, v0 = problem(4096, 5, seed=32, clusters=8, box=300.0)
    print("case: dflock forward vs TAY_GPU_Z_GRID smooth model", flush=True)
    with DiffFlock(gpu, len(p0), R, H2, DT, DAMP) as f:
        p, v = f.forward(p0, v0, GUESS, 5)
    rec = np.zeros(len(p0), W.A.SMOOTH_AGENT)
    rec["p"][:, :3] = p0
    rec["v"][:, :3] = v0
    rec["result_index"] = np.arange(len(p0))
    with W.Simulation(gpu, W.smooth_model(R, tuple(GUESS), H2, DT, DAMP), rec, api.TAY_GPU_Z_GRID) as s:
        s.run(5)
        out = s.agents()
    out = out[np.argsort(out["result_index"])]
    assert np.allclose(out["p"][:, :3], p, rtol=0, atol=1e-3)
    assert np.allclose(out["v"][:, :3], v, rtol=0, atol=1e-4)


def test_tay_path_trajectory_within_fp64_drift_bound(gpu, oracle64):
    """Trajectory drift of the Z-grid path itself: the smooth flocking model
    (C2 in every step) run through tay.h with TAY_GPU_Z_GRID for 20 steps
    stays within 1e-4 of the distance travelled of the FP64 trajectory of
    the same model, on every agent."""
    p0, v0 = problem(8192, 20, seed=35)
    rec = np.zeros(len(p0), W.A.SMOOTH_AGENT)
    rec["p"][:, :3] = p0
    rec["v"][:, :3] = v0
    rec["result_index"] = np.arange(len(p0))
    print("case: tay.h smooth model n=8192 steps=20 vs FP64", flush=True)
    with W.Simulation(gpu, W.smooth_model(R, tuple(GUESS), H2, DT, DAMP), rec, api.TAY_GPU_Z_GRID) as s:
        s.run(20)
        out = s.agents()
    out = out[np.argsort(out["result_index"])]
    p64, v64 = forward64(oracle64, p0, v0, GUESS, PRM64, 20)
    travelled = np.linalg.norm(p64 - p0, axis=1).mean()
    err = np.abs(out["p"][:, :3].astype(np.float64) - p64).max()
    assert err <= 1e-4 * travelled, (err, travelled)
    assert np.abs(out["v"][:, :3].astype(np.float64) - v64).max() <= 1e-4 * np.abs(v64).max()


def test_weight_free_motion_closed_form(gpu):
    p0, v0 = problem(2048, 10, seed=33)
    with DiffFlock(gpu, len(p0), R, H2, DT, DAMP) as f:
        p, v = f.forward(p0, v0, np.zeros(3, np.float32), 10)
        gp0, gv0, _ = f.backward(np.ones_like(p0))
    fct = 1.0 - DT * DAMP
    geo = sum(fct ** t for t in range(1, 11))
    assert np.allclose(v, v0 * fct ** 10, rtol=1e-5, atol=1e-6)
    assert np.allclose(p, p0 + DT * geo * v0, rtol=1e-6, atol=1e-4)
    assert np.allclose(gp0, 1.0, rtol=1e-6)
    assert np.allclose(gv0, DT * geo, rtol=1e-5)


def test_gradient_matches_fp64_oracle_full_size(gpu, oracle64):
    """65,536 agents in 16 clusters, 50 steps, checkpoints every 10: the
    weight gradient against the FP64 model's to 1e-4; the per-agent
    gradients within the bound the FP32 trajectory implies (the FP64
    adjoint along the GPU's own trajectory, same upstream gradient, against
    the FP64 model's gradient) + 1e-4."""
    n, steps = 65536, 50
    p0, v0 = problem(n, steps, seed=34)
    target = target_of(oracle64, p0, v0, steps)
    print("case: dflock gradient n=65536 steps=50 every=10 vs FP64", flush=True)
    with DiffFlock(gpu, n, R, H2, DT, DAMP, checkpoint_every=10) as f:
        sp, sv = [p0], [v0]
        for _ in range(steps):
            ps, vs = f.forward(sp[-1], sv[-1], GUESS, 1)
            sp.append(ps)
            sv.append(vs)
        p, _ = f.forward(p0, v0, GUESS, steps)
        assert np.array_equal(p, sp[-1])
        gp = (2.0 * (p.astype(np.float64) - target) / p.size).astype(np.float32)
        gp0, gv0, gw = f.backward(gp)
    loss64, gp64, gv64, gw64 = gradient64(oracle64, p0, v0, GUESS, PRM64, steps, target)
    assert abs(mse(p, target) - loss64) <= 1e-4 * loss64
    werr = np.abs(gw - gw64) / np.abs(gw64)
    print("end to end: weight gradient", gw, gw64, "rel", werr, "v0 rel", rel(gv0, gv64), "p0 rel", rel(gp0, gp64))
    assert (werr <= 1e-4).all(), (gw, gw64)
    ap, av, _ = backprop_along64(oracle64, np.stack(sp), np.stack(sv), GUESS, PRM64, gp)
    bound_v, bound_p = rel(av, gv64), rel(ap, gp64)
    print("drift bound (FP64 adjoint along the FP32 trajectory): v0 %.2e p0 %.2e" % (bound_v, bound_p))
    assert rel(gv0, gv64) <= bound_v + 1e-4 and rel(gv0, gv64) <= 1e-3
    assert rel(gp0, gp64) <= bound_p + 1e-4 and rel(gp0, gp64) <= 1e-3
    # the same upstream gradient, along the FP64 trajectory
    bp, bv, bw = backprop64(oracle64, p0, v0, GUESS, PRM64, steps, gp)
    print("same upstream: weight rel", np.abs(gw - bw) / np.abs(bw), "v0 rel", rel(gv0, bv), "p0 rel", rel(gp0, bp))
    assert rel(gv0, bv) <= 1e-3 and rel(gp0, bp) <= 1e-3


def test_adjoint_matches_fp64_along_the_gpu_trajectory(gpu, oracle64):
    """The FP64 adjoint along the GPU's own FP32 states (fetched one step at a
    time; a state round-trips through host memory exactly) and the GPU's
    checkpointed adjoint, same upstream gradient."""
    n, steps = 65536, 50
    p0, v0 = problem(n, steps, seed=34)
    target = target_of(oracle64, p0, v0, steps)
    print("case: dflock adjoint along the GPU trajectory n=65536 steps=50", flush=True)
    with DiffFlock(gpu, n, R, H2, DT, DAMP, checkpoint_every=10) as f:
        sp, sv = [p0], [v0]
        for _ in range(steps):
            p, v = f.forward(sp[-1], sv[-1], GUESS, 1)
            sp.append(p)
            sv.append(v)
        p, _ = f.forward(p0, v0, GUESS, steps)
        assert np.array_equal(p, sp[-1])
        gp = (2.0 * (p.astype(np.float64) - target) / p.size).astype(np.float32)
        gp0, gv0, gw = f.backward(gp)
    bp, bv, bw = backprop_along64(oracle64, np.stack(sp), np.stack(sv), GUESS, PRM64, gp)
    werr = np.abs(gw - bw) / np.abs(bw)
    print("along the GPU trajectory: weight rel", werr, "v0 rel", rel(gv0, bv), "p0 rel", rel(gp0, bp))
    assert (werr <= 1e-4).all()
    assert rel(gv0, bv) <= 1e-4
    assert rel(gp0, bp) <= 1e-4


def test_checkpointing_does_not_change_the_gradient(gpu, oracle64):
    n, steps = 8192, 24
    p0, v0 = problem(n, steps, seed=35)
    target = target_of(oracle64, p0, v0, steps)
    grads = []
    for every in (1, 5, 24):
        with DiffFlock(gpu, n, R, H2, DT, DAMP, checkpoint_every=every) as f:
            p, _ = f.forward(p0, v0, GUESS, steps)
            grads.append(f.backward((2.0 * (p - target) / p.size).astype(np.float32)))
    for g in grads[1:]:
        for a, b in zip(grads[0], g):
            assert np.array_equal(a, b)


def test_weight_gradient_matches_central_differences(gpu, oracle64):
    """Central differences of the GPU's FP32 forward, the loss summed in
    float64 (mse), against the GPU adjoint."""
    n, steps = 16384, 20
    p0, v0 = problem(n, steps, seed=36)
    target = target_of(oracle64, p0, v0, steps)
    print("case: dflock central differences n=16384 steps=20", flush=True)
    with DiffFlock(gpu, n, R, H2, DT, DAMP) as f:
        p, _ = f.forward(p0, v0, GUESS, steps)
        _, _, gw = f.backward((2.0 * (p.astype(np.float64) - target) / p.size).astype(np.float32))
        for eps, bound in ((1e-3, lambda fd: 2e-4 * np.linalg.norm(gw)), (3e-3, lambda fd: 1e-3 * abs(fd))):
            for k in range(3):
                wp, wm = GUESS.copy(), GUESS.copy()
                wp[k] += eps
                wm[k] -= eps
                fd = (mse(f.forward(p0, v0, wp, steps)[0], target) -
                      mse(f.forward(p0, v0, wm, steps)[0], target)) / (2 * eps)
                print("eps %g weight %d: adjoint %.8g difference quotient %.8g" % (eps, k, gw[k], fd))
                assert abs(gw[k] - fd) <= bound(fd), (eps, k, gw[k], fd)


def test_autograd_harness_matches_direct_call(gpu, oracle64):
    """loss.backward() through the autograd.Function gives exactly what the
    C-ABI backward gives for the upstream gradients torch computed."""
    n, steps = 4096, 10
    p0, v0 = problem(n, steps, seed=37)
    target = torch.from_numpy(target_of(oracle64, p0, v0, steps).astype(np.float32))

    def loss_of(p, v):
        return ((p - target) ** 2).mean() + 0.5 * (v ** 2).mean()

    with DiffFlock(gpu, n, R, H2, DT, DAMP) as f:
        tp0 = torch.from_numpy(p0).requires_grad_()
        tv0 = torch.from_numpy(v0).requires_grad_()
        tw = torch.from_numpy(GUESS.copy()).requires_grad_()
        p, v = simulate(f, tp0, tv0, tw, steps)
        loss_of(p, v).backward()
        pd, vd = f.forward(p0, v0, GUESS, steps)
        assert np.array_equal(pd, p.detach().numpy()) and np.array_equal(vd, v.detach().numpy())
        lp = torch.from_numpy(pd).requires_grad_()
        lv = torch.from_numpy(vd).requires_grad_()
        gp, gv = torch.autograd.grad(loss_of(lp, lv), (lp, lv))
        gp0, gv0, gw = f.backward(gp.numpy(), gv.numpy())
    assert np.array_equal(tw.grad.numpy(), gw)
    assert np.array_equal(tv0.grad.numpy(), gv0)
    assert np.array_equal(tp0.grad.numpy(), gp0)


def test_calibrates_weights_to_a_target_trajectory(gpu, oracle64):
    """Gradient descent on the three weights recovers the hidden ones."""
    n, steps = 16384, 30
    p0, v0 = problem(n, steps, seed=38)
    target = target_of(oracle64, p0, v0, steps)
    w = GUESS.astype(np.float64).copy()
    with DiffFlock(gpu, n, R, H2, DT, DAMP) as f:
        p, _ = f.forward(p0, v0, w.astype(np.float32), steps)
        loss0 = mse(p, target)
        m, s = np.zeros(3), np.zeros(3)
        for it in range(1, 201):
            p, _ = f.forward(p0, v0, w.astype(np.float32), steps)
            _, _, g = f.backward((2.0 * (p.astype(np.float64) - target) / p.size).astype(np.float32))
            m = 0.9 * m + 0.1 * g
            s = 0.999 * s + 0.001 * g * g
            lr = 0.05 if it <= 100 else 0.01
            w -= lr * (m / (1 - 0.9 ** it)) / (np.sqrt(s / (1 - 0.999 ** it)) + 1e-12)
        p, _ = f.forward(p0, v0, w.astype(np.float32), steps)
        loss1 = mse(p, target)
    print("calibration: loss %.4g -> %.4g, weights %s (hidden %s)" % (loss0, loss1, w, HIDDEN))
    assert loss1 < 1e-2 * loss0
    assert np.allclose(w, HIDDEN, rtol=0.05)


def brute_counts(p):
    """Per-agent neighbour counts: float32 offsets and squared lengths in
    the kernel's order, r^2 < R^2 strictly, the agent itself excluded."""
    p = np.asarray(p, np.float32)
    d = p[None, :, :] - p[:, None, :]
    r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    inside = r2 < np.float32(R) * np.float32(R)
    np.fill_diagonal(inside, False)
    return inside.sum(axis=1)


def test_neighbour_statistics_match_brute_force(gpu):
    """The pair total and the largest neighbour count of a forward run equal
    brute-force counts over the positions each step starts from."""
    p0, v0 = problem(3000, 2, seed=35, clusters=6, box=250.0)
    print("case: dflock neighbour statistics n=3000 steps=1,2", flush=True)
    with DiffFlock(gpu, len(p0), R, H2, DT, DAMP) as f:
        p1, _ = f.forward(p0, v0, GUESS, 1)
        c0 = brute_counts(p0)
        assert f.pair_count() == int(c0.sum()) > 0
        assert f.max_neighbours() == int(c0.max())
        f.forward(p0, v0, GUESS, 2)
        c1 = brute_counts(p1)
        assert f.pair_count() == int(c0.sum() + c1.sum())
        assert f.max_neighbours() == int(max(c0.max(), c1.max()))
