"""The FP64 oracle of the differentiable flocking model, pinned before it is
trusted as the reference gradient: its adjoint must agree with central
finite differences of its own forward model (FP64, so differences are
accurate to ~1e-9), and with the closed form of weight-free motion."""
import numpy as np
import pytest

from dflock_common import (DflockParams, backprop64, backprop_along64, clustered, forward64, gradient64,
                           load_oracle, mse)

PRM = DflockParams(10.0, 1.0, 0.1, 0.05)
W = np.array([1.2, 0.8, 0.6])


@pytest.fixture(scope="module")
def lib():
    return load_oracle()


@pytest.fixture(scope="module")
def case(lib):
    p0, v0 = clustered(300, 4, 6.0, 60.0, seed=21)
    target, _ = forward64(lib, p0, v0, [0.9, 1.1, 0.7], PRM, 8)
    target = target + np.random.default_rng(5).normal(scale=0.5, size=target.shape)
    return p0.astype(np.float64), v0.astype(np.float64), target


def loss_at(lib, p0, v0, w, target, steps=8):
    p, _ = forward64(lib, p0, v0, w, PRM, steps)
    return mse(p, target)


def test_case_has_neighbours(lib, case):
    p0, _, _ = case
    d = np.linalg.norm(p0[:, None, :] - p0[None, :, :], axis=2)
    per_agent = ((d < PRM.radius).sum(1) - 1)
    assert per_agent.mean() > 5


def test_loss_matches_forward(lib, case):
    p0, v0, target = case
    loss, _, _, _ = gradient64(lib, p0, v0, W, PRM, 8, target)
    assert abs(loss - loss_at(lib, p0, v0, W, target)) <= 1e-14 * max(1.0, loss)


def test_weight_gradient_matches_central_differences(lib, case):
    p0, v0, target = case
    _, _, _, gw = gradient64(lib, p0, v0, W, PRM, 8, target)
    for k in range(3):
        eps = 1e-5
        wp, wm = W.copy(), W.copy()
        wp[k] += eps
        wm[k] -= eps
        fd = (loss_at(lib, p0, v0, wp, target) - loss_at(lib, p0, v0, wm, target)) / (2 * eps)
        assert abs(gw[k] - fd) <= 1e-6 * abs(fd) + 1e-12, (k, gw[k], fd)


@pytest.mark.parametrize("which", ["v0", "p0"])
def test_state_gradient_matches_directional_differences(lib, case, which):
    p0, v0, target = case
    _, gp, gv, _ = gradient64(lib, p0, v0, W, PRM, 8, target)
    u = np.random.default_rng(9).normal(size=p0.shape)
    eps = 1e-5
    if which == "v0":
        fd = (loss_at(lib, p0, v0 + eps * u, W, target) - loss_at(lib, p0, v0 - eps * u, W, target)) / (2 * eps)
        ad = np.sum(gv * u)
    else:
        fd = (loss_at(lib, p0 + eps * u, v0, W, target) - loss_at(lib, p0 - eps * u, v0, W, target)) / (2 * eps)
        ad = np.sum(gp * u)
    assert abs(ad - fd) <= 1e-6 * abs(fd) + 1e-12, (ad, fd)


def test_weight_free_motion_closed_form(lib, case):
    p0, v0, target = case
    steps = 8
    p, v = forward64(lib, p0, v0, [0.0, 0.0, 0.0], PRM, steps)
    f = 1.0 - PRM.dt * PRM.damping
    assert np.allclose(v, v0 * f ** steps, rtol=1e-14, atol=0)
    geo = sum(f ** t for t in range(1, steps + 1))
    assert np.allclose(p, p0 + PRM.dt * geo * v0, rtol=1e-12, atol=1e-12)
    _, gp, gv, _ = gradient64(lib, p0, v0, [0.0, 0.0, 0.0], PRM, steps, target)
    lp = 2.0 * (p - target) / p.size
    assert np.allclose(gp, lp, rtol=1e-12, atol=1e-15)
    assert np.allclose(gv, lp * PRM.dt * geo, rtol=1e-12, atol=1e-15)


def test_backprop_of_the_mse_gradient_is_the_gradient(lib, case):
    p0, v0, target = case
    loss, gp, gv, gw = gradient64(lib, p0, v0, W, PRM, 8, target)
    p, _ = forward64(lib, p0, v0, W, PRM, 8)
    bp, bv, bw = backprop64(lib, p0, v0, W, PRM, 8, 2.0 * (p - target) / p.size)
    assert np.array_equal(bp, gp) and np.array_equal(bv, gv) and np.array_equal(bw, gw)


def test_velocity_loss_gradient_matches_differences(lib, case):
    """A loss of the final velocities, L = sum c . v_T: backprop with
    gv_final = c against central differences in the weights."""
    p0, v0, _ = case
    c = np.random.default_rng(12).normal(size=p0.shape)
    _, _, gw = backprop64(lib, p0, v0, W, PRM, 8, np.zeros_like(p0), c)
    for k in range(3):
        eps = 1e-5
        wp, wm = W.copy(), W.copy()
        wp[k] += eps
        wm[k] -= eps
        fd = (np.sum(c * forward64(lib, p0, v0, wp, PRM, 8)[1]) - np.sum(c * forward64(lib, p0, v0, wm, PRM, 8)[1])) / (2 * eps)
        assert abs(gw[k] - fd) <= 1e-6 * abs(fd) + 1e-12, (k, gw[k], fd)


def test_backprop_along_its_own_trajectory_is_backprop(lib, case):
    p0, v0, target = case
    sp, sv = [p0], [v0]
    for _ in range(8):
        p, v = forward64(lib, sp[-1], sv[-1], W, PRM, 1)
        sp.append(p)
        sv.append(v)
    g = 2.0 * (sp[-1] - target) / p0.size
    a = backprop_along64(lib, np.stack(sp), np.stack(sv), W, PRM, g)
    b = backprop64(lib, p0, v0, W, PRM, 8, g)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
