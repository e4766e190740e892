"""Pins of the CPU restatement of the pose solver (tests/gn_reference.py) that the GPU solver
tests compare against: the se(3) exponential against the matrix exponential of the twist
(scipy), the logarithm as its inverse, the adjoint identity, the Huber pieces in closed form,
and the motion prior's gradient and Gauss-Newton matrix against finite differences of its
cost under the solver's left perturbation T <- exp(delta^) T.
"""
import numpy as np
import pytest
from scipy.linalg import expm

from gn_reference import (HUBER, adjoint, huber_rho, huber_weight, motion_prior, se3_exp, se3_log, skew)


def hat(xi):
    X = np.zeros((4, 4))
    X[:3, :3], X[:3, 3] = skew(xi[3:]), xi[:3]
    return X


def random_xi(rng, angle):
    axis = rng.normal(size=3)
    return np.concatenate([rng.normal(0, 0.3, 3), angle * axis / np.linalg.norm(axis)])


@pytest.mark.parametrize("angle", [0.0, 1e-8, 1e-6, 1e-3, 0.3, 2.0, 3.0])
def test_exp_is_the_matrix_exponential_and_log_inverts_it(angle):
    rng = np.random.default_rng(7)
    for _ in range(20):
        xi = random_xi(rng, angle)
        T = se3_exp(xi)
        assert np.abs(T - expm(hat(xi))).max() < 1e-12
        assert np.abs(se3_log(T) - xi).max() < 1e-9


def test_adjoint_moves_a_twist_through_a_pose():
    rng = np.random.default_rng(8)
    for _ in range(20):
        T = se3_exp(random_xi(rng, 1.0))
        xi = random_xi(rng, 0.4)
        lhs = se3_exp(adjoint(T) @ xi)
        rhs = T @ se3_exp(xi) @ np.linalg.inv(T)
        assert np.abs(lhs - rhs).max() < 1e-12


def test_huber_pieces():
    e = np.array([-5.0, -HUBER, -1.0, 0.0, 0.5, HUBER, 3.0])
    assert np.allclose(huber_rho(e), [HUBER * (5 - HUBER / 2), HUBER ** 2 / 2, 0.5, 0.0, 0.125, HUBER ** 2 / 2,
                                      HUBER * (3 - HUBER / 2)])
    assert np.allclose(huber_weight(e), [HUBER / 5, 1, 1, 1, 1, 1, HUBER / 3])
    # psi(e) = weight(e) e is the derivative of rho (central differences)
    x = np.linspace(-6, 6, 97)
    h = 1e-6
    assert np.abs((huber_rho(x + h) - huber_rho(x - h)) / (2 * h) - huber_weight(x) * x).max() < 1e-6


def _prior_cost(T_start, T, delta, st, sr):
    return motion_prior(T_start, se3_exp(delta) @ T, st, sr)[0]


@pytest.mark.parametrize("offset", [0.0, 1e-3])
def test_motion_prior_gradient_and_matrix_against_finite_differences(offset):
    """At the start pose the prior's b is zero and A is the exact Hessian of its cost in delta;
    a small offset (1e-3, a few noise sigmas) leaves b within 1e-3 (relative) of the gradient
    (the solver drops Jr^-1(xi) = I + O(xi))."""
    rng = np.random.default_rng(9)
    st, sr = 0.005, float(np.deg2rad(0.5))
    for _ in range(5):
        T_start = se3_exp(random_xi(rng, 1.5))
        T = se3_exp(random_xi(rng, 1.0) * offset) @ T_start if offset else T_start
        cost, A, b = motion_prior(T_start, T, st, sr)
        h = 1e-6
        g = np.zeros(6)
        for k in range(6):
            d = np.zeros(6)
            d[k] = h
            g[k] = (_prior_cost(T_start, T, d, st, sr) - _prior_cost(T_start, T, -d, st, sr)) / (2 * h)
        scale = np.abs(g).max() + np.abs(b).max() + 1e-30
        if offset == 0.0:
            assert cost < 1e-20 and np.abs(b).max() < 1e-8 and np.abs(g).max() < 1e-3
            H = np.zeros((6, 6))
            h2 = 1e-4
            for i in range(6):
                for j in range(6):
                    di, dj = np.zeros(6), np.zeros(6)
                    di[i], dj[j] = h2, h2
                    H[i, j] = (_prior_cost(T_start, T, di + dj, st, sr) - _prior_cost(T_start, T, di - dj, st, sr)
                               - _prior_cost(T_start, T, dj - di, st, sr)
                               + _prior_cost(T_start, T, -di - dj, st, sr)) / (4 * h2 * h2)
            assert np.abs(H - A).max() < 1e-3 * np.abs(A).max()
        else:
            assert cost > 0.0 and np.abs(g - b).max() < 1e-3 * scale


def test_motion_prior_weights_translation_and_rotation_by_their_sigmas():
    """A pure camera-frame translation of 1 sigma_t costs 1/2; a pure rotation of 2 sigma_rot
    about the camera centre costs 2; a sigma of 0 drops its half."""
    rng = np.random.default_rng(10)
    T_start = se3_exp(random_xi(rng, 1.0))
    st, sr = np.float32(0.002), np.float32(0.01)
    st64, sr64 = float(st), float(sr)
    xt = np.array([st64, 0, 0, 0, 0, 0])
    assert abs(motion_prior(T_start, T_start @ se3_exp(xt), st, sr)[0] - 0.5) < 1e-9
    xr = np.array([0, 0, 0, 0, 2 * sr64, 0])
    assert abs(motion_prior(T_start, T_start @ se3_exp(xr), st, sr)[0] - 2.0) < 1e-9
    assert motion_prior(T_start, T_start @ se3_exp(xr), st, 0.0)[0] < 1e-20
    assert motion_prior(T_start, T_start @ se3_exp(xt), 0.0, sr)[0] < 1e-20


def test_motion_prior_closed_form_equals_the_adjoint_products():
    """The closed form of M^T Lam M and M^T Lam xi equals the products with M = Ad(T^-1)."""
    rng = np.random.default_rng(11)
    st, sr = 0.003, 0.02
    lt, lr = 1 / float(np.float32(st)) ** 2, 1 / float(np.float32(sr)) ** 2
    lam = np.array([lt] * 3 + [lr] * 3)
    for _ in range(10):
        T_start = se3_exp(random_xi(rng, 2.0))
        T = se3_exp(random_xi(rng, 0.3) * 0.01) @ T_start
        _, A, b = motion_prior(T_start, T, st, sr)
        M = adjoint(np.linalg.inv(T))
        xi = se3_log(np.linalg.inv(T_start) @ T)
        assert np.abs(A - M.T @ (lam[:, None] * M)).max() < 1e-9 * np.abs(A).max()
        assert np.abs(b - M.T @ (lam * xi)).max() < 1e-9 * np.abs(b).max()
