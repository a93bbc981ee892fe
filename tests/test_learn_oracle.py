"""Pins the float64 restatement of the learned-start network (oracle/
learn_oracle.py) by what the mathematics fixes: the features' closed forms on
hand-computed states, the backward pass against central finite differences
of the forward pass, and the oracle's two-shock solver against the
reference's twoshock.F on the same interface states."""
import numpy as np
import pytest

import learn_oracle as lo
import oracle


def random_weights(rng):
    w = rng.normal(0.0, 0.4, lo.N_FLOATS)
    w[lo.OFF_MEAN:lo.OFF_INV_STD] = rng.normal(0.0, 0.5, lo.N_IN)
    w[lo.OFF_INV_STD:lo.OFF_GATE] = rng.uniform(0.5, 2.0, lo.N_IN)
    w[lo.OFF_GATE] = 0.0
    return w


def random_records(rng, n):
    r = np.zeros((n, 10))
    r[:, 0] = np.exp(rng.normal(0, 1, n))
    r[:, 1] = np.exp(rng.normal(0, 1, n))
    r[:, 2] = np.exp(rng.normal(0, 1, n))
    r[:, 3] = np.exp(rng.normal(0, 1, n))
    r[:, 4] = rng.normal(0, 1, n)
    r[:, 5] = rng.normal(0, 1, n)
    r[:, 6] = np.sqrt(r[:, 2] * r[:, 3]) * np.exp(rng.normal(0, 0.1, n))
    return r


def test_features_closed_form():
    # equal states: no jumps, equal sound speeds, acoustic start = p
    X, Y = lo.features([[2.0, 2.0, 3.0, 3.0, 0.5, 0.5, 3.0]], 5.0 / 3.0)
    assert np.array_equal(X[0], [0.0, 0.0, 0.0, 0.0, 0.5, 0.5]) and Y[0] == 0.0
    # rho 1 | 4, p 1 | 2, u 1 | 0, gamma 1.4: c_l = sqrt(1.4), c_r = sqrt(0.7),
    # so c_l / (c_l + c_r) = sqrt2 / (sqrt2 + 1) = 2 - sqrt2
    X, _ = lo.features([[1.0, 4.0, 1.0, 2.0, 1.0, 0.0, 1.0]], 1.4)
    expect = [0.6931471805599453, 1.3862943611198906, 0.8451542547285166, 1.1952286093343936,
              0.5857864376269049, 0.4142135623730951]
    assert np.allclose(X[0], expect, rtol=1e-14, atol=0)
    # rho = p = 1 both sides, u_l - u_r = 2 / sqrt(1.4): the Lagrangian wave
    # speeds are sqrt(1.4), so the acoustic start is 1 + 1 = 2; p* = 2e
    # gives the target 1
    u = 0.8451542547285166
    r = [[1.0, 1.0, 1.0, 1.0, u, -u, 2.0 * 2.718281828459045]]
    assert abs(lo.acoustic_start(r, 1.4)[0] - 2.0) < 1e-15
    assert abs(lo.features(r, 1.4)[1][0] - 1.0) < 1e-15
    # cavitation: the acoustic start is negative, the record is not learnable
    r = [[1.0, 1.0, 1.0, 1.0, -5.0, 5.0, 1e-20], [1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 1.0]]
    assert list(lo.learnable(r, 1.4)) == [False, True]


def test_forward_standardises_inputs():
    # with W1 = identity rows on x0 and unit later layers the output follows
    # tanh(tanh((x0 - mean0) * inv_std0))
    w = np.zeros(lo.N_FLOATS)
    w[lo.OFF_W1 + 0] = 1.0            # W1[0][0]
    w[lo.OFF_W2 + 0] = 1.0            # W2[0][0]
    w[lo.OFF_W3 + 0] = 1.0            # W3[0]
    w[lo.OFF_MEAN] = 0.25
    w[lo.OFF_INV_STD:lo.OFF_GATE] = 2.0
    X = np.array([[0.75, 9, 9, 9, 9, 9]])
    y, _ = lo.forward(w, X)
    assert abs(y[0] - np.tanh(np.tanh(1.0))) < 1e-15


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = random_weights(rng)
    r = random_records(rng, 64)
    X, Y = lo.features(r[lo.learnable(r, 5.0 / 3.0)], 5.0 / 3.0)
    g = lo.gradient(w, X, Y)
    h = 1e-6
    fd = np.zeros(lo.N_PARAMS)
    for p in range(lo.N_PARAMS):
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        fd[p] = (lo.loss(wp, X, Y) - lo.loss(wm, X, Y)) / (2 * h)
    scale = np.max(np.abs(fd))
    assert np.max(np.abs(g - fd)) < 1e-7 * scale, np.max(np.abs(g - fd)) / scale


def test_finite_differences_catch_a_dropped_term():
    """The check above fails when the backward pass drops the tanh
    derivative of the first layer (planted in the oracle, on the CPU)."""
    rng = np.random.default_rng(4)
    w = random_weights(rng)
    r = random_records(rng, 48)
    X, Y = lo.features(r[lo.learnable(r, 5.0 / 3.0)], 5.0 / 3.0)

    def broken_gradient(w, X, Y):
        W1, b1, W2, b2, W3, b3, mean, inv_std = lo.unpack(w)
        y, (xn, h1, h2) = lo.forward(w, X)
        g = 2.0 * (y - Y) / len(Y)
        d2 = (g[:, None] * W3[None, :]) * (1.0 - h2 ** 2)
        d1 = d2 @ W2                       # (1 - h1^2) dropped
        return np.concatenate([(d1.T @ xn).ravel(), d1.sum(0), (d2.T @ h1).ravel(), d2.sum(0),
                               h2.T @ g, [g.sum()]])

    g = broken_gradient(w, X, Y)
    p = lo.OFF_W1 + 3
    wp, wm = w.copy(), w.copy()
    wp[p] += 1e-6
    wm[p] -= 1e-6
    fd = (lo.loss(wp, X, Y) - lo.loss(wm, X, Y)) / 2e-6
    assert abs(g[p] - fd) > 1e-3 * abs(fd)


@pytest.mark.skipif(not oracle.available("ref"), reason="reference build absent")
def test_twoshock_port_matches_reference():
    rng = np.random.default_rng(5)
    r = random_records(rng, 3000)
    pa, ua = lo.twoshock("port", r, 5.0 / 3.0, 1e-20)
    pb, ub = lo.twoshock("ref", r, 5.0 / 3.0, 1e-20)
    assert np.array_equal(pa, pb) and np.array_equal(ua, ub)
