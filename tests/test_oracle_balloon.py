"""CPU checks that pin the Balloon-Windkessel restatement (oracle/balloon.py)
by what the mathematics fixes: published coefficient values, the closed-form
fixed point for constant input, the rest state, first-order agreement of
the non-linear and linear BOLD equations, and call chaining."""
import numpy as np
import pytest

from oracle import balloon


def test_coefficients_stephan2007():
    """Revised (Stephan 2007 Eq. 12) and classical (Obata 2004) values at the
    reference's defaults: nu_0 40.3 Hz, E0 0.4, TE 0.04 s, epsilon 0.5, r_0 25 Hz."""
    k1, k2, k3 = balloon.coefficients(balloon.DEFAULTS, rbm=True)
    assert k1 == pytest.approx(2.77264, rel=1e-12)
    assert k2 == pytest.approx(0.2, rel=1e-12)
    assert k3 == pytest.approx(0.5, rel=1e-12)
    assert balloon.coefficients(balloon.DEFAULTS, rbm=False) == pytest.approx((2.8, 0.8, 0.5), rel=1e-12)


def test_rest_state_is_a_fixed_point():
    x = np.zeros((3000, 5))
    y, st = balloon.evaluate(x, 1.0)
    assert np.max(np.abs(y)) < 1e-12
    np.testing.assert_allclose(st, np.array([[0.0] * 5, [1.0] * 5, [1.0] * 5, [1.0] * 5]), atol=1e-12)


@pytest.mark.parametrize("x0", [0.05, 0.3, 1.0])
def test_constant_input_converges_to_closed_form_fixed_point(x0):
    """A wrong sign, exponent or time constant in any of the four equations
    moves the fixed point or the convergence."""
    x = np.full((90000, 1), x0)                 # 90 s at 1 ms
    y, st = balloon.evaluate(x, 1.0)
    ref = balloon.steady_state(np.array([x0]))
    np.testing.assert_allclose(st, ref, rtol=1e-8, atol=1e-10)
    k1, k2, k3 = balloon.coefficients(balloon.DEFAULTS)
    _, _, v, q = ref
    yb = 4.0 * (k1 * (1 - q) + k2 * (1 - q / v) + k3 * (1 - v))
    assert y[-1, 0] == pytest.approx(float(yb[0]), rel=1e-7)


def test_linear_and_nonlinear_agree_to_first_order():
    t = np.arange(20000) * 1e-3
    x = 1e-4 * np.exp(-((t - 1.0) / 0.2) ** 2)[:, None]
    ynl, _ = balloon.evaluate(x, 1.0, nonlinear=True)
    yl, _ = balloon.evaluate(x, 1.0, nonlinear=False)
    assert np.max(np.abs(ynl - yl)) <= 1e-3 * np.max(np.abs(ynl))
    # the haemodynamic response to a brief input peaks a few seconds later
    tp = t[np.argmax(np.abs(ynl[:, 0]))]
    assert 2.0 < tp < 8.0


def test_chained_calls_equal_one_call():
    rng = np.random.default_rng(3)
    x = rng.uniform(0.0, 1.0, size=(5000, 7))
    y, st = balloon.evaluate(x, 1.0)
    y1, s1 = balloon.evaluate(x[:1234], 1.0)
    y2, s2 = balloon.evaluate(x[1234:], 1.0, state=s1)
    np.testing.assert_array_equal(np.concatenate([y1, y2]), y)
    np.testing.assert_array_equal(s2, st)
