# SPDX-License-Identifier: MIT
"""Pins of the oracle's yield-curve rates, non-uniform time domain and
discrete-dividend events (CPU, no GPU): the reference's own known answers
(tests/yield_curve_test.cc, tests/discrete_dividend_event_test.cc,
tests/american_option_test.cc:183-221), closed forms the extensions reduce
to, and a European limit computed by quadrature that fixes the direction of
the curve's time axis and the timing of a dividend."""
import math

import numpy as np
import pytest
from scipy import integrate, stats

import pyoracle as O

CURVE_5_6 = [(0.0, 0.0), (1.0, -0.05), (2.0, -0.11)]  # yield_curve_test.cc:72-93


def test_curve_rate_interpolation_and_extrapolation():  # yield_curve_test.cc:72-101
    ok, r, ld, _, _ = O.curve_eval(CURVE_5_6, [0.0, 0.5, 0.99, 1.0, 1.5, 2.0, 3.0])
    assert ok
    np.testing.assert_allclose(r, [0.05, 0.05, 0.05, 0.06, 0.06, 0.06, 0.06], atol=1e-10)
    # beyond the last tenor ln D continues with the last forward rate
    assert abs(ld[-1] - (-0.11 - 0.06 * 1.0)) < 1e-12
    ok, r, _, d, _ = O.curve_eval([(0.0, 0.0), (100.0, -5.0)], [0.0, 0.5, 1.0, 10.0, 50.0, 100.0, 2.0])
    assert ok
    np.testing.assert_allclose(r, 0.05, rtol=0, atol=1e-15)
    assert abs(d[2] - math.exp(-0.05)) < 1e-10 and abs(d[6] - math.exp(-0.10)) < 1e-10


def test_curve_log_linear_discount():  # yield_curve_test.cc:103-119
    ok, _, _, d, _ = O.curve_eval([(0.0, 0.0), (1.0, math.log(0.95)), (2.0, math.log(0.90))], [0.5, 0.0])
    assert ok
    assert abs(d[0] - math.exp(0.5 * math.log(0.95))) < 1e-12
    assert d[1] == 1.0


@pytest.mark.parametrize("points, msg", [
    ([(0.5, 0.0), (1.0, -0.05)], "t=0"),                        # :40-49
    ([(0.0, 0.01), (1.0, -0.05)], "log_discount at t=0"),      # :139-148
    ([(0.0, 0.0), (1.0, -0.05), (1.0, -0.06)], "strictly increasing"),   # :156-167
    ([(0.0, 0.0), (1.0, -0.05), (1.0 + 1e-11, -0.06)], "strictly increasing"),  # :170-181
    ([], "Empty"),
])
def test_curve_validation(points, msg):
    ok, _, _, _, why = O.curve_eval(points, [0.5])
    assert not ok and msg in why


def test_curve_points_are_sorted():  # yield_curve_test.cc:184-196
    ok, _, _, d, _ = O.curve_eval([(2.0, -0.10), (0.0, 0.0), (1.0, -0.05)], [1.0])
    assert ok and abs(d[0] - math.exp(-0.05)) < 1e-10


def test_time_domain_with_mandatory_points():
    pts, dt_avg = O.time_points(0.0, 1.0, 0.1, [0.25])
    # [0, 0.25] in ceil(2.5) = 3 steps, [0.25, 1] in ceil(7.5) = 8 steps
    assert len(pts) == 12 and abs(dt_avg - 1.0 / 11.0) < 1e-15
    assert 0.25 in pts and pts[0] == 0.0 and pts[-1] == 1.0
    np.testing.assert_allclose(np.diff(pts[:4]), 0.25 / 3, rtol=1e-14)
    np.testing.assert_allclose(np.diff(pts[3:]), 0.75 / 8, rtol=1e-14)
    # points outside (t0, t1) or within 1e-14 of another are dropped
    a, _ = O.time_points(0.0, 1.0, 0.1, [0.0, 1.0, 0.5, 0.5 + 1e-15, 1.5])
    b, _ = O.time_points(0.0, 1.0, 0.1, [0.5])
    np.testing.assert_array_equal(a, b)
    # no interior point: the uniform domain of ceil((t1 - t0) / dt) steps
    c, dta = O.time_points(0.0, 1.0, 0.3, [])
    assert len(c) == 5 and dta == 0.25


def test_dividend_shift_on_linear_data():
    # the natural spline reproduces linear data, so the shift is exact:
    # u(x) = a + b x -> a + b ln(e^x - d) where the shifted point is on the grid
    x = np.linspace(-1.0, 1.0, 41)
    u = 0.3 + 0.7 * x
    d = 0.5
    out = O.dividend_shift(x, u, d, 1.0)
    s_adj = np.exp(x) - d
    for i in range(len(x)):
        if s_adj[i] > 1e-10 and math.log(s_adj[i]) >= x[0]:
            assert abs(out[i] - (0.3 + 0.7 * math.log(s_adj[i]))) < 1e-13
        else:
            assert out[i] == 1.0
    assert np.array_equal(O.dividend_shift(x, u, 0.0, 1.0), u)


def _price(*a, **k):
    st, g, ns, dt = O.price_single(O.Params(*a, **k))
    assert st == 0
    return g, ns, dt


def test_zero_or_outside_dividend_matches_no_dividend():  # discrete_dividend_event_test.cc:72-81
    base, _, _ = _price(100, 100, 1.0, 0.2, 0.05, 0.0, True)
    for divs in ([(0.25, 0.0)], [(0.0, 5.0)], [(1.0, 5.0)]):
        g, _, _ = _price(100, 100, 1.0, 0.2, 0.05, 0.0, True, dividends=divs)
        assert abs(g[0] - base[0]) < 1e-10


def test_dividend_moves_put_up_and_call_down():  # discrete_dividend_event_test.cc:48-93
    p0, _, _ = _price(100, 100, 1.0, 0.2, 0.05, 0.0, True)
    p1, _, _ = _price(100, 100, 1.0, 0.2, 0.05, 0.0, True, dividends=[(0.25, 5.0)])
    c0, _, _ = _price(100, 100, 1.0, 0.2, 0.05, 0.0, False)
    c1, _, _ = _price(100, 100, 1.0, 0.2, 0.05, 0.0, False, dividends=[(0.25, 5.0)])
    assert p1[0] > p0[0] and c1[0] < c0[0]
    small, _, _ = _price(80, 100, 1.0, 0.2, 0.05, 0.0, True, dividends=[(0.5, 2.0)])
    large, _, _ = _price(80, 100, 1.0, 0.2, 0.05, 0.0, True, dividends=[(0.5, 10.0)])
    assert large[0] > small[0]


def _european_put_with_dividend(S, K, T, sig, D, t_div):
    """r = q = 0: E[P_BS(S_td - D, T - td)], the stock worthless when S_td <= D."""
    def f(z):
        y = S * math.exp(-0.5 * sig * sig * t_div + sig * math.sqrt(t_div) * z)
        v = K if y <= D else O.bs_price(y - D, K, T - t_div, sig, 0.0, 0.0, put=True)
        return v * stats.norm.pdf(z)
    return integrate.quad(f, -12, 12, limit=400, epsabs=1e-12)[0]


def test_zero_rate_put_with_dividend_is_european():
    # with r = q = 0 early exercise never pays, also with a cash dividend
    # ahead, so the American put is the European one; the shift at the
    # dividend's time-to-expiry T - t_div is what makes them agree
    S, K, T, sig, D, t_div = 100.0, 100.0, 1.0, 0.2, 5.0, 0.25
    g, _, _ = _price(S, K, T, sig, 0.0, 0.0, True, dividends=[(t_div, D)])
    eu = _european_put_with_dividend(S, K, T, sig, D, t_div)
    wrong = _european_put_with_dividend(S, K, T, sig, D, T - t_div)
    assert abs(g[0] - eu) < 5e-3
    assert abs(wrong - eu) > 20 * 5e-3  # the check separates the two timings


def test_flat_curve_and_constant_forward_curve_equal_constant_rate():
    base, _, _ = _price(100, 100, 1.0, 0.2, 0.05, 0.02, True)
    flat, _, _ = _price(100, 100, 1.0, 0.2, 0.0, 0.02, True, curve=[(0.0, 0.0), (100.0, -5.0)])
    steps, _, _ = _price(100, 100, 1.0, 0.2, 0.0, 0.02, True,
                         curve=[(0.0, 0.0), (0.3, -0.015), (0.7, -0.035), (2.0, -0.1)])
    np.testing.assert_allclose(flat, base, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(steps, base, rtol=1e-12, atol=1e-14)


def test_sloped_curve_near_average_flat_rate():  # american_option_test.cc:183-221
    g, _, _ = _price(100, 100, 1.0, 0.2, 0.0, 0.02, True, curve=[(0.0, 0.0), (0.5, -0.025), (1.0, -0.055)])
    f, _, _ = _price(100, 100, 1.0, 0.2, 0.055, 0.02, True)
    assert 0.0 < g[0] < 100.0
    assert abs(g[0] - f[0]) <= 0.02 * f[0]


def test_curve_call_is_european_and_theta_reads_the_near_rate():
    # q = 0, rates >= 0: the American call is European with discount D(T);
    # theta differences the state one step before the end of the march,
    # the value at calendar time dt with discount D(T)/D(dt), so it pins the
    # direction of the curve's time axis (0% near today, 10% after 0.5y)
    curve = [(0.0, 0.0), (0.5, 0.0), (5.0, -0.45)]
    g, _, dt = _price(100, 100, 1.0, 0.2, 0.0, 0.0, False, curve=curve)
    _, _, ld, _, _ = O.curve_eval(curve, [1.0, dt])
    d_T, d_dt = math.exp(ld[0]), math.exp(ld[1])
    rz = -math.log(d_T)
    eu = O.bs_price(100, 100, 1.0, 0.2, rz, 0.0, put=False)
    assert abs(g[0] - eu) <= 0.05
    rz2 = -math.log(d_T / d_dt) / (1.0 - dt)
    theta = (O.bs_price(100, 100, 1.0 - dt, 0.2, rz2, 0.0, put=False) - eu) / dt
    assert abs(g[3] - theta) < 0.01 * abs(theta)
    # the reversed axis (10% near today) would give a theta far outside that
    rev = (O.bs_price(100, 100, 1.0 - dt, 0.2, (0.05 - 0.1 * dt) / (1.0 - dt), 0.0, put=False) -
           O.bs_price(100, 100, 1.0, 0.2, 0.05, 0.0, put=False)) / dt
    assert abs(rev - theta) > 0.2 * abs(theta)


def test_batch_dividend_dates_per_contract():
    # the batch resolves each contract's dividend dates as mandatory points of
    # its own time domain (american_option_batch.cpp:448-470): a shared-grid
    # batch equals per-contract explicit solves on that grid
    items = [O.Params(100, 100, 1.0, 0.2, 0.05, 0.0, True, dividends=[(0.3, 2.0)]),
             O.Params(95, 100, 0.8, 0.25, 0.03, 0.01, False, dividends=[(0.1, 1.0), (0.5, 1.5)]),
             O.Params(105, 100, 1.0, 0.2, 0.05, 0.0, True)]
    arr, n = O.params_array(items)
    g, st, failed, _ = O.price_batch(None, use_shared_grid=True, params_array=(arr, n))
    assert failed == 0
    spec, nt = O.estimate_batch_pde_grid(arr, n)
    for i, it in enumerate(items):
        T = it.p.maturity
        mand = [T - it._divs[k].calendar_time for k in range(it.p.n_dividends)] if it._divs else []
        st1, res = O.solve_ex(it, spec, nt, T / nt, mand)
        assert st1 == 0
        np.testing.assert_array_equal(res["greeks"], g[i])
