# SPDX-License-Identifier: MIT
"""Pins of the oracle's full American price (CPU, no GPU): the reference's
QuantLib-validated constants, the invariants of its pricing tests, closed
forms, exact identities of the chain route and self-convergence."""
import math

import numpy as np
import pytest

import pyoracle as O
import mango_option as M
from mango_option import _abi, workloads as W


def _auto(spot, strike, T, vol, r, q, put=True):
    p = O.one_param(spot, strike, T, vol, r, q, put)
    spec, nt = O.estimate_pde_grid(p)
    st, res = O.solve(p, spec, nt)
    assert st == 0
    return res, spec, nt


def _golden_ok(computed, expected):  # benchmarks/readme_benchmarks.cc:46-58
    abs_err = abs(computed - expected)
    return abs_err <= 0.01 or abs_err / expected <= 0.001


def test_quantlib_golden_atm_put():  # benchmarks/readme_benchmarks.cc:32-33, 180-208
    res, spec, nt = _auto(100, 100, 1.0, 0.20, 0.05, 0.02)
    assert _golden_ok(res["greeks"][0], 6.65996306)


def test_quantlib_golden_batch_first_strike():  # benchmarks/readme_benchmarks.cc:37-44, 259
    # The reference validates only K=90.0 of its five constants.
    res, _, _ = _auto(100, 90.0, 1.0, 0.20, 0.05, 0.02)
    assert _golden_ok(res["greeks"][0], 2.82114033)


def test_auto_grid_of_readme_put_is_101x120():
    # The current estimator (src/option/grid_spec_types.cpp:37-118) gives
    # 101x120 for the README put, not the "101x498" label of README.md:114.
    _, spec, nt = _auto(100, 100, 1.0, 0.20, 0.05, 0.02)
    assert (spec.n_points, nt) == (101, 120)
    assert spec.n_clusters == 1 and abs(spec.clusters[0].center_x) < 1e-15


def test_primary_header_atm():  # tests/american_option_test.cc:325-333
    res, _, _ = _auto(100, 100, 1.0, 0.20, 0.05, 0.02)
    assert abs(res["greeks"][0] - 6.35) <= 0.5


def test_put_respects_intrinsic():  # tests/american_option_test.cc:44-55
    res, _, _ = _auto(100, 110, 1.0, 0.25, 0.03, 0.0)
    assert 10.0 - 1e-6 <= res["greeks"][0] < 110.0


def test_call_increases_with_vol():  # tests/american_option_test.cc:57-81
    vals = [_auto(100, 100, 1.0, v, 0.05, 0.0, put=False)[0]["greeks"][0] for v in (0.15, 0.25, 0.4)]
    assert vals[0] < vals[1] < vals[2]


def test_put_increases_with_maturity():  # tests/american_option_test.cc:83-99
    vals = [_auto(100, 95, T, 0.2, 0.02, 0.0)[0]["greeks"][0] for T in (0.25, 0.5, 1.0, 2.0)]
    assert all(b >= a - 5e-3 for a, b in zip(vals, vals[1:]))


def test_deep_itm_put_is_intrinsic():  # tests/american_option_test.cc:145-161
    res, _, _ = _auto(0.25, 100, 0.75, 0.2, 0.05, 0.0)
    assert abs(res["greeks"][0] - 99.75) <= 1e-3


def test_atm_put_keeps_time_value():  # tests/american_option_test.cc:163-181
    res, _, _ = _auto(100, 100, 1.0, 0.25, 0.05, 0.0)
    assert 7.0 < res["greeks"][0] < 12.0


@pytest.mark.parametrize("spot", [150.0, 200.0, 300.0])
def test_no_dividend_call_is_european(spot):  # tests/american_option_test.cc:391-410
    res, _, _ = _auto(spot, 100, 1.0, 0.2, 0.05, 0.0, put=False)
    eu = O.bs_price(spot, 100, 1.0, 0.2, 0.05, 0.0, put=False)
    assert res["greeks"][0] >= spot - 100 - 1e-6
    assert abs(res["greeks"][0] - eu) <= 0.05


def test_zero_rate_deep_itm_put_is_european():  # tests/american_option_test.cc:416-430
    res, _, _ = _auto(4.0, 100, 1.0, 0.2, 0.0, 0.04)
    eu = O.bs_price(4.0, 100, 1.0, 0.2, 0.0, 0.04, put=True)
    assert abs(res["greeks"][0] - eu) <= 0.05


def test_uniform_grid_call_is_deterministic():  # tests/american_option_test.cc:346-389
    p = O.one_param(100, 100, 1.0, 0.2, 0.05, 0.0, False)
    spec = M.GridSpec.uniform(-0.6, 0.6, 201).to_c()
    _, a = O.solve(p, spec, 200)
    _, b = O.solve(p, M.GridSpec.uniform(-0.6, 0.6, 301).to_c(), 200)
    _, c = O.solve(p, spec, 200)
    assert a["greeks"][0] == c["greeks"][0]


def test_european_closed_form_parity_and_vega():
    S, K, T, s, r, q = 105.0, 100.0, 0.7, 0.3, 0.04, 0.015
    c = O.bs_price(S, K, T, s, r, q, put=False)
    p = O.bs_price(S, K, T, s, r, q, put=True)
    assert abs((c - p) - (S * math.exp(-q * T) - K * math.exp(-r * T))) < 1e-12
    h = 1e-5
    fd = (O.bs_price(S, K, T, s + h, r, q) - O.bs_price(S, K, T, s - h, r, q)) / (2 * h)
    assert abs(fd - O.bs_vega(S, K, T, s, r, q)) < 1e-6


def test_greek_identities():
    # V(S,K) = K v(ln(S/K)) on a fixed grid: delta from the spline derivative,
    # homogeneity S*delta + K*dV/dK = V checked against a strike bump on the
    # same grid (the PDE solution does not change with K).
    p = O.one_param(100, 100, 1.0, 0.25, 0.03, 0.01, True)
    spec = M.GridSpec.sinh_spaced(-2.0, 2.0, 201, 2.0).to_c()
    _, r0 = O.solve(p, spec, 400)
    h = 1e-3
    vs = []
    for K in (100 - h, 100 + h):
        _, rk = O.solve(O.one_param(100, K, 1.0, 0.25, 0.03, 0.01, True), spec, 400)
        vs.append(rk["greeks"][0])
    dvdk = (vs[1] - vs[0]) / (2 * h)
    V, D = r0["greeks"][0], r0["greeks"][1]
    assert abs(100 * D + 100 * dvdk - V) < 1e-5


def test_normalized_chain_equals_regular_bitwise():
    # strike-free PDE: one S=K=1 solve on the shared grid reproduces every
    # member of the regular route (src/option/american_option_batch.cpp:313-382)
    opts, chain = W.c2_chains(2, 16, seed=1)
    sub = opts.take(np.where((chain == 0) & (opts.option_type == W.PUT))[0])
    arr, n = O.c_params(sub)
    acc = M.GridAccuracyParams(min_spatial_points=101, max_spatial_points=101).to_c()
    spec, _ = O.estimate_batch_pde_grid(arr, n, acc)
    cfg = M.grid_config(_abi.MANGO_GRID_MODE_EXPLICIT, spec=M.GridSpec.from_c(spec), n_time=498)
    g1, s1, _, solves1 = O.price_batch(sub, True, cfg, use_normalized=True)
    g2, s2, _, solves2 = O.price_batch(sub, True, cfg, use_normalized=False)
    assert solves1 == 1 and solves2 == n
    assert np.array_equal(g1, g2)


def test_second_order_self_convergence():
    # benchmarks/BENCHMARK_RESULTS.md:116-124 reports ~4x error reduction per
    # doubling of (n, N) for the ATM put.
    p = O.one_param(100, 100, 1.0, 0.2, 0.05, 0.0, True)
    vals = []
    for n in (51, 101, 201, 401):
        _, r = O.solve(p, M.GridSpec.sinh_spaced(-3.0, 3.0, n, _abi.default_sinh_alpha()).to_c(), 10 * (n - 1))
        vals.append(r["greeks"][0])
    d = np.diff(vals)
    ratios = d[:-1] / d[1:]
    assert all(2.5 < q < 6.0 for q in ratios), ratios
