# SPDX-License-Identifier: MIT
"""GPU parity of the event kernel (yield curves, mandatory time points,
discrete dividends) against the oracle, through the C ABI: the single-option
entry (mango_price_american), per-option and shared-grid batches, the
normalized chain with a shared curve, the fixed-step grid mode and mixed
batches that route plain and event jobs together.

Bar: prices within 1e-9 absolute, Greeks within 1e-7 relative (1e-10 floor).
"""
import sys

import numpy as np
import pytest

import pyoracle as O
import mango_option as M
from mango_option import _abi

pytestmark = pytest.mark.gpu

PRICE_ABS = 1e-9
GREEK_REL = 1e-7
GREEK_FLOOR = 1e-10

CURVE_UP = [(0.0, 0.0), (0.5, -0.025), (1.0, -0.055)]                  # american_option_test.cc:186-190
CURVE_STEP = [(0.0, 0.0), (0.25, -0.0025), (0.6, -0.0305), (2.0, -0.1425)]  # 1%, 8%, 8%
CURVE_NEG = [(0.0, 0.0), (0.3, 0.003), (1.5, -0.045)]                  # -1% then 4%

# (spot, strike, T, vol, r, q, put, curve, dividends)
CASES = [
    (100, 100, 1.0, 0.20, 0.0, 0.02, True, CURVE_UP, None),
    (100, 100, 1.0, 0.20, 0.0, 0.00, False, CURVE_UP, None),
    (90, 100, 1.5, 0.30, 0.0, 0.01, True, CURVE_STEP, None),
    (110, 100, 0.7, 0.25, 0.0, 0.03, False, CURVE_NEG, None),
    (100, 100, 1.0, 0.20, 0.05, 0.0, True, None, [(0.25, 5.0)]),
    (100, 100, 1.0, 0.20, 0.05, 0.0, False, None, [(0.25, 5.0)]),
    (80, 100, 1.0, 0.20, 0.05, 0.0, True, None, [(0.5, 10.0)]),
    (120, 100, 1.0, 0.25, 0.03, 0.0, False, None, [(0.2, 2.0), (0.45, 2.0), (0.7, 2.0), (0.95, 2.0)]),
    (100, 100, 1.0, 0.20, 0.05, 0.0, True, None, [(0.5, 1.0), (0.5, 1.5)]),          # same-date merge
    (30, 100, 1.0, 0.40, 0.05, 0.0, True, None, [(0.3, 40.0)]),                      # D > S: fallback nodes
    (100, 100, 1.0, 0.20, 0.05, 0.0, True, None, [(0.9999, 3.0)]),                   # inside the Rannacher step
    (100, 100, 1.0, 0.20, 0.05, 0.0, False, None, [(0.0, 3.0), (1.0, 3.0), (0.4, 0.0)]),  # all filtered
    (105, 100, 1.2, 0.22, 0.0, 0.01, True, CURVE_STEP, [(0.3, 1.5), (0.8, 1.5)]),
    (95, 100, 1.2, 0.22, 0.0, 0.00, False, CURVE_UP, [(0.55, 4.0)]),
]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mp(spot, strike, T, vol, r, q, put, curve, divs):
    rate = M.YieldCurve.from_points(curve) if curve else r
    return M.PricingParams(spot=spot, strike=strike, maturity=T, rate=rate, dividend_yield=q,
                           option_type=M.OptionType.PUT if put else M.OptionType.CALL, volatility=vol,
                           discrete_dividends=[M.Dividend(a, b) for a, b in (divs or [])])


def _op(spot, strike, T, vol, r, q, put, curve, divs):
    return O.Params(spot, strike, T, vol, r, q, put, curve=curve, dividends=divs)


def _check(g_gpu, s_gpu, g_ref, s_ref, label):
    g_gpu, g_ref = np.atleast_2d(g_gpu), np.atleast_2d(g_ref)
    assert np.array_equal(s_gpu, s_ref), f"{label}: status mismatch {s_gpu} vs {s_ref}"
    ok = np.asarray(s_ref) == 0
    dv = np.abs(g_gpu[ok, 0] - g_ref[ok, 0])
    assert dv.max(initial=0.0) <= PRICE_ABS, f"{label}: price diff {dv.max()}"
    for j, name in ((1, "delta"), (2, "gamma"), (3, "theta")):
        d = np.abs(g_gpu[ok, j] - g_ref[ok, j])
        tol = GREEK_REL * np.abs(g_ref[ok, j]) + GREEK_FLOOR
        bad = d > tol
        assert not bad.any(), f"{label}: {name} diff {d[bad][:4]} ref {g_ref[ok, j][bad][:4]}"
    return float(dv.max(initial=0.0))


def test_single_option_entry():
    _say(f"test_single_option_entry: {len(CASES)} curve/dividend cases through mango_price_american")
    worst = 0.0
    for i, c in enumerate(CASES):
        _say(f"  case {i}: {c}")
        r = M.american_option_price(_mp(*c))
        g = np.array([r.value(), r.delta(), r.gamma(), r.theta()])
        st, gr, _, _ = O.price_single(_op(*c))
        assert st == 0
        worst = max(worst, _check(g[None], [0], gr[None], [0], f"case {i}"))
    _say(f"  max |dprice| = {worst:.3e}")


def test_batch_per_option_and_shared_grid():
    _say("test_batch_per_option_and_shared_grid")
    mp = [_mp(*c) for c in CASES]
    op = [_op(*c) for c in CASES]
    for shared in (False, True):
        g, s, failed = M.price_batch(mp, shared, None)
        gr, sr, fr, _ = O.price_batch(None, shared, None, params_array=O.params_array(op))
        assert failed == fr == 0
        err = _check(g, s, gr, sr, f"shared={shared}")
        _say(f"  shared={shared}: max |dprice| = {err:.3e}")


def test_fixed_time_and_explicit_grids():
    _say("test_fixed_time_and_explicit_grids")
    mp = [_mp(*c) for c in CASES]
    op = [_op(*c) for c in CASES]
    acc = M.GridAccuracyParams(min_spatial_points=101, max_spatial_points=101)
    cfgs = [M.grid_config(_abi.MANGO_GRID_MODE_ACCURACY_FIXED_TIME, acc, n_time=300),
            M.grid_config(_abi.MANGO_GRID_MODE_EXPLICIT, spec=M.GridSpec.sinh_spaced(-3.0, 3.0, 151, 2.0),
                          n_time=400)]
    for k, cfg in enumerate(cfgs):
        g, s, _ = M.price_batch(mp, False, cfg, use_normalized=False)
        gr, sr, _, _ = O.price_batch(None, False, cfg, use_normalized=False, params_array=O.params_array(op))
        err = _check(g, s, gr, sr, f"grid mode {k}")
        _say(f"  grid mode {k}: max |dprice| = {err:.3e}")


def test_normalized_chain_with_curve():
    _say("test_normalized_chain_with_curve: one curve, 24 strikes, shared grid")
    cases = [(100.0, K, 1.0, 0.25, 0.0, 0.01, True, CURVE_STEP, None) for K in np.linspace(85, 115, 24)]
    mp = [_mp(*c) for c in cases]
    op = [_op(*c) for c in cases]
    g, s, failed = M.price_batch(mp, True, None)
    gr, sr, fr, solves = O.price_batch(None, True, None, params_array=O.params_array(op))
    assert solves == 1 and failed == fr == 0
    err = _check(g, s, gr, sr, "chain")
    _say(f"  max |dprice| = {err:.3e}")


def test_mixed_batch_routes_both_kernels():
    _say("test_mixed_batch_routes_both_kernels: plain + curve + dividend options")
    cases = []
    rng = np.random.default_rng(11)
    for i in range(40):
        S = float(rng.uniform(80, 120))
        base = (S, 100.0, float(rng.uniform(0.3, 1.5)), float(rng.uniform(0.15, 0.4)), 0.04, 0.01, bool(i % 2))
        kind = i % 4
        if kind == 0:
            cases.append(base + (None, None))
        elif kind == 1:
            cases.append(base[:4] + (0.0,) + base[5:] + (CURVE_STEP, None))
        elif kind == 2:
            cases.append(base + (None, [(0.5 * base[2], 2.0)]))
        else:
            cases.append(base[:4] + (0.0,) + base[5:] + (CURVE_UP, [(0.25 * base[2], 1.0), (0.75 * base[2], 1.0)]))
    mp = [_mp(*c) for c in cases]
    op = [_op(*c) for c in cases]
    g, s, failed = M.price_batch(mp, False, None, use_normalized=False)
    gr, sr, fr, _ = O.price_batch(None, False, None, use_normalized=False, params_array=O.params_array(op))
    assert failed == fr == 0
    err = _check(g, s, gr, sr, "mixed")
    _say(f"  max |dprice| = {err:.3e}")


def test_invalid_curves_and_dividends_fail_per_slot():
    _say("test_invalid_curves_and_dividends_fail_per_slot")
    good = _mp(100, 100, 1.0, 0.2, 0.0, 0.0, True, CURVE_UP, None)
    bad_curve = _mp(100, 100, 1.0, 0.2, 0.0, 0.0, True, None, None)
    bad_curve.rate = M.YieldCurve([(0.5, 0.0), (1.0, -0.05)])  # no t = 0 point
    bad_div = _mp(100, 100, 1.0, 0.2, 0.05, 0.0, True, None, [(1.5, 1.0)])  # after maturity
    neg_div = _mp(100, 100, 1.0, 0.2, 0.05, 0.0, True, None, [(0.5, -1.0)])
    g, s, failed = M.price_batch([good, bad_curve, bad_div, neg_div], False, None)
    assert failed == 3
    assert s[0] == 0 and all(s[k] == _abi.MANGO_ERR_VALIDATION for k in (1, 2, 3))
    assert np.isnan(g[1:]).all()
    with pytest.raises(ValueError):
        M.american_option_price(bad_div)


def test_large_event_batch_uses_packed_groups():
    # more than 1024 event jobs take the several-jobs-per-wavefront variants
    _say("test_large_event_batch_uses_packed_groups: 1100 puts/calls with a curve and a dividend, 101x300")
    rng = np.random.default_rng(21)
    cases = []
    for i in range(1100):
        T = float(rng.uniform(0.2, 1.5))
        cases.append((float(rng.uniform(80, 120)), 100.0, T, float(rng.uniform(0.15, 0.45)), 0.0,
                      float(rng.uniform(0.0, 0.03)), bool(i % 2), CURVE_STEP, [(float(rng.uniform(0.05, 0.95)) * T, 1.5)]))
    mp = [_mp(*c) for c in cases]
    op = [_op(*c) for c in cases]
    acc = M.GridAccuracyParams(min_spatial_points=101, max_spatial_points=101)
    for n_pts in (101, 121):
        acc = M.GridAccuracyParams(min_spatial_points=n_pts, max_spatial_points=n_pts)
        cfg = M.grid_config(_abi.MANGO_GRID_MODE_ACCURACY_FIXED_TIME, acc, n_time=300)
        g, s, failed = M.price_batch(mp, False, cfg, use_normalized=False)
        gr, sr, fr, _ = O.price_batch(None, False, cfg, use_normalized=False, n_threads=8,
                                      params_array=O.params_array(op))
        assert failed == fr == 0
        err = _check(g, s, gr, sr, f"{n_pts} points")
        _say(f"  {n_pts} points: max |dprice| = {err:.3e}")
