# SPDX-License-Identifier: MIT
"""Pins of the oracle's FDM implied volatility (IVSolver: Brent over full PDE
solves, CPU): the reference's own IVSolver tests (tests/iv_solver_test.cc:
fixture S = K = 100, T = 1, r = 0.05, put, market 10.45, Brent tolerance
1e-6, at most 100 iterations), its error codes, and round trips of the
oracle's own prices, also with a yield curve and discrete dividends."""
import math

import pytest

import pyoracle as O

OK, VALIDATION, ARBITRAGE = 0, 1, 2


def _iv(spot=100.0, strike=100.0, T=1.0, r=0.05, q=0.0, put=True, market=10.45, curve=None, divs=None):
    return O.fdm_iv(O.Params(spot, strike, T, 0.2, r, q, put, curve=curve, dividends=divs), market)


@pytest.mark.parametrize("kw, target, tol", [
    (dict(), 0.316, 0.05),                                    # ATMPutIVCalculation (:33-43)
    (dict(strike=110.0, market=15.0), 0.28, 0.08),            # ITMPutIVCalculation (:93-103)
    (dict(strike=90.0, market=2.5), 0.20, 0.08),              # OTMPutIVCalculation (:105-115)
    (dict(spot=50.0, market=54.5), 0.875, 0.10),              # DeepITMPutIVCalculation (:117-129)
])
def test_reference_iv_values(kw, target, tol):
    st, iv, _, it = _iv(**kw)
    assert st == OK and it > 0
    assert abs(iv - target) <= tol


def test_deep_otm_put_and_atm_call():  # :131-159
    st, iv, _, _ = _iv(spot=200.0, market=1.0)
    assert st == OK and 0.0 < iv < 1.5
    st, iv, _, _ = _iv(put=False, market=10.0)
    assert st == OK and 0.0 < iv < 1.0


@pytest.mark.parametrize("kw, status", [
    (dict(spot=-100.0), VALIDATION),      # NegativeSpot (:45-55)
    (dict(strike=0.0), VALIDATION),       # NegativeStrike (:57-67)
    (dict(T=-1.0), VALIDATION),           # NegativeMaturity (:69-79)
    (dict(market=-5.0), VALIDATION),      # NegativeMarketPrice (:81-91)
    (dict(market=120.0), ARBITRAGE),      # above the put's upper bound K
    (dict(spot=50.0, market=40.0), ARBITRAGE),  # below intrinsic
    (dict(q=-0.01), ARBITRAGE),           # InvalidDividend -> ArbitrageViolation (iv_result.hpp:86-89)
])
def test_error_codes(kw, status):
    st, _, _, _ = _iv(**kw)
    assert st == status


@pytest.mark.parametrize("sigma, put, curve, divs", [
    (0.25, True, None, None),
    (0.45, False, None, None),
    (0.30, True, [(0.0, 0.0), (0.5, -0.02), (2.0, -0.11)], None),
    (0.22, True, None, [(0.5, 2.0)]),               # DiscreteDividendIVRoundTrip (:262-297)
    (0.35, False, [(0.0, 0.0), (1.0, -0.03)], [(0.3, 1.0), (0.6, 1.0)]),
])
def test_round_trip_of_own_prices(sigma, put, curve, divs):
    # price at sigma with the objective's own grid resolution, then invert
    p = O.Params(100.0, 105.0, 1.0, sigma, 0.04, 0.01, put, curve=curve, dividends=divs)
    spec, nt = O.estimate_pde_grid(p.p)
    mand = [1.0 - t for t, _ in (divs or []) if 0.0 < t < 1.0]
    st, res = O.solve_ex(p, spec, nt, 1.0 / nt, mand)
    assert st == 0
    market = float(res["greeks"][0])
    st, iv, err, it = O.fdm_iv(p, market)
    assert st == OK and it > 0
    assert abs(iv - sigma) < 1e-5, (iv, sigma, err)
