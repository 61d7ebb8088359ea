# SPDX-License-Identifier: MIT
"""GPU parity of the FDM implied volatility (mango_solve_iv,
mango_solve_iv_batch: Brent over full PDE solves on the GPU) against the
oracle's IVSolver: same status and Brent iteration count, implied vol within
1e-9 (the objective values agree to ~1e-12, far inside Brent's 1e-6)."""
import sys

import numpy as np
import pytest

import pyoracle as O
import mango_option as M

pytestmark = pytest.mark.gpu

IV_ABS = 1e-9
CURVE = [(0.0, 0.0), (0.5, -0.02), (2.0, -0.11)]

# (spot, strike, T, r, q, put, market, curve, dividends): tests/iv_solver_test.cc cases and extensions
CASES = [
    (100.0, 100.0, 1.0, 0.05, 0.0, True, 10.45, None, None),
    (100.0, 110.0, 1.0, 0.05, 0.0, True, 15.0, None, None),
    (100.0, 90.0, 1.0, 0.05, 0.0, True, 2.5, None, None),
    (50.0, 100.0, 1.0, 0.05, 0.0, True, 54.5, None, None),
    (200.0, 100.0, 1.0, 0.05, 0.0, True, 1.0, None, None),
    (100.0, 100.0, 1.0, 0.05, 0.0, False, 10.0, None, None),
    (100.0, 105.0, 0.75, 0.0, 0.01, True, 9.0, CURVE, None),
    (100.0, 100.0, 1.0, 0.05, 0.0, True, 11.0, None, [(0.5, 2.0)]),
    (110.0, 100.0, 1.2, 0.0, 0.0, False, 16.0, CURVE, [(0.3, 1.0), (0.8, 1.0)]),
    (100.0, 100.0, 1.0, 0.05, 0.0, True, 120.0, None, None),     # arbitrage
    (100.0, 100.0, 1.0, 0.05, 0.0, True, -5.0, None, None),      # validation
    (100.0, 100.0, 1.0, 0.05, 0.0, True, 1e-9, None, None),      # below the model's range: bracketing
]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mq(spot, strike, T, r, q, put, market, curve, divs):
    rate = M.YieldCurve.from_points(curve) if curve else r
    return M.IVQuery(spot=spot, strike=strike, maturity=T, rate=rate, dividend_yield=q,
                     option_type=M.OptionType.PUT if put else M.OptionType.CALL, market_price=market,
                     discrete_dividends=[M.Dividend(a, b) for a, b in (divs or [])])


def _oracle(c):
    spot, strike, T, r, q, put, market, curve, divs = c
    return O.fdm_iv(O.Params(spot, strike, T, 0.2, r, q, put, curve=curve, dividends=divs), market)


def test_single_queries_match_oracle():
    _say(f"test_single_queries_match_oracle: {len(CASES)} queries through mango_solve_iv")
    solver = M.IVSolver()
    for i, c in enumerate(CASES):
        _say(f"  case {i}: {c}")
        ok, s, e = solver.solve(_mq(*c))
        rst, riv, _, rit = _oracle(c)
        if rst == 0:
            assert ok, (i, e)
            assert s.iterations == rit and abs(s.implied_vol - riv) <= IV_ABS, (i, s, riv, rit)
        else:
            assert not ok and e.code == rst, (i, e.code, rst)


def test_batch_matches_single_and_oracle():
    _say("test_batch_matches_single_and_oracle: one batch of all cases")
    solver = M.IVSolver()
    res, failed = solver.solve_batch([_mq(*c) for c in CASES])
    ref = [_oracle(c) for c in CASES]
    assert failed == sum(1 for r in ref if r[0] != 0)
    for i, ((ok, s, e), (rst, riv, _, rit)) in enumerate(zip(res, ref)):
        if rst == 0:
            assert ok and s.iterations == rit and abs(s.implied_vol - riv) <= IV_ABS, (i, s, riv)
        else:
            assert not ok and e.code == rst, (i, e.code, rst)


def test_config_limits():
    _say("test_config_limits: max_iter and tolerance")
    tight = M.IVSolverConfig(M.RootFindingConfig(max_iter=2, brent_tol_abs=1e-12))
    ok, _, e = M.IVSolver(tight).solve(_mq(*CASES[0]))
    assert not ok and e.code == 3  # MANGO_ERR_NO_CONVERGENCE
    loose = M.IVSolverConfig(M.RootFindingConfig(max_iter=100, brent_tol_abs=1e-3))
    ok, s, _ = M.IVSolver(loose).solve(_mq(*CASES[0]))
    rst, riv, _, rit = O.fdm_iv(O.Params(100.0, 100.0, 1.0, 0.2, 0.05, 0.0, True), 10.45, 100, 1e-3)
    assert ok and rst == 0 and s.iterations == rit and abs(s.implied_vol - riv) <= IV_ABS
