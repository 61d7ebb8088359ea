# SPDX-License-Identifier: MIT
"""CPU tests pinning the oracle's reverse-mode sensitivities (the checker of
the device adjoint):

* central finite differences of the frozen-exercise-set map (the forward
  replaying the base run's locked and clamped rows) at the adjoint
  configuration's steps h_sigma = 1e-4, h_r = h_q = 1e-5: 1e-6 relative;
* central differences of the true map with steps small enough that no row
  changes its exercise state (1e-7): agreement to the FD rounding level;
* dV/dK = (V - S Delta)/K against a strike bump on the frozen grid;
* European limits (call with q = 0, put with r = 0: no early exercise) against
  the Black-Scholes vega, rho and dividend sensitivity, converging with the grid.
"""
import ctypes as C
import math

import numpy as np
import pytest

import pyoracle as O
import mango_option as M
from mango_option import _abi, workloads as W

N_SPACE, N_TIME = 101, 498


def _cfg(n=N_SPACE, nt=N_TIME):
    acc = M.GridAccuracyParams(min_spatial_points=n, max_spatial_points=n)
    return acc, M.grid_config(_abi.MANGO_GRID_MODE_ACCURACY_FIXED_TIME, acc, n_time=nt)


def _batch():
    puts = W.c1_puts(20, seed=5)
    chains, _ = W.c2_chains(1, 12, seed=1)
    return W.OptionBatch(*(np.concatenate([getattr(puts, f), getattr(chains, f)]) for f in
                           ("spot", "strike", "maturity", "volatility", "rate", "dividend_yield", "option_type")))


FIELDS = (("volatility", 2), ("rate_const", 3), ("dividend_yield", 4), ("strike", 5))


def _fd(p, spec, field, h, frozen):
    lib = O.load()

    def value(sign):
        q = _abi.MangoPricingParams.from_buffer_copy(p)
        setattr(q, field, getattr(q, field) + sign * h)
        g = _abi.MangoAmericanGreeks()
        if frozen:
            rc = lib.oracle_solve_frozen(C.byref(p), C.byref(q), C.byref(spec), N_TIME, C.byref(g))
        else:
            rc = lib.oracle_solve(C.byref(q), C.byref(spec), N_TIME, None, None, None, C.byref(g))
        assert rc == 0
        return g.value

    return (value(1) - value(-1)) / (2 * h)


def test_adjoint_matches_frozen_set_finite_differences():
    opts = _batch()
    acc, cfg = _cfg()
    out, st = O.adjoint_batch(opts, cfg, n_threads=8)
    assert (st == 0).all()
    arr, n = O.c_params(opts)
    steps = {"volatility": 1e-4, "rate_const": 1e-5, "dividend_yield": 1e-5, "strike": 1e-4}
    for i in range(n):
        spec, _ = O.estimate_pde_grid(arr[i], acc.to_c())
        for field, col in FIELDS:
            fd = _fd(arr[i], spec, field, steps[field], frozen=True)
            # 1e-6 relative; the floor covers sensitivities near zero (deep
            # in the money, held at intrinsic), where FD rounding dominates
            assert abs(out[i, col] - fd) <= 1e-6 * abs(fd) + 1e-8, (i, field, out[i, col], fd)


def test_adjoint_matches_true_map_at_small_steps():
    opts = _batch().take(np.arange(0, 32, 3))
    acc, cfg = _cfg()
    out, _ = O.adjoint_batch(opts, cfg, n_threads=8)
    arr, n = O.c_params(opts)
    for i in range(n):
        spec, _ = O.estimate_pde_grid(arr[i], acc.to_c())
        for field, col in FIELDS[:3]:
            fd = _fd(arr[i], spec, field, 1e-7, frozen=False)
            assert abs(out[i, col] - fd) <= 1e-6 * abs(fd) + 1e-7, (i, field, out[i, col], fd)


def test_strike_sensitivity_closed_form():
    opts = _batch().take(np.arange(0, 32, 4))
    acc, cfg = _cfg()
    out, _ = O.adjoint_batch(opts, cfg)
    np.testing.assert_allclose(out[:, 5], (out[:, 0] - opts.spot * out[:, 1]) / opts.strike, rtol=1e-14)
    arr, n = O.c_params(opts)
    for i in range(n):
        spec, _ = O.estimate_pde_grid(arr[i], acc.to_c())
        fd = _fd(arr[i], spec, "strike", 1e-4, frozen=False)
        assert abs(out[i, 5] - fd) <= 1e-8 * abs(fd) + 1e-10


def _bs_sens(S, K, T, sig, r, q, put):
    d1 = (math.log(S / K) + (r - q + 0.5 * sig * sig) * T) / (sig * math.sqrt(T))
    d2 = d1 - sig * math.sqrt(T)
    cdf = lambda z: 0.5 * math.erfc(-z / math.sqrt(2.0))
    pdf = lambda z: math.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)
    vega = S * math.exp(-q * T) * pdf(d1) * math.sqrt(T)
    if put:
        return vega, -K * T * math.exp(-r * T) * cdf(-d2), S * T * math.exp(-q * T) * cdf(-d1)
    return vega, K * T * math.exp(-r * T) * cdf(d2), -S * T * math.exp(-q * T) * cdf(d1)


@pytest.mark.parametrize("grid,tol", [((101, 498), 1e-3), ((401, 2000), 1e-4)])
def test_european_limits_match_black_scholes(grid, tol):
    cases = [(100, 110, 1.0, 0.25, 0.05, 0.0, 0), (100, 90, 0.5, 0.3, 0.03, 0.0, 0),
             (100, 100, 1.0, 0.2, 0.0, 0.02, 1), (100, 120, 2.0, 0.35, 0.0, 0.03, 1)]
    b = W.OptionBatch(*(np.array([c[j] for c in cases], dtype=float) for j in range(6)),
                      np.array([c[6] for c in cases], dtype=np.int32))
    _, cfg = _cfg(*grid)
    out, st = O.adjoint_batch(b, cfg)
    assert (st == 0).all()
    for c, o in zip(cases, out):
        ref = _bs_sens(*c[:6], c[6] == 1)
        for k in range(3):
            assert abs(o[2 + k] - ref[k]) <= tol * abs(ref[k]), (c, k, o[2 + k], ref[k])


def test_invalid_option_reports_validation():
    opts = W.c1_puts(3, seed=5)
    opts.volatility[1] = 0.0
    _, cfg = _cfg()
    out, st = O.adjoint_batch(opts, cfg)
    assert st[1] == _abi.MANGO_ERR_VALIDATION and np.isnan(out[1]).all()
    assert st[0] == 0 and st[2] == 0
