# SPDX-License-Identifier: MIT
"""CPU tests pinning the oracle's price-table path (the checker of the device
table build and the interpolated-IV batch):

* clamped cubic B-spline collocation reproduces tensor products of cubic
  polynomials exactly (values and first partials), interpolates arbitrary
  data at the nodes, and a table of ones is a partition of unity;
* analytic partials against central differences of the spline;
* eep_floor closed values (softplus with sharpness 100, identity above 5);
* the tensor at (m, tau, sigma, r) equals the natural spline of the solve's
  snapshot at ln m, with the reference's snapshot indexing (entry j of the
  sorted, deduplicated list; columns past its end repeat the last one);
* the table price at the nodes is the American price with the EEP floor
  applied; the interpolated IV inverts the table price; the statuses follow
  the reference's validation order.
"""
import math

import numpy as np
import pytest

import pyoracle as O
import mango_option as M
from mango_option import _abi, workloads as W


def _grids(rng, sizes):
    out = []
    for n in sizes:
        g = np.sort(rng.uniform(-1.0, 1.0, n))
        g[0], g[-1] = -1.0, 1.0
        out.append(g)
    return out


def test_bspline_reproduces_tensor_cubics():
    rng = np.random.default_rng(11)
    ax = _grids(rng, (7, 6, 5, 4))
    ca = [rng.normal(size=4) for _ in range(4)]  # cubic per axis
    f = lambda a, x: ca[a][0] + x * (ca[a][1] + x * (ca[a][2] + x * ca[a][3]))
    df = lambda a, x: ca[a][1] + x * (2 * ca[a][2] + 3 * x * ca[a][3])
    X = np.meshgrid(*ax, indexing="ij")
    vals = f(0, X[0]) * f(1, X[1]) * f(2, X[2]) * f(3, X[3])
    coef, res = O.bspline_fit4(ax, vals)
    assert res < 1e-12
    pts = np.stack([rng.uniform(-1, 1, 300) for _ in range(4)], 1)
    want = np.prod([f(a, pts[:, a]) for a in range(4)], axis=0)
    np.testing.assert_allclose(O.bspline_eval4(ax, coef, pts), want, rtol=0, atol=1e-12)
    for a in range(4):
        d = np.prod([df(b, pts[:, b]) if b == a else f(b, pts[:, b]) for b in range(4)], axis=0)
        np.testing.assert_allclose(O.bspline_eval4(ax, coef, pts, a), d, rtol=0, atol=1e-10)


def test_bspline_interpolates_nodes_and_partition_of_unity():
    rng = np.random.default_rng(12)
    ax = _grids(rng, (9, 5, 6, 4))
    vals = rng.normal(size=[len(a) for a in ax])
    coef, _ = O.bspline_fit4(ax, vals)
    X = np.meshgrid(*ax, indexing="ij")
    nodes = np.stack([x.ravel() for x in X], 1)
    np.testing.assert_allclose(O.bspline_eval4(ax, coef, nodes), vals.ravel(), rtol=0, atol=1e-11)
    ones = np.ones([len(a) for a in ax])
    pts = np.stack([rng.uniform(-1.2, 1.2, 200) for _ in range(4)], 1)  # includes clamped queries
    np.testing.assert_allclose(O.bspline_eval4(ax, ones, pts), 1.0, rtol=0, atol=1e-14)
    for a in range(4):
        np.testing.assert_allclose(O.bspline_eval4(ax, ones, pts[np.all(np.abs(pts) < 1, 1)], a), 0.0, atol=1e-12)


def test_bspline_partials_match_central_differences():
    rng = np.random.default_rng(13)
    ax = _grids(rng, (6, 7, 5, 5))
    coef = rng.normal(size=[len(a) for a in ax])
    pts = np.stack([rng.uniform(-0.9, 0.9, 100) for _ in range(4)], 1)
    h = 1e-6
    for a in range(4):
        e = np.zeros(4)
        e[a] = h
        fd = (O.bspline_eval4(ax, coef, pts + e) - O.bspline_eval4(ax, coef, pts - e)) / (2 * h)
        an = O.bspline_eval4(ax, coef, pts, a)
        np.testing.assert_allclose(an, fd, rtol=1e-6, atol=1e-6)


def test_eep_floor_closed_values():
    assert O.eep_floor(0.0) == 0.0
    assert O.eep_floor(-1.0) == 0.0
    assert O.eep_floor(0.01) == pytest.approx((math.log1p(math.e) - math.log(2.0)) / 100.0, rel=1e-15)
    # 100 x = 500 is still on the softplus branch; above it the floor is the identity
    assert O.eep_floor(5.0) == pytest.approx(5.0 - math.log(2.0) / 100.0, rel=1e-15)
    assert O.eep_floor(5.25) == 5.25


@pytest.fixture(scope="module")
def c4_table():
    ax = W.c4_axes()
    spec = M.GridSpec(n_points=201).to_c()
    return O.table_build(ax, W.C4_K_REF, W.C4_DIVIDEND, True, spec=spec, n_time=1000, n_threads=4)


def _fallback_accuracy(ax):
    """The explicit 201 x 1000 grid fails the width check, so the builder
    estimates per-solve grids with min = max = 201 points, 1000 steps, the
    grid's concentration and n_sigma covering the grid half-width."""
    max_sst = ax[2][-1] * math.sqrt(ax[1][-1])
    acc = M.GridAccuracyParams(min_spatial_points=201, max_spatial_points=201, max_time_steps=1000)
    acc.n_sigma = max(5.0, 3.0 / max_sst * 1.1, max(abs(ax[0][0]), abs(ax[0][-1])) / max_sst * 1.1)
    return acc


def test_c4_fallback_route(c4_table):
    assert c4_table.fallback and c4_table.pde_solves == 200
    assert c4_table.max_residual < 1e-9


@pytest.mark.parametrize("si,ri", [(0, 0), (8, 4), (19, 9)])
def test_c4_tensor_is_snapshot_spline(c4_table, si, ri):
    ax = W.c4_axes()
    p = O.one_param(1.0, 1.0, ax[1][-1], ax[2][si], ax[3][ri], W.C4_DIVIDEND)
    arr = (_abi.MangoPricingParams * 1)(p)
    spec, nt = O.estimate_batch_pde_grid(arr, 1, _fallback_accuracy(ax).to_c())
    st, snaps, _ = O.solve_snapshots(p, spec, nt, ax[1])
    assert st == 0
    x = O.generate(spec)
    for j in range(len(ax[1])):
        k = min(j, len(snaps) - 1)
        ok, want, _ = O.spline_eval(x, snaps[k], ax[0])
        assert ok
        np.testing.assert_allclose(c4_table.tensor[:, j, si, ri], want, rtol=0, atol=1e-13)
    if si == 19:  # few steps: two maturities share a step, the last column repeats
        assert len(snaps) == len(ax[1]) - 1
        assert np.array_equal(c4_table.tensor[:, -1, si, ri], c4_table.tensor[:, -2, si, ri])


def test_c4_price_at_nodes_is_floored_american(c4_table):
    ax = W.c4_axes()
    rng = np.random.default_rng(14)
    idx = np.stack([rng.integers(0, len(a), 200) for a in ax], 1)
    lm, tau, sig, rate = (ax[a][idx[:, a]] for a in range(4))
    S = W.C4_K_REF * np.exp(lm)
    K = np.full(len(S), W.C4_K_REF)
    price, _ = c4_table.price(np.stack([S, K, tau, sig, rate], 1))
    am = W.C4_K_REF * c4_table.tensor[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]]
    eu = np.array([O.bs_price(S[i], K[i], tau[i], sig[i], rate[i], W.C4_DIVIDEND) for i in range(len(S))])
    want = np.array([max(0.0, O.eep_floor(a - e)) for a, e in zip(am, eu)]) + eu
    np.testing.assert_allclose(price, want, rtol=0, atol=1e-9)


def test_c4_iv_inverts_table_price(c4_table):
    pts, _ = W.c4_query_points(400, seed=21)
    n = len(pts)
    S = W.C4_K_REF * np.exp(pts[:, 0])
    K = np.full(n, W.C4_K_REF)
    tau, r = pts[:, 1], pts[:, 3]
    price, _ = c4_table.price(np.stack([S, K, tau, pts[:, 2], r], 1))
    rows = np.stack([S, K, tau, r, np.full(n, W.C4_DIVIDEND), price], 1)
    iv, it, err, st = c4_table.iv(rows)
    ok = st == 0
    assert ok.mean() > 0.6
    # Brent stops at |price(iv) - market| < 1e-6 or a bracket narrower than 1e-6
    back, _ = c4_table.price(np.stack([S, K, tau, iv, r], 1)[ok])
    assert np.all((np.abs(back - price[ok]) < 1e-6) | (np.abs(iv[ok] - pts[ok, 2]) < 2e-6))
    assert np.all((it[ok] >= 1) & (it[ok] <= 50))
    # Each failure has its reason: below intrinsic (table price under the
    # payoff), no sign change over the bracket [0.05, 0.8] (the table is not
    # monotone in sigma there), or a vega pre-check below 1e-4.
    intr = np.maximum(K - S, 0.0)
    lo, _ = c4_table.price(np.stack([S, K, tau, np.full(n, 0.05), r], 1))
    hi, _ = c4_table.price(np.stack([S, K, tau, np.full(n, 0.80), r], 1))
    probe = [c4_table.price(np.stack([S, K, tau, np.full(n, v), r], 1))[1] for v in (0.10, 0.25, 0.50)]
    max_vega = np.max(np.abs(probe), axis=0)
    assert np.all((price < intr)[st == 2])
    assert np.all(((lo - price) * (hi - price) > 0)[st == 4])
    assert np.all((max_vega < 1e-4)[st == 5])
    assert np.all((max_vega >= 1e-4)[(st == 0) | (st == 4)])
    assert set(np.unique(st)) <= {0, 2, 4, 5}


def test_c4_iv_statuses(c4_table):
    base = [100.0, 100.0, 1.0, 0.05, W.C4_DIVIDEND, 10.0]
    cases = {
        "ok": (base, True, 0),
        "type mismatch": (base, False, 1),
        "dividend mismatch": (base[:4] + [0.03, 10.0], True, 1),
        "negative price": (base[:5] + [-1.0], True, 1),
        "below intrinsic": ([80.0, 100.0, 1.0, 0.05, W.C4_DIVIDEND, 19.0], True, 2),
        "above strike": (base[:5] + [101.0], True, 2),
        "maturity outside axis": ([100.0, 100.0, 3.0, 0.05, W.C4_DIVIDEND, 10.0], True, 1),
        "moneyness outside axis": ([300.0, 100.0, 1.0, 0.05, W.C4_DIVIDEND, 0.5], True, 1),
        "no bracket": (base[:5] + [40.0], True, 4),
    }
    for name, (row, put, want) in cases.items():
        _, _, _, st = c4_table.iv(np.array([row]), put=put)
        assert st[0] == want, name


def test_bspline_second_partial():
    rng = np.random.default_rng(15)
    ax = _grids(rng, (7, 6, 5, 5))
    ca = [rng.normal(size=4) for _ in range(4)]
    f = lambda a, x: ca[a][0] + x * (ca[a][1] + x * (ca[a][2] + x * ca[a][3]))
    d2f = lambda a, x: 2 * ca[a][2] + 6 * x * ca[a][3]
    X = np.meshgrid(*ax, indexing="ij")
    coef, _ = O.bspline_fit4(ax, f(0, X[0]) * f(1, X[1]) * f(2, X[2]) * f(3, X[3]))
    pts = np.stack([rng.uniform(-0.95, 0.95, 200) for _ in range(4)], 1)
    for a in range(4):
        want = np.prod([d2f(b, pts[:, b]) if b == a else f(b, pts[:, b]) for b in range(4)], axis=0)
        np.testing.assert_allclose(O.bspline_eval4_d2(ax, coef, pts, a), want, rtol=0, atol=1e-9)
    # arbitrary coefficients: central difference of the first partial
    coef = rng.normal(size=[len(a) for a in ax])
    h = 1e-6
    e = np.zeros(4)
    e[0] = h
    fd = (O.bspline_eval4(ax, coef, pts + e, 0) - O.bspline_eval4(ax, coef, pts - e, 0)) / (2 * h)
    np.testing.assert_allclose(O.bspline_eval4_d2(ax, coef, pts, 0), fd, rtol=1e-5, atol=1e-5)


def test_table_greeks_match_price_differences(c4_table):
    """delta, gamma, theta, rho of the table against central differences of
    the table price (away from the EEP floor's kink, where both are smooth)."""
    rng = np.random.default_rng(16)
    n = 60
    S = 100.0 * np.exp(rng.uniform(-0.3, 0.25, n))
    K = np.full(n, 100.0)
    tau = rng.uniform(0.3, 1.6, n)
    sig = rng.uniform(0.25, 0.6, n)
    r = rng.uniform(0.01, 0.09, n)
    rows = np.stack([S, K, tau, sig, r], 1)
    g = c4_table.greeks(rows)

    def p(dS=0.0, dt=0.0, dr=0.0):
        return c4_table.price(np.stack([S + dS, K, tau + dt, sig, r + dr], 1))[0]

    hS, ht, hr = 1e-3 * S, 1e-5, 1e-6
    delta = (p(dS=hS) - p(dS=-hS)) / (2 * hS)
    gamma = (p(dS=hS) - 2 * p() + p(dS=-hS)) / hS ** 2
    theta = -(p(dt=ht) - p(dt=-ht)) / (2 * ht)
    rho = (p(dr=hr) - p(dr=-hr)) / (2 * hr)
    np.testing.assert_allclose(g[:, 0], delta, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(g[:, 1], gamma, rtol=1e-3, atol=1e-5)
    np.testing.assert_allclose(g[:, 2], theta, rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(g[:, 3], rho, rtol=1e-4, atol=1e-4)


def test_european_greek_identities(c4_table):
    """Where the EEP reads zero the table Greeks are the Black-Scholes Greeks:
    check put-call relations of the closed forms through a call and a put
    table of zero coefficients."""
    ax = W.c4_axes()
    zero = np.zeros([len(a) for a in ax])
    put = O.Table(ax, zero, 100.0, 0.02, True)
    call = O.Table(ax, zero, 100.0, 0.02, False)
    rows = np.array([[95.0, 100.0, 0.8, 0.3, 0.04], [120.0, 100.0, 1.5, 0.2, 0.01], [70.0, 100.0, 0.1, 0.5, 0.08]])
    gp, gc = put.greeks(rows), call.greeks(rows)
    S, K, tau, r, q = rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 4], 0.02
    np.testing.assert_allclose(gc[:, 0] - gp[:, 0], np.exp(-q * tau), rtol=1e-12)
    np.testing.assert_allclose(gc[:, 1], gp[:, 1], rtol=1e-12)
    np.testing.assert_allclose(gc[:, 3] - gp[:, 3], K * tau * np.exp(-r * tau), rtol=1e-12)
    # theta of the parity C - P = S e^{-q tau} - K e^{-r tau}: d/dt = q S e^{-q tau} - r K e^{-r tau}
    np.testing.assert_allclose(gc[:, 2] - gp[:, 2], q * S * np.exp(-q * tau) - r * K * np.exp(-r * tau),
                               rtol=1e-10, atol=1e-12)
    # gamma: central difference of the closed-form delta
    h = 1e-4 * S
    up = put.greeks(np.stack([S + h, K, tau, rows[:, 3], r], 1))[:, 0]
    dn = put.greeks(np.stack([S - h, K, tau, rows[:, 3], r], 1))[:, 0]
    np.testing.assert_allclose(gp[:, 1], (up - dn) / (2 * h), rtol=1e-6)
