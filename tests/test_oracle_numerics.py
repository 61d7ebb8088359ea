# SPDX-License-Identifier: MIT
"""Pins of the oracle's numerical kernels (CPU, no GPU).

Known-answer tests are the reference's own (file:line cited per test); the
rest pin what mathematics fixes: dense solves, LCP active-set enumeration,
exactness of 3-point stencils on quadratics, spline reproduction of lines.
"""
import itertools

import numpy as np
import pytest

import pyoracle as O


# --- Thomas (src/math/thomas_solver.hpp:86-198) --------------------------------------------
def test_thomas_3x3_kat():  # tests/tridiagonal_solver_test.cc:7-30
    ok, x = O.thomas([1.0, 1.0], [2.0, 2.0, 2.0], [1.0, 1.0], [1.0, 0.0, 1.0])
    assert ok
    np.testing.assert_allclose(x, [1.0, -1.0, 1.0], atol=1e-10)


def test_thomas_kat_spline_file():  # tests/thomas_cubic_spline_test.cc:19-50
    ok, x = O.thomas([1.0, 1.0], [2.0, 3.0, 2.0], [1.0, 1.0], [1.0, 2.0, 3.0])
    assert ok
    # residual of the system the reference solves
    A = np.diag([2.0, 3.0, 2.0]) + np.diag([1.0, 1.0], 1) + np.diag([1.0, 1.0], -1)
    np.testing.assert_allclose(A @ x, [1.0, 2.0, 3.0], atol=1e-12)


def test_thomas_singular_detected():  # tests/tridiagonal_solver_test.cc:32-47
    ok, _ = O.thomas([1.0], [0.0, 0.0], [1.0], [1.0, 1.0])
    assert not ok


def test_thomas_heat_row_residual():  # tests/tridiagonal_solver_test.cc:49-71
    n = 5
    lo, dg, up = np.full(n - 1, -1.0), np.full(n, 3.0), np.full(n - 1, -1.0)
    rhs = np.array([1.0, 2.0, 3.0, 2.0, 1.0])
    ok, x = O.thomas(lo, dg, up, rhs)
    assert ok
    assert abs(lo[1] * x[1] + dg[2] * x[2] + up[2] * x[3] - rhs[2]) < 1e-9


@pytest.mark.parametrize("n", [2, 3, 5, 9, 16])
def test_thomas_matches_dense_solve(n):
    rng = np.random.default_rng(n)
    lo, up = rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n - 1)
    dg = 3.0 + rng.uniform(0, 1, n)
    rhs = rng.normal(size=n)
    A = np.diag(dg) + np.diag(up, 1) + np.diag(lo, -1)
    ok, x = O.thomas(lo, dg, up, rhs)
    assert ok
    np.testing.assert_allclose(x, np.linalg.solve(A, rhs), rtol=1e-12, atol=1e-12)


# --- projected Thomas (src/math/thomas_solver.hpp:321-420) ---------------------------------
def test_projected_no_active_constraints():  # tests/tridiagonal_solver_test.cc:103-121
    ok, x = O.thomas_projected([1, 1], [2, 2, 2], [1, 1], [1, 0, 1], [-100] * 3)
    assert ok
    np.testing.assert_allclose(x, [1, -1, 1], atol=1e-10)


def test_projected_all_active():  # tests/tridiagonal_solver_test.cc:127-148
    ok, x = O.thomas_projected([1, 1], [2, 2, 2], [1, 1], [1, 0, 1], [100] * 3)
    assert ok
    np.testing.assert_allclose(x, [100] * 3, atol=1e-10)


def test_projected_partial():  # tests/tridiagonal_solver_test.cc:150-166
    ok, x = O.thomas_projected([1, 1], [2, 2, 2], [1, 1], [1, 0, 1], [-100, 5, -100])
    assert ok and x[1] >= 5 - 1e-10


def test_projected_respects_oscillating_bound():  # tests/tridiagonal_solver_test.cc:168-190
    n = 20
    psi = 0.3 * np.sin(np.arange(n))
    ok, x = O.thomas_projected(np.full(n - 1, -0.5), np.full(n, 2.0), np.full(n - 1, -0.5), np.ones(n), psi)
    assert ok and (x >= psi - 1e-10).all()


def test_projected_singular():  # tests/tridiagonal_solver_test.cc:198-217
    ok, _ = O.thomas_projected([1.0], [0.0, 0.0], [1.0], [1.0, 1.0], [0.0, 0.0])
    assert not ok


def _lcp_bruteforce(A, b, psi):
    """u >= psi, A u - b >= 0, (u - psi)'(A u - b) = 0, by active-set enumeration."""
    n = len(b)
    for mask in itertools.product([0, 1], repeat=n):
        act = np.array(mask, dtype=bool)
        u = psi.copy()
        free = ~act
        if free.any():
            Aff = A[np.ix_(free, free)]
            rhs = b[free] - A[np.ix_(free, act)] @ psi[act]
            u[free] = np.linalg.solve(Aff, rhs)
        w = A @ u - b
        if (u >= psi - 1e-12).all() and (w >= -1e-12).all() and abs((u - psi) @ w) < 1e-10:
            return u
    raise AssertionError("no LCP solution")


def _mmatrix_case(seed, n=8):
    rng = np.random.default_rng(seed)
    k = rng.uniform(0.2, 1.5)
    dg = np.full(n, 1 + 2 * k)
    lo = np.full(n - 1, -k * rng.uniform(0.8, 1.2))
    up = np.full(n - 1, -k * rng.uniform(0.8, 1.2))
    return rng, dg, lo, up, np.diag(dg) + np.diag(up, 1) + np.diag(lo, -1)


@pytest.mark.parametrize("seed", range(6))
def test_projected_equals_lcp_when_contact_is_right_block(seed):
    # Top-down elimination then bottom-up projection (the reference's order)
    # solves the LCP exactly when the contact set is a right-hand block
    # (call-like obstacle): rows left of it are genuine equations.
    rng, dg, lo, up, A = _mmatrix_case(seed)
    xs = np.linspace(-1, 1, len(dg))
    psi = np.maximum(np.exp(xs) - 1, 0.0)
    b = psi * rng.uniform(0.2, 0.9) + 0.02
    ok, x = O.thomas_projected(lo, dg, up, b, psi)
    assert ok
    np.testing.assert_allclose(x, _lcp_bruteforce(A, b, psi), atol=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_projected_put_obstacle_is_feasible_not_lcp(seed):
    # For a left contact block the same order is only a feasible projection
    # (u >= psi everywhere), not the LCP solution: this is the property the
    # reference compensates with its deep-ITM row locking
    # (src/pde/internal/pde_solver.hpp:631-680). Pinned so that a change of
    # elimination order would be noticed.
    rng, dg, lo, up, A = _mmatrix_case(seed)
    xs = np.linspace(-1, 1, len(dg))
    psi = np.maximum(1 - np.exp(xs), 0.0)
    b = psi * rng.uniform(0.2, 0.9) + 0.02
    ok, x = O.thomas_projected(lo, dg, up, b, psi)
    assert ok and (x >= psi).all()
    lcp = _lcp_bruteforce(A, b, psi)
    assert np.max(np.abs(x - lcp)) > 1e-6


# --- cubic spline (src/math/cubic_spline_solver.hpp:74-370) ---------------------------------
def test_spline_reproduces_line():  # tests/thomas_cubic_spline_test.cc:145-168
    ok, v, d = O.spline_eval([0.0, 1.0, 2.0, 3.0], [0.0, 2.0, 4.0, 6.0], [0.5, 1.5, 2.5])
    assert ok
    np.testing.assert_allclose(v, [1.0, 3.0, 5.0], atol=1e-10)
    np.testing.assert_allclose(d, [2.0, 2.0, 2.0], atol=1e-10)


def test_spline_quadratic_tolerances():  # tests/thomas_cubic_spline_test.cc:170-194
    x = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    ok, v, d = O.spline_eval(x, x * x, [0.5, 1.5, 2.5, 1.0, 2.0, 3.0])
    assert ok
    assert abs(v[0] - 0.25) <= 0.1 and abs(v[1] - 2.25) <= 0.05 and abs(v[2] - 6.25) <= 0.05
    np.testing.assert_allclose(d[3:], [2.0, 4.0, 6.0], atol=0.2)


def test_spline_interpolates_knots_and_sine():  # tests/thomas_cubic_spline_test.cc:196-230
    x = np.linspace(0, np.pi, 50)
    ok, v, _ = O.spline_eval(x, np.sin(x), x)
    np.testing.assert_allclose(v, np.sin(x), atol=1e-14)
    xe = np.linspace(0.3, 2.8, 37)
    ok, v, d = O.spline_eval(x, np.sin(x), xe)
    np.testing.assert_allclose(v, np.sin(xe), atol=1e-4)
    np.testing.assert_allclose(d, np.cos(xe), atol=1e-3)


# --- stencil and analytic Jacobian ----------------------------------------------------------
def _sinh_grid(x_min, x_max, n, c):
    from mango_option._abi import MangoGridSpec, MANGO_GRID_SINH
    s = MangoGridSpec()
    s.type = MANGO_GRID_SINH; s.x_min = x_min; s.x_max = x_max; s.n_points = n; s.concentration = c
    return O.generate(s)


def test_jacobian_matches_stencil():  # tests/internal/spatial_operator_jacobian_test.cc:26-208
    x = _sinh_grid(-1.0, 1.0, 21, 4.0)
    p = O.one_param(1, 1, 1, 0.30, 0.10, 0.01, True)
    rng = np.random.default_rng(0)
    for _ in range(5):
        u = rng.normal(size=len(x))
        lu, jl, jd, ju = O.operator(p, x, u)
        via_bands = jl[1:-1] * u[:-2] + jd[1:-1] * u[1:-1] + ju[1:-1] * u[2:]
        np.testing.assert_allclose(lu[1:-1], via_bands, rtol=1e-12, atol=1e-11)


def test_stencil_exact_on_quadratics():
    x = _sinh_grid(-2.0, 1.5, 31, 3.0)
    sig, r, q = 0.25, 0.05, 0.02
    p = O.one_param(1, 1, 1, sig, r, q, True)
    u = 0.7 * x * x - 0.3 * x + 0.2
    lu, *_ = O.operator(p, x, u)
    hss = 0.5 * sig * sig
    exact = hss * 1.4 + (r - q - hss) * (1.4 * x - 0.3) - r * u
    np.testing.assert_allclose(lu[1:-1], exact[1:-1], rtol=1e-10, atol=1e-10)
