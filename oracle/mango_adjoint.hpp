// SPDX-License-Identifier: MIT
//
// ORACLE — test infrastructure only (see mango_oracle.hpp).
//
// Reverse-mode sensitivities of the American price on a frozen grid: the
// adjoint of the discrete TR-BDF2 march that solve_american() performs, with
// the forward's exercise sets frozen (deep-ITM locked rows and the rows the
// projected back substitution clamped to the obstacle). The reference has no
// such function; its counterparts are central finite differences of the
// forward solve on the same grid (the pin used by the tests) and, for
// dV/dK, the strike-free form of the PDE: V = K s(ln(S/K)) with s the spline
// of the normalized solution, so dV/dK = s - s' = (V - S Delta)/K.
//
// Parameters enter the discrete map only through
//   * the operator L = a d2 + b d1 + c I with a = sigma^2/2,
//     b = r - q - sigma^2/2, c = -r (black_scholes_pde.hpp:72-104), both in
//     the stage matrices A = I - w J (spatial_operator.hpp:123-164) and in the
//     stage-1 right-hand side u + w1 L(u) (pde_solver.hpp:342-362);
//   * the call's right boundary value exp(x_max) - exp(-r t)
//     (american_option.cpp:276-288).
// The reverse sweep accumulates dV/da, dV/db, dV/dc and the boundary term,
// then dV/dsigma = sigma (dV/da - dV/db), dV/dr = dV/db - dV/dc + boundary,
// dV/dq = -dV/db.
//
// Stage reverse (frozen sets), for the projected Thomas of thomas_solver.hpp:321-420:
//   back substitution u_i = max(d'_i - c'_i u_{i+1}, psi_i): carry upward in i,
//     g_i = [row i not clamped] (ubar_i + carry), carry_{i+1} = -c'_i g_i;
//   elimination d'_i = (rhs_i - l_i d'_{i-1}) / den_i, c'_i = u_i / den_i,
//     den_i = diag_i - l_i c'_{i-1}: downward in i, with
//     rhsbar_i = dbar'_i / den_i, denbar_i = -rhsbar_i d'_i - cbar'_i c'_i / den_i.
#pragma once

#include "mango_oracle.hpp"

namespace oracle {

struct AdjointResult {
    double value = 0.0, delta = 0.0;
    double dv_dsigma = 0.0, dv_drate = 0.0, dv_ddividend = 0.0, dv_dstrike = 0.0;
};

// Derivative of the natural-spline value s(xe) with respect to the knot values
// (CubicSpline::build / eval, cubic_spline_solver.hpp:74-370), scaled by `scale`
// and added to ybar.
inline void spline_value_adjoint(const std::vector<double>& x, double xe, double scale, std::vector<double>& ybar) {
    const size_t n = x.size();
    CubicSpline probe;
    probe.x = x;
    const size_t i = probe.find_interval(xe);
    const double h = x[i + 1] - x[i];
    const double dx = xe - x[i];
    ybar[i] += scale * (1.0 - dx / h);
    ybar[i + 1] += scale * (dx / h);
    if (n <= 2) return;
    // s depends on c_i, c_{i+1} (interior knots 1..n-2 solve M c = R y)
    std::vector<double> cbar(n, 0.0);
    cbar[i] += scale * (-2.0 * h * dx / 3.0 + dx * dx - dx * dx * dx / (3.0 * h));
    cbar[i + 1] += scale * (-h * dx / 3.0 + dx * dx * dx / (3.0 * h));
    const size_t m = n - 2;
    std::vector<double> hh(n - 1);
    for (size_t k = 0; k + 1 < n; ++k) hh[k] = x[k + 1] - x[k];
    // M is symmetric: diag 2(h_{k-1} + h_k), off-diagonal h_k
    std::vector<double> lo(m > 0 ? m - 1 : 0), dg(m), up(m > 0 ? m - 1 : 0), rhs(m), z(m), ws(2 * m);
    for (size_t r = 0; r < m; ++r) {
        const size_t k = r + 1;
        dg[r] = 2.0 * (hh[k - 1] + hh[k]);
        if (r > 0) lo[r - 1] = hh[k - 1];
        if (r < m - 1) up[r] = hh[k];
        rhs[r] = cbar[k];
    }
    solve_thomas(lo.data(), dg.data(), up.data(), rhs.data(), z.data(), m, ws.data());
    // rhs_k = 3 ((y_{k+1} - y_k)/h_k - (y_k - y_{k-1})/h_{k-1})
    for (size_t r = 0; r < m; ++r) {
        const size_t k = r + 1;
        ybar[k + 1] += 3.0 * z[r] / hh[k];
        ybar[k] -= 3.0 * z[r] * (1.0 / hh[k] + 1.0 / hh[k - 1]);
        ybar[k - 1] += 3.0 * z[r] / hh[k - 1];
    }
}

inline int32_t adjoint_american(const PricingParams& p, const GridSpec& spec, const TimeDomain& td,
                                AdjointResult& out) {
    Solution sol;
    StageTape tape;
    const int32_t st = solve_american(p, spec, td, sol, {}, &tape);
    if (st != kOk) return st;
    const Greeks g = compute_greeks(sol, p.spot, p.strike);
    out = AdjointResult{};
    out.value = g.value;
    out.delta = g.delta;
    out.dv_dstrike = (g.value - p.spot * g.delta) / p.strike;

    const std::vector<double>& x = sol.x;
    const size_t n = x.size();
    const double K = p.strike;
    const bool put = p.type == OptionType::PUT;
    const double r = p.rate;
    auto dgr_dr = [&](double t) { return put ? 0.0 : t * std::exp(-r * t); };

    // seed: adjoint of the final solution for V = K max(0, s(ln(S/K)))
    std::vector<double> ub(n, 0.0);
    const double xs = std::log(p.spot / p.strike);
    if (xs <= x[0]) ub[0] = K;
    else if (xs >= x[n - 1]) ub[n - 1] = K;
    else {
        CubicSpline spl;
        spl.build(x.data(), sol.u.data(), n);
        if (spl.eval(xs) > 0.0) spline_value_adjoint(x, xs, K, ub);
    }

    // second- and first-derivative stencil weights per interior node (the
    // Jacobian formulas of spatial_operator.hpp:123-164 with a = 1 or b = 1)
    std::vector<double> d2l(n, 0.0), d2d(n, 0.0), d2u(n, 0.0), d1l(n, 0.0), d1d(n, 0.0), d1u(n, 0.0);
    for (size_t i = 1; i + 1 < n; ++i) {
        const double dl = x[i] - x[i - 1], dr = x[i + 1] - x[i], dxa = (dl + dr) / 2.0;
        d2l[i] = 1.0 / (dl * dxa);
        d2d[i] = -(1.0 / dl + 1.0 / dr) / dxa;
        d2u[i] = 1.0 / (dr * dxa);
        d1l[i] = -dr / (dl * (dl + dr));
        d1d[i] = (dr - dl) / (dl * dr);
        d1u[i] = dl / (dr * (dl + dr));
    }
    const double a = 0.5 * p.volatility * p.volatility;
    const double b = p.rate - p.dividend_yield - a;
    const double c = -p.rate;
    auto jl = [&](size_t i) { return a * d2l[i] + b * d1l[i]; };
    auto jd = [&](size_t i) { return a * d2d[i] + b * d1d[i] + c; };
    auto ju = [&](size_t i) { return a * d2u[i] + b * d1u[i]; };

    double SA = 0.0, SB = 0.0, SC = 0.0, Rbc = 0.0;
    std::vector<double> pending(n, 0.0);  // beta-term adjoint of u^n from stage 2
    std::vector<double> gg(n), dpb(n), cpb(n), lob(n, 0.0), dgb(n), upb(n), rhb(n), nu(n);
    const Spacing spacing(x);
    std::vector<double> d2u_n(n, 0.0), d1u_n(n, 0.0);

    for (size_t si = tape.stages.size(); si-- > 0;) {
        const StageRecord& R = tape.stages[si];
        // boundary values re-applied after the solve
        Rbc += ub[n - 1] * dgr_dr(R.t);
        ub[0] = 0.0;
        ub[n - 1] = 0.0;
        // back substitution, reversed (carry upward)
        std::fill(cpb.begin(), cpb.end(), 0.0);
        double carry = 0.0;
        for (size_t i = 0; i < n; ++i) {
            const double tot = ub[i] + carry;
            gg[i] = R.clamped[i] ? 0.0 : tot;
            dpb[i] = gg[i];
            if (i + 1 < n) {
                cpb[i] = -gg[i] * R.sol[i + 1];
                carry = -R.cp[i] * gg[i];
            }
        }
        // elimination, reversed
        std::fill(lob.begin(), lob.end(), 0.0);
        for (size_t i = n; i-- > 0;) {
            double den;
            if (i == 0) den = R.diag[0];
            else den = R.diag[i] - R.lower[i - 1] * R.cp[i - 1];
            const double delta = dpb[i] / den;
            rhb[i] = delta;
            double denbar = -delta * R.dp[i];
            if (i + 1 < n) {
                upb[i] = cpb[i] / den;
                denbar -= cpb[i] * R.cp[i] / den;
            } else {
                upb[i] = 0.0;
            }
            dgb[i] = denbar;
            if (i > 0) {
                lob[i - 1] += -delta * R.dp[i - 1] - denbar * R.cp[i - 1];
                dpb[i - 1] += -delta * R.lower[i - 1];
                cpb[i - 1] += -denbar * R.lower[i - 1];
            }
        }
        // parameters through the stage matrix (free interior rows only)
        for (size_t i = 1; i + 1 < n; ++i) {
            if (!R.free_row[i]) continue;
            const double jlb = -R.w * lob[i - 1];
            const double jdb = -R.w * dgb[i];
            const double jub = -R.w * upb[i];
            SA += jlb * d2l[i] + jdb * d2d[i] + jub * d2u[i];
            SB += jlb * d1l[i] + jdb * d1d[i] + jub * d1u[i];
            SC += jdb;
        }
        // boundary rows of the right-hand side
        Rbc += rhb[n - 1] * dgr_dr(R.t);
        // free rows of the right-hand side -> stage inputs
        for (size_t i = 0; i < n; ++i) rhb[i] = R.free_row[i] ? rhb[i] : 0.0;
        if (R.kind == 2) {
            const double gamma = 2.0 - std::sqrt(2.0);
            const double den = gamma * (2.0 - gamma);
            const double alpha = 1.0 / den;
            const double beta = -((1.0 - gamma) * (1.0 - gamma)) / den;
            for (size_t i = 0; i < n; ++i) {
                ub[i] = alpha * rhb[i];
                pending[i] = beta * rhb[i];
            }
        } else if (R.kind == 1) {
            // rhs = u^n + w1 L(u^n); w1 is the stage weight
            const double w1 = R.w;
            std::fill(d2u_n.begin(), d2u_n.end(), 0.0);
            std::fill(d1u_n.begin(), d1u_n.end(), 0.0);
            spacing.second_derivative(R.u_n.data(), d2u_n.data(), 1, n - 1);
            spacing.first_derivative(R.u_n.data(), d1u_n.data(), 1, n - 1);
            std::fill(nu.begin(), nu.end(), 0.0);
            for (size_t i = 1; i + 1 < n; ++i) {
                const double y = rhb[i];
                if (y == 0.0) continue;
                SA += w1 * y * d2u_n[i];
                SB += w1 * y * d1u_n[i];
                SC += w1 * y * R.u_n[i];
                nu[i - 1] += w1 * y * jl(i);
                nu[i] += y + w1 * y * jd(i);
                nu[i + 1] += w1 * y * ju(i);
            }
            for (size_t i = 0; i < n; ++i) {
                ub[i] = pending[i] + nu[i];
                pending[i] = 0.0;
            }
        } else {
            ub = rhb;
        }
    }
    const double sigma = p.volatility;
    out.dv_dsigma = sigma * (SA - SB);
    out.dv_drate = SB - SC + Rbc;
    out.dv_ddividend = -SB;
    return kOk;
}

}  // namespace oracle
