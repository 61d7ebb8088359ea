// SPDX-License-Identifier: MIT
//
// ORACLE — test infrastructure only. IVSolver (src/option/iv_solver.cpp:
// 20-209): implied volatility by Brent's method where every objective
// evaluation is a full American PDE solve, restated on top of the oracle's
// solver (mango_oracle.hpp) and Brent (mango_table.hpp).
//
// Pins (tests/test_oracle_fdm_iv.py): the reference's own IVSolver tests
// (tests/iv_solver_test.cc: ATM/ITM/OTM/deep-ITM put values and tolerances,
// the error codes of invalid inputs, the discrete-dividend round trip), and
// round trips of the oracle's own prices.
#pragma once

#include "mango_oracle.hpp"
#include "mango_table.hpp"

namespace oracle {

struct FdmIvResult {
    int32_t status = 0;  // MangoStatus: 0 ok, 1 validation, 2 arbitrage, 3 no convergence, 4 bracketing, 5 solver
    double implied_vol = std::numeric_limits<double>::quiet_NaN();
    double final_error = 0.0;
    size_t iterations = 0;
};

// objective_function (iv_solver.cpp:48-118): V(sigma) at the spot on the
// estimated grid, the dividend dates as mandatory points of a time domain of
// the estimate's step count, minus the market price.
inline double fdm_objective(PricingParams p, double market, double sigma) {
    p.volatility = sigma;
    auto [spec, td] = estimate_pde_grid(p);
    std::vector<double> mand = mandatory_taus(p);
    const TimeDomain t2 = mand.empty()
        ? TimeDomain::from_n_steps(0.0, p.maturity, td.n_steps)
        : TimeDomain::with_mandatory_points(0.0, p.maturity, p.maturity / static_cast<double>(td.n_steps), mand);
    if (!validate_pricing_params(p)) return std::numeric_limits<double>::quiet_NaN();
    Solution s;
    if (solve_american(p, spec, t2, s) != kOk) return std::numeric_limits<double>::quiet_NaN();
    return value_at(s, p.spot, p.strike) - market;
}

// validate_iv_query (option_spec.cpp:52-115) with validation_error_to_iv_error
// (iv_result.hpp:55-99) and the C ABI's status mapping (mango_c_api.cpp:124-141).
inline int32_t fdm_iv_validate(const PricingParams& p, double market) {
    if (p.spot <= 0.0 || !std::isfinite(p.spot)) return 1;
    if (p.strike <= 0.0 || !std::isfinite(p.strike)) return 1;
    if (p.maturity <= 0.0 || !std::isfinite(p.maturity)) return 1;
    if (!p.curve && !std::isfinite(p.rate)) return 2;
    if (p.dividend_yield < 0.0 || !std::isfinite(p.dividend_yield)) return 2;
    if (!std::isfinite(market) || market <= 0.0) return 1;
    const bool put = p.type == OptionType::PUT;
    const double intrinsic = put ? std::max(p.strike - p.spot, 0.0) : std::max(p.spot - p.strike, 0.0);
    const double upper = put ? p.strike : p.spot;
    if (market < intrinsic || market > upper) return 2;
    for (const auto& d : p.dividends) {
        if (d.calendar_time < 0.0 || d.calendar_time > p.maturity || d.amount < 0.0 ||
            !std::isfinite(d.calendar_time) || !std::isfinite(d.amount))
            return 2;
    }
    return 0;
}

// IVSolver::solve (iv_solver.cpp:124-209): adaptive upper bound from the
// time-value ratio, lower bound 0.01, Brent with max_iter and brent_tol_abs.
inline FdmIvResult fdm_iv(const PricingParams& p, double market, size_t max_iter = 100, double tol = 1e-6) {
    FdmIvResult r;
    r.status = fdm_iv_validate(p, market);
    if (r.status != 0) return r;
    const bool put = p.type == OptionType::PUT;
    const double intrinsic = put ? std::max(p.strike - p.spot, 0.0) : std::max(p.spot - p.strike, 0.0);
    const double ratio = (market - intrinsic) / market;
    const double hi = ratio > 0.5 ? 3.0 : (ratio > 0.2 ? 2.0 : 1.5);
    auto f = [&](double s) { return fdm_objective(p, market, s); };
    double root = 0.0, err = 0.0;
    size_t it = 0;
    const int32_t st = brent(f, 0.01, hi, tol, max_iter, root, it, err);
    r.iterations = it;
    r.final_error = err;
    if (st != 0) { r.status = st; return r; }
    r.implied_vol = root;
    return r;
}

}  // namespace oracle
