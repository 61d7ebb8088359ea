// SPDX-License-Identifier: MIT
//
// FDM implied volatility (IVSolver, src/option/iv_solver.cpp:20-209): Brent's
// method (src/math/root_finding.hpp:145-296) where each objective value
// V(sigma) - market is a full American PDE solve on the estimated grid.
//
// Each query's Brent iteration is a small state machine on the host; a round
// collects the sigma every unfinished query needs next and prices them all in
// one batched GPU solve (mango_price_american_batch_ex with the default grid
// accuracy, per-option grids: exactly IVSolver::objective_function's
// PDEGridConfig{estimated grid, estimated step count, dividend dates}).
#include "../../include/mango_amd.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <string>
#include <utility>
#include <vector>

#include "mango_events.hpp"
#include "mango_host.hpp"

namespace mango_amd {
namespace {

constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();

// mango_c_api.cpp:388-437 pre-validation, then validate_iv_query
// (option_spec.cpp:52-115) with validation_error_to_iv_error (iv_result.hpp:
// 55-99) and map_iv_error (mango_c_api.cpp:124-141).
MangoStatus validate_fdm_query(const MangoIvQuery& q, const char** why) {
    auto bad = [&](MangoStatus st, const char* m) { *why = m; return st; };
    if (q.option_type != MANGO_CALL && q.option_type != MANGO_PUT) return bad(MANGO_ERR_VALIDATION, "invalid option_type");
    if (!validate_curve(q.tenor_points, q.n_tenor_points, why)) return MANGO_ERR_VALIDATION;
    if (q.dividends == nullptr && q.n_dividends > 0) return bad(MANGO_ERR_VALIDATION, "dividends is null but n_dividends > 0");
    if (!std::isfinite(q.dividend_yield)) return bad(MANGO_ERR_VALIDATION, "dividend_yield must be finite");
    if (q.n_tenor_points == 0 && !std::isfinite(q.rate_const)) return bad(MANGO_ERR_VALIDATION, "rate must be finite");
    for (uint64_t i = 0; i < q.n_dividends; ++i) {
        const double ct = q.dividends[i].calendar_time, amt = q.dividends[i].amount;
        if (!std::isfinite(ct) || ct < 0.0 || ct > q.maturity || !std::isfinite(amt) || amt < 0.0)
            return bad(MANGO_ERR_VALIDATION, "invalid discrete dividend");
    }
    if (q.spot <= 0.0 || !std::isfinite(q.spot)) return bad(MANGO_ERR_VALIDATION, "Invalid spot price");
    if (q.strike <= 0.0 || !std::isfinite(q.strike)) return bad(MANGO_ERR_VALIDATION, "Invalid strike");
    if (q.maturity <= 0.0 || !std::isfinite(q.maturity)) return bad(MANGO_ERR_VALIDATION, "Invalid maturity");
    if (q.dividend_yield < 0.0) return bad(MANGO_ERR_ARBITRAGE, "Invalid dividend yield");
    if (!std::isfinite(q.market_price) || q.market_price <= 0.0) return bad(MANGO_ERR_VALIDATION, "Invalid market price");
    const bool put = q.option_type == MANGO_PUT;
    const double intrinsic = put ? std::max(q.strike - q.spot, 0.0) : std::max(q.spot - q.strike, 0.0);
    const double upper = put ? q.strike : q.spot;
    if (q.market_price < intrinsic || q.market_price > upper)
        return bad(MANGO_ERR_ARBITRAGE, "Market price violates arbitrage bounds");
    *why = nullptr;
    return MANGO_OK;
}

// brent_find_root as a resumable state machine: `want` is the next abscissa
// to evaluate (NaN when finished).
struct Brent {
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0, fa = 0.0, fb = 0.0, fc = 0.0, s = 0.0;
    bool mflag = true;
    int phase = 0;  // 0: f(a) pending, 1: f(b) pending, 2: f(s) pending, 3: done
    size_t iter = 0, max_iter = 100;
    double tol = 1e-6;
    MangoStatus status = MANGO_OK;
    double root = kNaN, final_error = 0.0;
    size_t iterations = 0;

    double want() const { return phase == 0 ? a : (phase == 1 ? b : (phase == 2 ? s : kNaN)); }

    void finish(MangoStatus st, double r, size_t it, double err) {
        status = st; root = r; iterations = it; final_error = err; phase = 3;
    }
    // one Brent iteration up to the next evaluation (iv loop head)
    void step() {
        if (iter >= max_iter) { finish(MANGO_ERR_NO_CONVERGENCE, b, max_iter, std::fabs(fb)); return; }
        if (std::fabs(fb) < tol || std::fabs(b - a) < tol) { finish(MANGO_OK, b, iter + 1, std::fabs(fb)); return; }
        double t;
        if (fa != fc && fb != fc)
            t = a * fb * fc / ((fa - fb) * (fa - fc)) + b * fa * fc / ((fb - fa) * (fb - fc)) +
                c * fa * fb / ((fc - fa) * (fc - fb));
        else
            t = b - fb * (b - a) / (fb - fa);
        const double bisect = (3.0 * a + b) / 4.0;
        const bool c1 = !((t > bisect && t < b) || (t < bisect && t > b));
        const bool c2 = mflag && std::fabs(t - b) >= std::fabs(b - c) / 2.0;
        const bool c3 = !mflag && std::fabs(t - b) >= std::fabs(c - d) / 2.0;
        const bool c4 = mflag && std::fabs(b - c) < tol;
        const bool c5 = !mflag && std::fabs(c - d) < tol;
        if (c1 || c2 || c3 || c4 || c5) { t = (a + b) / 2.0; mflag = true; }
        else mflag = false;
        s = t;
        phase = 2;
    }
    void feed(double f) {
        if (phase == 0) { fa = f; phase = 1; return; }
        if (phase == 1) {
            fb = f;
            if (!std::isfinite(fa) || !std::isfinite(fb)) { finish(MANGO_ERR_SOLVER, kNaN, 0, kNaN); return; }
            if (fa * fb > 0.0) { finish(MANGO_ERR_BRACKETING, kNaN, 0, std::min(std::fabs(fa), std::fabs(fb))); return; }
            if (std::fabs(fa) < tol) { finish(MANGO_OK, a, 0, std::fabs(fa)); return; }
            if (std::fabs(fb) < tol) { finish(MANGO_OK, b, 0, std::fabs(fb)); return; }
            if (std::fabs(fa) < std::fabs(fb)) { std::swap(a, b); std::swap(fa, fb); }
            c = a; fc = fa; mflag = true; d = 0.0; iter = 0;
            step();
            return;
        }
        // phase 2: f(s)
        if (!std::isfinite(f)) { finish(MANGO_ERR_SOLVER, kNaN, iter + 1, kNaN); return; }
        d = c; c = b; fc = fb;
        if (fa * f < 0.0) { b = s; fb = f; } else { a = s; fa = f; }
        if (std::fabs(fa) < std::fabs(fb)) { std::swap(a, b); std::swap(fa, fb); }
        ++iter;
        step();
    }
};

const char* iv_message(MangoStatus st) {
    switch (st) {
        case MANGO_ERR_BRACKETING: return "Root not bracketed: market price outside the model's range on [0.01, upper]";
        case MANGO_ERR_NO_CONVERGENCE: return "Brent's method did not converge within max_iter";
        case MANGO_ERR_SOLVER: return "Numerical instability: the PDE solve failed for a trial volatility";
        default: return "IV solve failed";
    }
}

// Solves n queries; slots receive status and success. Returns a call-level
// error (device, allocation) or MANGO_OK.
MangoStatus fdm_iv_run(const MangoIvQuery* qs, uint64_t n, const MangoIvConfig* cfg, MangoIvBatchSlot* slots,
                       const char** first_why, MangoError* err) {
    size_t max_iter = 100;
    double tol = 1e-6;
    if (cfg) {
        if (cfg->max_iter > 0) max_iter = static_cast<size_t>(cfg->max_iter);
        if (cfg->brent_tol_abs > 0.0) tol = cfg->brent_tol_abs;
    }
    std::vector<Brent> st(n);
    std::vector<char> live(n, 0);
    for (uint64_t i = 0; i < n; ++i) {
        const char* why = nullptr;
        const MangoStatus v = validate_fdm_query(qs[i], &why);
        slots[i].status = v;
        slots[i].success = MangoIvSuccess{};
        if (v != MANGO_OK) {
            if (first_why && !*first_why) *first_why = why;
            continue;
        }
        // adaptive bounds (iv_solver.cpp:129-149)
        const bool put = qs[i].option_type == MANGO_PUT;
        const double intrinsic = put ? std::max(qs[i].strike - qs[i].spot, 0.0) : std::max(qs[i].spot - qs[i].strike, 0.0);
        const double ratio = (qs[i].market_price - intrinsic) / qs[i].market_price;
        st[i].a = 0.01;
        st[i].b = ratio > 0.5 ? 3.0 : (ratio > 0.2 ? 2.0 : 1.5);
        st[i].max_iter = max_iter;
        st[i].tol = tol;
        live[i] = 1;
    }
    std::vector<MangoPricingParams> ps;
    std::vector<uint64_t> who;
    std::vector<MangoAmericanGreeks> out;
    std::vector<int32_t> ost;
    for (;;) {
        ps.clear();
        who.clear();
        for (uint64_t i = 0; i < n; ++i) {
            if (!live[i]) continue;
            // the first round evaluates both bracket ends
            const int evals = st[i].phase == 0 ? 2 : 1;
            for (int e = 0; e < evals; ++e) {
                MangoPricingParams p{};
                p.spot = qs[i].spot; p.strike = qs[i].strike; p.maturity = qs[i].maturity;
                p.dividend_yield = qs[i].dividend_yield; p.rate_const = qs[i].rate_const;
                p.tenor_points = qs[i].tenor_points; p.n_tenor_points = qs[i].n_tenor_points;
                p.dividends = qs[i].dividends; p.n_dividends = qs[i].n_dividends;
                p.option_type = qs[i].option_type;
                p.volatility = e == 0 ? st[i].want() : st[i].b;
                ps.push_back(p);
                who.push_back(i);
            }
        }
        if (ps.empty()) break;
        out.assign(ps.size(), MangoAmericanGreeks{});
        ost.assign(ps.size(), 0);
        uint64_t failed = 0;
        const MangoStatus rc = mango_price_american_batch_ex(ps.data(), ps.size(), 0, 0, nullptr, out.data(), ost.data(),
                                                             &failed, err);
        if (rc != MANGO_OK) return rc;
        for (size_t k = 0; k < ps.size(); ++k) {
            Brent& b = st[who[k]];
            const double f = ost[k] == MANGO_OK ? out[k].value - qs[who[k]].market_price : kNaN;
            b.feed(f);
        }
        for (uint64_t i = 0; i < n; ++i) {
            if (!live[i] || st[i].phase != 3) continue;
            live[i] = 0;
            slots[i].status = st[i].status;
            slots[i].success.iterations = st[i].iterations;
            slots[i].success.final_error = st[i].final_error;
            if (st[i].status == MANGO_OK) slots[i].success.implied_vol = st[i].root;
            else if (first_why && !*first_why) *first_why = iv_message(st[i].status);
        }
    }
    return MANGO_OK;
}

}  // namespace
}  // namespace mango_amd

using namespace mango_amd;

extern "C" {

MangoStatus mango_solve_iv(const MangoIvQuery* query, const MangoIvConfig* config, MangoIvSuccess* out_success,
                           MangoError* out_err) {
    if (!query || !out_success) {
        set_err(out_err, MANGO_ERR_VALIDATION, "null query or out_success");
        return MANGO_ERR_VALIDATION;
    }
    if (config && config->max_iter < 0) {
        set_err(out_err, MANGO_ERR_VALIDATION, "max_iter must be >= 0");
        return MANGO_ERR_VALIDATION;
    }
    try {
        MangoIvBatchSlot slot{};
        const char* why = nullptr;
        const MangoStatus rc = fdm_iv_run(query, 1, config, &slot, &why, out_err);
        if (rc != MANGO_OK) return rc;
        if (slot.status != MANGO_OK) {
            set_err(out_err, static_cast<MangoStatus>(slot.status), why ? why : "IV solve failed");
            return static_cast<MangoStatus>(slot.status);
        }
        *out_success = slot.success;
        return MANGO_OK;
    } catch (const std::exception& ex) {
        set_err(out_err, MANGO_ERR_SOLVER, ex.what());
        return MANGO_ERR_SOLVER;
    }
}

MangoStatus mango_solve_iv_batch(const MangoIvQuery* queries, uint64_t n, const MangoIvConfig* config,
                                 MangoIvBatchSlot* out_slots, uint64_t* out_failed_count, MangoError* err) {
    if ((!queries || !out_slots) && n > 0) {
        set_err(err, MANGO_ERR_VALIDATION, "null queries or out_slots");
        return MANGO_ERR_VALIDATION;
    }
    if (config && config->max_iter < 0) {
        set_err(err, MANGO_ERR_VALIDATION, "max_iter must be >= 0");
        return MANGO_ERR_VALIDATION;
    }
    if (out_failed_count) *out_failed_count = 0;
    if (n == 0) return MANGO_OK;
    try {
        const MangoStatus rc = fdm_iv_run(queries, n, config, out_slots, nullptr, err);
        if (rc != MANGO_OK) return rc;
        if (out_failed_count) {
            uint64_t failed = 0;
            for (uint64_t i = 0; i < n; ++i) failed += out_slots[i].status != MANGO_OK;
            *out_failed_count = failed;
        }
        return MANGO_OK;
    } catch (const std::exception& ex) {
        set_err(err, MANGO_ERR_SOLVER, ex.what());
        return MANGO_ERR_SOLVER;
    }
}

}  // extern "C"
