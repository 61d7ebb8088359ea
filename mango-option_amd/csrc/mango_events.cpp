// SPDX-License-Identifier: MIT
//
// Host side of yield curves, mandatory time points and discrete dividends
// (declarations and reference citations in mango_events.hpp).
#include "mango_events.hpp"

#include <algorithm>
#include <cmath>
#include <utility>

namespace mango_amd {

HostTimeDomain time_uniform(double t0, double t1, int64_t n) {
    HostTimeDomain td;
    td.n_steps = n;
    td.dt_avg = (t1 - t0) / static_cast<double>(n);
    return td;
}

HostTimeDomain time_with_mandatory(double t0, double t1, double dt, std::vector<double> mandatory) {
    std::vector<double> breaks;
    for (double t : mandatory)
        if (t > t0 + 1e-14 && t < t1 - 1e-14) breaks.push_back(t);
    std::sort(breaks.begin(), breaks.end());
    breaks.erase(std::unique(breaks.begin(), breaks.end(), [](double a, double b) { return std::abs(a - b) < 1e-14; }),
                 breaks.end());
    if (breaks.empty()) return time_uniform(t0, t1, static_cast<int64_t>(std::ceil((t1 - t0) / dt)));
    std::vector<double> bounds{t0};
    bounds.insert(bounds.end(), breaks.begin(), breaks.end());
    bounds.push_back(t1);
    HostTimeDomain td;
    for (size_t s = 0; s + 1 < bounds.size(); ++s) {
        const double len = bounds[s + 1] - bounds[s];
        const int64_t n_sub = std::max<int64_t>(1, static_cast<int64_t>(std::ceil(len / dt)));
        td.seg_start.push_back(bounds[s]);
        td.sub_dt.push_back(len / static_cast<double>(n_sub));
        td.n_sub.push_back(n_sub);
        td.n_steps += n_sub;
    }
    td.dt_avg = (t1 - t0) / static_cast<double>(td.n_steps);
    return td;
}

bool has_events(const MangoPricingParams& p) { return p.n_tenor_points > 0 || p.n_dividends > 0; }

std::vector<double> mandatory_taus(const MangoPricingParams& p) {
    std::vector<double> out;
    for (uint64_t i = 0; i < p.n_dividends && p.dividends; ++i) {
        const double tau = p.maturity - p.dividends[i].calendar_time;
        if (tau > 0.0 && tau < p.maturity) out.push_back(tau);
    }
    return out;
}

double max_dividend_over_strike(const MangoPricingParams& p) {
    double m = 0.0;
    for (uint64_t i = 0; i < p.n_dividends && p.dividends; ++i) {
        const double tau = p.maturity - p.dividends[i].calendar_time;
        if (tau > 0.0 && tau < p.maturity) m = std::max(m, p.dividends[i].amount / p.strike);
    }
    return m;
}

namespace {
std::vector<MangoTenorPoint> sorted_curve(const MangoPricingParams& p) {
    std::vector<MangoTenorPoint> v(p.tenor_points, p.tenor_points + p.n_tenor_points);
    std::sort(v.begin(), v.end(), [](const MangoTenorPoint& a, const MangoTenorPoint& b) { return a.tenor < b.tenor; });
    return v;
}
}  // namespace

bool validate_curve(const MangoTenorPoint* pts, uint64_t n, const char** why) {
    if (n == 0) return true;
    if (!pts) { *why = "tenor_points is null but n_tenor_points > 0"; return false; }
    std::vector<MangoTenorPoint> v(pts, pts + n);
    std::sort(v.begin(), v.end(), [](const MangoTenorPoint& a, const MangoTenorPoint& b) { return a.tenor < b.tenor; });
    if (v[0].tenor != 0.0) { *why = "invalid yield curve: First point must have t=0"; return false; }
    if (std::abs(v[0].log_discount) > 1e-10) { *why = "invalid yield curve: log_discount at t=0 must be 0"; return false; }
    for (size_t i = 1; i < v.size(); ++i)
        if (v[i].tenor <= v[i - 1].tenor + 1e-10) { *why = "invalid yield curve: Tenors must be strictly increasing"; return false; }
    return true;
}

double curve_zero_rate(const MangoTenorPoint* pts, uint64_t n, double T) {
    std::vector<MangoTenorPoint> v(pts, pts + n);
    std::sort(v.begin(), v.end(), [](const MangoTenorPoint& a, const MangoTenorPoint& b) { return a.tenor < b.tenor; });
    const size_t m = v.size();
    auto rate_between = [&](size_t i) {
        if (i + 1 >= m) return 0.0;
        const double dt = v[i + 1].tenor - v[i].tenor;
        return dt <= 0.0 ? 0.0 : -(v[i + 1].log_discount - v[i].log_discount) / dt;
    };
    if (T <= 0.0) return m < 2 ? 0.0 : rate_between(0);
    if (m < 2) return -0.0 / T;
    const size_t ub = static_cast<size_t>(std::upper_bound(v.begin(), v.end(), T, [](double t, const MangoTenorPoint& q) {
                                              return t < q.tenor; }) - v.begin());
    double ld;
    if (ub == 0) ld = 0.0;
    else if (ub == m) {
        const double r = -(v[m - 1].log_discount - v[m - 2].log_discount) / (v[m - 1].tenor - v[m - 2].tenor);
        ld = v[m - 1].log_discount - r * (T - v[m - 1].tenor);
    } else {
        const double a = (T - v[ub - 1].tenor) / (v[ub].tenor - v[ub - 1].tenor);
        ld = v[ub - 1].log_discount + a * (v[ub].log_discount - v[ub - 1].log_discount);
    }
    return -ld / T;
}

MangoStatus validate_events(const MangoPricingParams& p, const char** why) {
    auto bad = [&](const char* m) { *why = m; return MANGO_ERR_VALIDATION; };
    if (p.n_tenor_points > 0 && !validate_curve(p.tenor_points, p.n_tenor_points, why)) return MANGO_ERR_VALIDATION;
    if (p.n_dividends > 0 && !p.dividends) return bad("dividends is null but n_dividends > 0");
    for (uint64_t i = 0; i < p.n_dividends; ++i) {
        const MangoDividend& d = p.dividends[i];
        if (d.calendar_time < 0.0 || d.calendar_time > p.maturity) return bad("invalid discrete dividend");
        if (d.amount < 0.0) return bad("invalid discrete dividend");
        if (!std::isfinite(d.calendar_time) || !std::isfinite(d.amount)) return bad("invalid discrete dividend");
    }
    *why = nullptr;
    return MANGO_OK;
}

bool same_curve(const MangoPricingParams& a, const MangoPricingParams& b) {
    if (a.n_tenor_points != b.n_tenor_points) return false;
    const auto va = sorted_curve(a), vb = sorted_curve(b);
    for (size_t i = 0; i < va.size(); ++i)
        if (std::abs(va[i].tenor - vb[i].tenor) > 1e-12 || std::abs(va[i].log_discount - vb[i].log_discount) > 1e-12)
            return false;
    return true;
}

HostTimeDomain contract_time(const MangoPricingParams& p, int64_t n_time) {
    const std::vector<double> mand = mandatory_taus(p);
    if (mand.empty()) return time_uniform(0.0, p.maturity, n_time);
    return time_with_mandatory(0.0, p.maturity, p.maturity / static_cast<double>(n_time), mand);
}

int64_t append_event_record(std::vector<double>& pool, const MangoPricingParams& p, const HostTimeDomain& td) {
    const int64_t o = static_cast<int64_t>(pool.size());
    std::vector<MangoTenorPoint> curve;
    if (p.n_tenor_points > 0) curve = sorted_curve(p);
    // filter_and_merge_dividends, then the event list sorted by time-to-expiry
    std::vector<MangoDividend> f;
    for (uint64_t i = 0; i < p.n_dividends; ++i) {
        const MangoDividend& d = p.dividends[i];
        if (d.calendar_time > 0.0 && d.calendar_time < p.maturity && d.amount > 0.0) f.push_back(d);
    }
    std::sort(f.begin(), f.end(), [](const MangoDividend& a, const MangoDividend& b) { return a.calendar_time < b.calendar_time; });
    std::vector<MangoDividend> m;
    for (const auto& d : f) {
        if (!m.empty() && std::abs(m.back().calendar_time - d.calendar_time) < 1e-12) m.back().amount += d.amount;
        else m.push_back(d);
    }
    std::vector<std::pair<double, double>> ev;
    for (const auto& d : m) ev.push_back({p.maturity - d.calendar_time, d.amount / p.strike});
    std::sort(ev.begin(), ev.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    pool.push_back(static_cast<double>(curve.size()));
    pool.push_back(static_cast<double>(td.seg_start.size()));
    pool.push_back(static_cast<double>(ev.size()));
    pool.push_back(static_cast<double>(td.n_steps));
    pool.push_back(td.dt_avg);
    for (const auto& c : curve) pool.push_back(c.tenor);
    for (const auto& c : curve) pool.push_back(c.log_discount);
    for (double v : td.seg_start) pool.push_back(v);
    for (double v : td.sub_dt) pool.push_back(v);
    for (int64_t v : td.n_sub) pool.push_back(static_cast<double>(v));
    for (const auto& e : ev) pool.push_back(e.first);
    for (const auto& e : ev) pool.push_back(e.second);
    return o + 1;
}

}  // namespace mango_amd
