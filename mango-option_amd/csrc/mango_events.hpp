// SPDX-License-Identifier: MIT
//
// Yield curves, non-uniform time domains and discrete dividends: the host
// side (validation, time domains, the per-job record the event kernel reads)
// and the device kernel's launcher.
//
// Reference behaviour restated here:
//   YieldCurve::from_points validation        src/option/yield_curve.hpp:47-79
//   make_rate_fn / make_forward_discount_fn    src/option/option_spec.hpp:50-91
//   TimeDomain::with_mandatory_points          src/pde/core/time_domain.hpp:55-107
//   filter_and_merge_dividends                 src/option/dividend_utils.hpp:14-40
//   dividend widening + mandatory times        src/option/grid_spec_types.cpp:56-115, 157-175
//   per-contract mandatory times of a batch    src/option/american_option_batch.cpp:448-470
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/mango_amd.h"
#include "american_kernels.hpp"

namespace mango_amd {

// A time domain as the march walks it: uniform (n_steps steps of dt) or
// piecewise uniform, segment s starting at seg_start[s] with n_sub[s] steps
// of sub_dt[s]. dt_avg = (t1 - t0) / n_steps is TimeDomain::dt().
struct HostTimeDomain {
    int64_t n_steps = 0;
    double dt_avg = 0.0;
    std::vector<double> seg_start, sub_dt;
    std::vector<int64_t> n_sub;
};
HostTimeDomain time_uniform(double t0, double t1, int64_t n);
HostTimeDomain time_with_mandatory(double t0, double t1, double dt, std::vector<double> mandatory);

// Whether a solve needs the event kernel (a yield curve or dividends).
bool has_events(const MangoPricingParams& p);
// Dividend times as times-to-expiry strictly inside (0, T).
std::vector<double> mandatory_taus(const MangoPricingParams& p);
// Largest D/K over the dividends strictly inside (0, T) (0 if none).
double max_dividend_over_strike(const MangoPricingParams& p);
// The curve and dividend checks of mango_c_api.cpp:42-78 and
// validate_pricing_params (option_spec.cpp:131-152); *why is static text.
MangoStatus validate_events(const MangoPricingParams& p, const char** why);
// YieldCurve::from_points checks of a tenor list (false with *why set).
bool validate_curve(const MangoTenorPoint* pts, uint64_t n, const char** why);
// get_zero_rate (option_spec.hpp:99-112) of a valid curve: -ln D(T) / T, or
// the forward rate at 0 for T <= 0 (YieldCurve::zero_rate, yield_curve.hpp:126-130).
double curve_zero_rate(const MangoTenorPoint* pts, uint64_t n, double T);
// Two curves equal within 1e-12 per entry after sorting (YieldCurve::operator==).
bool same_curve(const MangoPricingParams& a, const MangoPricingParams& b);
// PDEGridConfig{spec, n_time, dividend dates} resolved per contract:
// uniform n_time steps, or with_mandatory_points(0, T, T / n_time, dates).
HostTimeDomain contract_time(const MangoPricingParams& p, int64_t n_time);

// Record of one event job in a double pool, at offset o:
//   [o]   n_curve   (0: constant rate)       [o+1] n_seg (0: uniform)
//   [o+2] n_events  [o+3] n_steps            [o+4] dt_avg
//   [o+5 ...] tenor[n_curve], log_discount[n_curve],
//             seg_start[n_seg], sub_dt[n_seg], n_sub[n_seg],
//             tau[n_events] ascending, d[n_events] (= amount / strike)
// Returns 1 + o (the PdeJob::ext value).
int64_t append_event_record(std::vector<double>& pool, const MangoPricingParams& p, const HostTimeDomain& td);

// Event kernel: one job per group of v.lanes lanes holding v.nodes nodes
// each; jobs must have ext != 0 and n_points <= lanes * nodes.
int launch_american_events(const SolveVariant& v, const PdeJob* d_jobs, int n_jobs, const double* d_pool,
                           Greeks* d_greeks, int32_t* d_status, const StateOut& state, hipStream_t stream);
// Variant for n points in a launch of n_jobs jobs ({0, 0} if n is too large).
SolveVariant events_variant_for(int64_t n_points, int64_t n_jobs);

}  // namespace mango_amd
