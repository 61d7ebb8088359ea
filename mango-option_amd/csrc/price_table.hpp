// SPDX-License-Identifier: MIT
//
// Device side of the price table: the 4-D EEP B-spline over (log-moneyness,
// maturity, volatility, rate) and the interpolated implied-volatility batch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "american_kernels.hpp"

namespace mango_amd {

constexpr int kTableDims = 4;

// A fitted table resident in device memory. All pointers are device
// pointers into one allocation owned by the host table object.
struct DeviceTable {
    const double* grid[kTableDims];   // axis points (ln m, tau, sigma, r)
    const double* knots[kTableDims];  // clamped cubic knots, n + 4 per axis
    int32_t n[kTableDims];
    const double* coeffs;             // [n0][n1][n2][n3] row-major
    double K_ref;
    double dividend_yield;
    int32_t option_type;              // 0 call, 1 put
    int32_t pad;
};

// InterpolatedIVSolverConfig (interpolated_iv_solver.hpp:44-57)
struct IvConfig {
    int64_t max_iter;
    double tolerance;
    double sigma_min;
    double sigma_max;
    double vega_threshold;
};

struct IvQueryRec {  // layout of MangoDeviceIvQuery
    double spot, strike, maturity, rate, dividend_yield, market_price;
    int32_t option_type;
    int32_t reserved;
};

struct IvResultRec {  // layout of MangoDeviceIvResult
    double implied_vol;
    double final_error;
    int32_t iterations;
    int32_t status;
};

// Snapshot solves of the table build: job j writes its grid to
// x[j * row] and the state at step steps[j * max_snap + k] (k < n_snap[j],
// ascending) to snaps[(j * max_snap + k) * row].
int launch_table_snapshots(const SolveVariant& v, const PdeJob* d_jobs, int n_jobs, const int32_t* d_steps,
                           const int32_t* d_n_snap, int max_snap, int row, double* d_x, double* d_snaps,
                           int32_t* d_status, hipStream_t stream);
bool table_variant_supported(const SolveVariant& v);

// Tensor extraction and EEP transform: one thread per (slice s = sigma-index
// * n_rate + rate-index, maturity column j). Slice s reads the solve
// src_job[s]; column j takes that solve's snapshot min(j, n_snap - 1).
// raw (nullable) receives K_ref-normalized American prices, eep the floored
// early-exercise premium, both [n_m][n_tau][n_sigma][n_rate].
struct ExtractArgs {
    const double* lm;
    const double* tau;
    const double* sigma;
    const double* rate;
    int32_t n_m, n_tau, n_sigma, n_rate;
    const int32_t* src_job;
    const int32_t* n_snap;
    const int64_t* n_space;
    int32_t max_snap;
    int32_t row;
    const double* x;
    const double* snaps;
    double* spline_scratch;  // 2 * row doubles per thread
    double K_ref;
    double dividend_yield;
    int32_t option_type;
    int32_t pad;
    double* raw;
    double* eep;
};
int launch_table_extract(const ExtractArgs& a, hipStream_t stream);

// One axis of the separable collocation fit: every 1-D slice along `axis`
// of src solves B c = f with the axis' LU factors (row pivots piv), writing
// c to dst; max |B c - f| over all slices goes to *d_max_res (as the bit
// pattern of a non-negative double, combined with atomicMax).
int launch_fit_axis(const double* d_src, double* d_dst, int64_t total, int na, int64_t stride,
                    const double* d_lu, const int32_t* d_piv, const double* d_b, unsigned long long* d_max_res,
                    hipStream_t stream);

// Table price and vega at points rows[i] = (S, K, tau, sigma, r).
int launch_table_price(const DeviceTable& t, const double* d_rows, int64_t n, double* d_price, double* d_vega,
                       hipStream_t stream);

// Table delta, gamma, theta, rho at points rows[i] = (S, K, tau, sigma, r):
// out[4 i + k].
int launch_table_greeks(const DeviceTable& t, const double* d_rows, int64_t n, double* d_out, hipStream_t stream);

// Interpolated IV for n queries.
int launch_interp_iv(const DeviceTable& t, const IvConfig& cfg, const IvQueryRec* d_q, int64_t n,
                     IvResultRec* d_out, hipStream_t stream);

}  // namespace mango_amd
