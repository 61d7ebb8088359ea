// SPDX-License-Identifier: MIT
//
// ORACLE — test infrastructure only (see mango_oracle.hpp).
//
// Price table and interpolated implied volatility, restated:
//   clamped cubic knots, span search, Cox-de Boor basis and derivative
//       (src/math/bspline/bspline_basis.hpp:47-257)
//   1-D collocation B c = f and its residual check
//       (src/math/bspline/bspline_collocation.hpp:326-456); the reference
//       factors the 4-band matrix with LAPACK dgbtrf (partial pivoting); this
//       restatement factors the same matrix densely with partial pivoting,
//       which gives the same solution up to rounding
//   separable N-D fit, axes N-1 .. 0 (bspline_nd_separable.hpp:170-384)
//   tensor-product evaluation and partials (bspline_nd.hpp:133-340)
//   PriceTableBuilderND::build for N = 4: make_batch, solve_batch with the
//       explicit-grid constraint check and auto fallback, extract_tensor,
//       EEP transform, fit (bspline_builder.cpp:26-600)
//   eep_floor (eep_decomposer.hpp:22-33), the EEP leaf price / vega
//       (transform_leaf.hpp, eep_layer.hpp)
//   InterpolatedIVSolver::solve (interpolated_iv_solver.hpp:431-556) with
//       Brent's method (root_finding.hpp:145-296)
//
// Pins: cubic-polynomial reproduction by clamped cubic interpolation (1-D and
// separable 4-D), interpolation at the data points, partition of unity, the
// derivative against finite differences, eep_floor closed values, Brent
// against known roots, and the table against direct PDE prices.
#pragma once

#include "mango_oracle.hpp"

#include <array>

namespace oracle {

// ---------------------------------------------------------------------------
// B-spline basis
// ---------------------------------------------------------------------------
inline std::vector<double> clamped_knots_cubic(const std::vector<double>& x) {
    const int n = static_cast<int>(x.size());
    std::vector<double> t(static_cast<size_t>(n + 4));
    std::fill_n(t.begin(), 4, x.front());
    if (n > 4) {
        const int interior = n - 4;
        const int intervals = n - 1;
        for (int idx = 0; idx < interior; ++idx) {
            const double ratio = static_cast<double>(idx + 1) / static_cast<double>(interior + 1);
            const double pos = ratio * static_cast<double>(intervals);
            int low = static_cast<int>(std::floor(pos));
            if (low >= intervals) low = intervals - 1;
            const double frac = pos - static_cast<double>(low);
            const double left = x[static_cast<size_t>(low)];
            const double right = x[static_cast<size_t>(low + 1)];
            double knot = (1.0 - frac) * left + frac * right;
            const double spacing = right - left;
            const double eps = std::max(128.0 * std::numeric_limits<double>::epsilon() * spacing,
                                        std::numeric_limits<double>::epsilon() * std::max(std::abs(right), 1.0));
            knot = std::clamp(knot, left + eps, right - eps);
            t[static_cast<size_t>(4 + idx)] = knot;
        }
    }
    std::fill_n(t.end() - 4, 4, x.back());
    return t;
}

inline int find_span_cubic(const std::vector<double>& t, double x) {
    const int n_ctrl = static_cast<int>(t.size()) - 4;
    const int min_span = 3;
    const int max_span = std::max(min_span, n_ctrl - 1);
    if (x <= t[static_cast<size_t>(min_span)]) return min_span;
    if (x >= t[static_cast<size_t>(n_ctrl)]) return max_span;
    auto it = std::upper_bound(t.begin() + min_span, t.begin() + n_ctrl + 1, x);
    int i = static_cast<int>(std::distance(t.begin(), it)) - 1;
    return std::clamp(i, min_span, max_span);
}

// N[k] is the weight of control point i - k. deriv: first derivative.
inline void cubic_basis(const std::vector<double>& t, int i, double x, double N[4], bool deriv = false) {
    const int n = static_cast<int>(t.size());
    if (!deriv && std::abs(x - t.back()) < 64.0 * std::numeric_limits<double>::epsilon() * std::max(std::abs(t.back()), 1.0)) {
        N[0] = 1.0; N[1] = N[2] = N[3] = 0.0;
        return;
    }
    double N0[4] = {0, 0, 0, 0}, N1[4] = {0, 0, 0, 0}, N2[4] = {0, 0, 0, 0};
    auto T = [&](int j) { return t[static_cast<size_t>(j)]; };
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        if (idx >= 0 && idx + 1 < n) N0[k] = (T(idx) <= x && x < T(idx + 1)) ? 1.0 : 0.0;
    }
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        if (idx >= 0 && idx + 2 < n) {
            const double ld = T(idx + 1) - T(idx), rd = T(idx + 2) - T(idx + 1);
            const double l = ld > 0.0 ? (x - T(idx)) / ld * N0[k] : 0.0;
            const double r = (rd > 0.0 && k > 0) ? (T(idx + 2) - x) / rd * N0[k - 1] : 0.0;
            N1[k] = l + r;
        }
    }
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        if (idx >= 0 && idx + 3 < n) {
            const double ld = T(idx + 2) - T(idx), rd = T(idx + 3) - T(idx + 1);
            const double l = ld > 0.0 ? (x - T(idx)) / ld * N1[k] : 0.0;
            const double r = (rd > 0.0 && k > 0) ? (T(idx + 3) - x) / rd * N1[k - 1] : 0.0;
            N2[k] = l + r;
        }
    }
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        if (idx >= 0 && idx + 4 < n) {
            const double ld = T(idx + 3) - T(idx), rd = T(idx + 4) - T(idx + 1);
            if (deriv) {
                const double l = ld > 0.0 ? (3.0 / ld) * N2[k] : 0.0;
                const double r = (rd > 0.0 && k > 0) ? (3.0 / rd) * N2[k - 1] : 0.0;
                N[k] = l - r;
            } else {
                const double l = ld > 0.0 ? (x - T(idx)) / ld * N2[k] : 0.0;
                const double r = (rd > 0.0 && k > 0) ? (T(idx + 4) - x) / rd * N2[k - 1] : 0.0;
                N[k] = l + r;
            }
        } else {
            N[k] = 0.0;
        }
    }
}

// Second derivative of the cubic basis (cubic_basis_second_derivative_nonuniform,
// bspline_basis.hpp:343-409): degree-0/1 basis, first derivative of the
// quadratic basis, then the cubic's second derivative.
inline void cubic_basis_d2(const std::vector<double>& t, int i, double x, double N[4]) {
    const int n = static_cast<int>(t.size());
    double N0[4] = {0, 0, 0, 0}, N1[4] = {0, 0, 0, 0}, dN2[4] = {0, 0, 0, 0};
    auto T = [&](int j) { return t[static_cast<size_t>(j)]; };
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        if (idx >= 0 && idx + 1 < n) N0[k] = (T(idx) <= x && x < T(idx + 1)) ? 1.0 : 0.0;
    }
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        if (idx >= 0 && idx + 2 < n) {
            const double ld = T(idx + 1) - T(idx), rd = T(idx + 2) - T(idx + 1);
            const double l = ld > 0.0 ? (x - T(idx)) / ld * N0[k] : 0.0;
            const double r = (rd > 0.0 && k > 0) ? (T(idx + 2) - x) / rd * N0[k - 1] : 0.0;
            N1[k] = l + r;
        }
    }
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        if (idx >= 0 && idx + 3 < n) {
            const double ld = T(idx + 2) - T(idx), rd = T(idx + 3) - T(idx + 1);
            const double l = ld > 0.0 ? (2.0 / ld) * N1[k] : 0.0;
            const double r = (rd > 0.0 && k > 0) ? (2.0 / rd) * N1[k - 1] : 0.0;
            dN2[k] = l - r;
        }
    }
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        if (idx >= 0 && idx + 4 < n) {
            const double ld = T(idx + 3) - T(idx), rd = T(idx + 4) - T(idx + 1);
            const double l = ld > 0.0 ? (3.0 / ld) * dN2[k] : 0.0;
            const double r = (rd > 0.0 && k > 0) ? (3.0 / rd) * dN2[k - 1] : 0.0;
            N[k] = l - r;
        } else {
            N[k] = 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// 1-D collocation: B c = f, B[i][j] = N_j(x_i)
// ---------------------------------------------------------------------------
struct Collocation1D {
    std::vector<double> grid, knots, B, LU;
    std::vector<int> piv;
    int n = 0;
    bool build(const std::vector<double>& x) {
        grid = x;
        n = static_cast<int>(x.size());
        if (n < 4) return false;
        for (int i = 1; i < n; ++i) if (x[static_cast<size_t>(i)] - x[static_cast<size_t>(i - 1)] < 1e-14) return false;
        knots = clamped_knots_cubic(x);
        B.assign(static_cast<size_t>(n * n), 0.0);
        for (int i = 0; i < n; ++i) {
            const int span = find_span_cubic(knots, x[static_cast<size_t>(i)]);
            double basis[4];
            cubic_basis(knots, span, x[static_cast<size_t>(i)], basis);
            for (int k = 0; k < 4; ++k) {
                const int col = span - k;
                if (col >= 0 && col < n) B[static_cast<size_t>(i * n + col)] = basis[k];
            }
        }
        // dense LU with partial pivoting
        LU = B;
        piv.assign(static_cast<size_t>(n), 0);
        for (int c = 0; c < n; ++c) {
            int p = c;
            for (int r = c + 1; r < n; ++r)
                if (std::abs(LU[static_cast<size_t>(r * n + c)]) > std::abs(LU[static_cast<size_t>(p * n + c)])) p = r;
            piv[static_cast<size_t>(c)] = p;
            if (LU[static_cast<size_t>(p * n + c)] == 0.0) return false;
            if (p != c) for (int j = 0; j < n; ++j) std::swap(LU[static_cast<size_t>(c * n + j)], LU[static_cast<size_t>(p * n + j)]);
            for (int r = c + 1; r < n; ++r) {
                const double f = LU[static_cast<size_t>(r * n + c)] / LU[static_cast<size_t>(c * n + c)];
                LU[static_cast<size_t>(r * n + c)] = f;
                for (int j = c + 1; j < n; ++j) LU[static_cast<size_t>(r * n + j)] -= f * LU[static_cast<size_t>(c * n + j)];
            }
        }
        return true;
    }
    // c = B^-1 f; returns the max residual |B c - f|
    double solve(const double* f, double* c) const {
        std::vector<double> y(f, f + n);
        for (int k = 0; k < n; ++k) std::swap(y[static_cast<size_t>(k)], y[static_cast<size_t>(piv[static_cast<size_t>(k)])]);
        for (int r = 0; r < n; ++r)
            for (int j = 0; j < r; ++j) y[static_cast<size_t>(r)] -= LU[static_cast<size_t>(r * n + j)] * y[static_cast<size_t>(j)];
        for (int r = n - 1; r >= 0; --r) {
            for (int j = r + 1; j < n; ++j) y[static_cast<size_t>(r)] -= LU[static_cast<size_t>(r * n + j)] * y[static_cast<size_t>(j)];
            y[static_cast<size_t>(r)] /= LU[static_cast<size_t>(r * n + r)];
        }
        double res = 0.0;
        for (int i = 0; i < n; ++i) {
            double bc = 0.0;
            for (int j = 0; j < n; ++j) bc = std::fma(B[static_cast<size_t>(i * n + j)], y[static_cast<size_t>(j)], bc);
            res = std::max(res, std::abs(bc - f[i]));
        }
        std::copy(y.begin(), y.end(), c);
        return res;
    }
};

// Separable fit of values[i0][i1][i2][i3] (row-major), axes 3, 2, 1, 0.
inline bool fit_separable4(const std::array<std::vector<double>, 4>& grids, std::vector<double>& v,
                           double tol = 1e-6, double* max_residual = nullptr) {
    std::array<size_t, 4> dims{}, strides{};
    for (int a = 0; a < 4; ++a) dims[static_cast<size_t>(a)] = grids[static_cast<size_t>(a)].size();
    strides[3] = 1;
    for (int a = 2; a >= 0; --a) strides[static_cast<size_t>(a)] = strides[static_cast<size_t>(a + 1)] * dims[static_cast<size_t>(a + 1)];
    double worst = 0.0;
    for (int axis = 3; axis >= 0; --axis) {
        Collocation1D col;
        if (!col.build(grids[static_cast<size_t>(axis)])) return false;
        const size_t na = dims[static_cast<size_t>(axis)], st = strides[static_cast<size_t>(axis)];
        const size_t total = v.size();
        std::vector<double> f(na), c(na);
        for (size_t base = 0; base < total; ++base) {
            // base offsets whose index along `axis` is 0
            if ((base / st) % na != 0) continue;
            for (size_t i = 0; i < na; ++i) f[i] = v[base + i * st];
            const double r = col.solve(f.data(), c.data());
            if (!(r <= tol)) return false;
            worst = std::max(worst, r);
            for (size_t i = 0; i < na; ++i) v[base + i * st] = c[i];
        }
    }
    if (max_residual) *max_residual = worst;
    return true;
}

struct BSpline4 {
    std::array<std::vector<double>, 4> grids, knots;
    std::vector<double> coeffs;
    void set(const std::array<std::vector<double>, 4>& g, std::vector<double> c) {
        grids = g;
        for (int a = 0; a < 4; ++a) knots[static_cast<size_t>(a)] = clamped_knots_cubic(g[static_cast<size_t>(a)]);
        coeffs = std::move(c);
    }
    // clamp_bspline_query + tensor-product sum (bspline_nd.hpp:40-62, 133-160, 287-314)
    // order 1: first partial along partial_axis, 2: second partial
    double eval(const double q[4], int partial_axis = -1, int order = 1) const {
        int span[4];
        double w[4][4];
        for (int a = 0; a < 4; ++a) {
            const auto& g = grids[static_cast<size_t>(a)];
            double x = q[a];
            if (x <= g.front()) x = g.front();
            else if (x >= g.back()) x = std::nextafter(g.back(), -std::numeric_limits<double>::infinity());
            span[a] = find_span_cubic(knots[static_cast<size_t>(a)], x);
            if (a == partial_axis && order == 2) cubic_basis_d2(knots[static_cast<size_t>(a)], span[a], x, w[a]);
            else cubic_basis(knots[static_cast<size_t>(a)], span[a], x, w[a], a == partial_axis);
        }
        const size_t d1 = grids[1].size(), d2 = grids[2].size(), d3 = grids[3].size();
        const int dims[4] = {static_cast<int>(grids[0].size()), static_cast<int>(d1), static_cast<int>(d2), static_cast<int>(d3)};
        double s0 = 0.0;
        for (int o0 = 0; o0 < 4; ++o0) {
            const int i0 = span[0] - o0;
            if (static_cast<unsigned>(i0) >= static_cast<unsigned>(dims[0])) continue;
            double s1 = 0.0;
            for (int o1 = 0; o1 < 4; ++o1) {
                const int i1 = span[1] - o1;
                if (static_cast<unsigned>(i1) >= static_cast<unsigned>(dims[1])) continue;
                double s2 = 0.0;
                for (int o2 = 0; o2 < 4; ++o2) {
                    const int i2 = span[2] - o2;
                    if (static_cast<unsigned>(i2) >= static_cast<unsigned>(dims[2])) continue;
                    double s3 = 0.0;
                    for (int o3 = 0; o3 < 4; ++o3) {
                        const int i3 = span[3] - o3;
                        if (static_cast<unsigned>(i3) >= static_cast<unsigned>(dims[3])) continue;
                        const size_t idx = ((static_cast<size_t>(i0) * d1 + static_cast<size_t>(i1)) * d2 + static_cast<size_t>(i2)) * d3 + static_cast<size_t>(i3);
                        s3 = std::fma(coeffs[idx], w[3][o3], s3);
                    }
                    s2 = std::fma(s3, w[2][o2], s2);
                }
                s1 = std::fma(s2, w[1][o1], s1);
            }
            s0 = std::fma(s1, w[0][o0], s0);
        }
        return s0;
    }
};

inline double eep_floor(double eep_raw) {
    constexpr double kSharpness = 100.0;
    if (kSharpness * eep_raw > 500.0) return eep_raw;
    const double softplus = std::log1p(std::exp(kSharpness * eep_raw)) / kSharpness;
    const double bias = std::log(2.0) / kSharpness;
    return std::max(0.0, softplus - bias);
}

// European closed form with the reference's edge cases (european_option.hpp:24-66)
inline double intrinsic(double S, double K, OptionType type) {
    return type == OptionType::CALL ? std::max(S - K, 0.0) : std::max(K - S, 0.0);
}
inline double european_price(double S, double K, double tau, double sigma, double r, double q, OptionType type) {
    if (tau <= 0.0 || sigma <= 0.0) {
        if (tau <= 0.0) return intrinsic(S, K, type);
        return intrinsic(S * std::exp(-q * tau), K * std::exp(-r * tau), type);
    }
    return bs_price(S, K, tau, sigma, r, q, type);
}
inline double european_vega(double S, double K, double tau, double sigma, double r, double q) {
    if (tau <= 0.0 || sigma <= 0.0) return 0.0;
    return bs_vega(S, K, tau, sigma, r, q);
}

// EuropeanOptionResult delta, gamma, theta, rho (european_option.cpp:60-160)
inline double bs_d1(double S, double K, double tau, double sigma, double r, double q) {
    const double sst = sigma * std::sqrt(tau);
    return (std::log(S / K) + (r - q + 0.5 * sigma * sigma) * tau) / sst;
}
inline double european_delta(double S, double K, double tau, double sigma, double r, double q, OptionType type) {
    const double eq = std::exp(-q * tau);
    if (tau <= 0.0 || sigma <= 0.0) {
        if (type == OptionType::PUT) return (S < K) ? -eq : 0.0;
        return (S > K) ? eq : 0.0;
    }
    const double d1 = bs_d1(S, K, tau, sigma, r, q);
    return type == OptionType::PUT ? -eq * norm_cdf(-d1) : eq * norm_cdf(d1);
}
inline double european_gamma(double S, double K, double tau, double sigma, double r, double q) {
    if (tau <= 0.0 || sigma <= 0.0) return 0.0;
    const double d1 = bs_d1(S, K, tau, sigma, r, q);
    return std::exp(-q * tau) * norm_pdf(d1) / (S * sigma * std::sqrt(tau));
}
inline double european_theta(double S, double K, double tau, double sigma, double r, double q, OptionType type) {
    if (tau <= 0.0 || sigma <= 0.0) return 0.0;
    const double d1 = bs_d1(S, K, tau, sigma, r, q);
    const double d2 = d1 - sigma * std::sqrt(tau);
    const double sqrt_tau = std::sqrt(tau);
    const double eq = std::exp(-q * tau), er = std::exp(-r * tau);
    const double common = -S * eq * norm_pdf(d1) * sigma / (2.0 * sqrt_tau);
    if (type == OptionType::PUT) return common + r * K * er * norm_cdf(-d2) - q * S * eq * norm_cdf(-d1);
    return common - r * K * er * norm_cdf(d2) + q * S * eq * norm_cdf(d1);
}
inline double european_rho(double S, double K, double tau, double sigma, double r, double q, OptionType type) {
    if (tau <= 0.0 || sigma <= 0.0) return 0.0;
    const double d1 = bs_d1(S, K, tau, sigma, r, q);
    const double d2 = d1 - sigma * std::sqrt(tau);
    const double er = std::exp(-r * tau);
    return type == OptionType::PUT ? -K * tau * er * norm_cdf(-d2) : K * tau * er * norm_cdf(d2);
}

// ---------------------------------------------------------------------------
// Price table: EEP B-spline over (ln(S/K), tau, sigma, r) at K_ref
// ---------------------------------------------------------------------------
struct PriceTable {
    BSpline4 spline;
    double K_ref = 100.0, q = 0.0;
    OptionType type = OptionType::PUT;
    // EEPLayer<TransformLeaf> price and vega
    double price(double S, double K, double tau, double sigma, double r) const {
        const double c[4] = {std::log(S / K), tau, sigma, r};
        const double raw = spline.eval(c);
        return std::max(0.0, raw) * K / K_ref + european_price(S, K, tau, sigma, r, q, type);
    }
    double vega(double S, double K, double tau, double sigma, double r) const {
        const double c[4] = {std::log(S / K), tau, sigma, r};
        const double raw = spline.eval(c);
        const double leaf = raw <= 0.0 ? 0.0 : spline.eval(c, 2) * K / K_ref;
        return leaf + european_vega(S, K, tau, sigma, r, q);
    }
    // EEPLayer::greek / gamma over TransformLeaf with StandardTransform4D
    // weights (eep_layer.hpp:33-77, transform_leaf.hpp:48-90,
    // transforms/standard_4d.hpp:20-31): delta = f_x / S, theta = -f_tau,
    // rho = f_r, gamma = (f_xx - f_x) / S^2, each times K / K_ref, plus the
    // European Greek; the European Greek alone where the EEP reads <= 0.
    void greeks(double S, double K, double tau, double sigma, double r, double out[4]) const {
        const double c[4] = {std::log(S / K), tau, sigma, r};
        const double raw = spline.eval(c);
        out[0] = european_delta(S, K, tau, sigma, r, q, type);
        out[1] = european_gamma(S, K, tau, sigma, r, q);
        out[2] = european_theta(S, K, tau, sigma, r, q, type);
        out[3] = european_rho(S, K, tau, sigma, r, q, type);
        if (raw <= 0.0) return;
        const double fx = spline.eval(c, 0);
        const double fxx = spline.eval(c, 0, 2);
        double d = 0.0;
        d += (1.0 / S) * fx;
        out[0] = d * K / K_ref + out[0];
        out[1] = (fxx - fx) / (S * S) * K / K_ref + out[1];
        double th = 0.0;
        th += -1.0 * spline.eval(c, 1);
        out[2] = th * K / K_ref + out[2];
        double rh = 0.0;
        rh += 1.0 * spline.eval(c, 3);
        out[3] = rh * K / K_ref + out[3];
    }
};

struct TableBuildInfo {
    size_t pde_solves = 0, failed_solves = 0;
    size_t repaired_points = 0, repaired_slices = 0;  // repair_failed_slices diagnostics
    double max_residual = 0.0;
    bool fallback = false;
    std::vector<double> tensor;  // normalized American prices before the EEP transform
};

// PriceTableBuilderND<4>::build with an explicit PDEGridConfig (the
// constraint check and auto fallback of solve_batch) or GridAccuracyParams.
inline bool build_price_table(const std::array<std::vector<double>, 4>& axes, double K_ref, double q, OptionType type,
                              const std::optional<PDEGridConfig>& explicit_grid, const GridAccuracyParams& acc_in,
                              PriceTable& out, TableBuildInfo& info, int n_threads = 1) {
    const auto& lm = axes[0];
    const auto& tau = axes[1];
    const auto& sig = axes[2];
    const auto& rate = axes[3];
    for (const auto& g : axes) if (g.size() < 4) return false;
    // build() checks (bspline_builder.cpp:46-87): positive tau and sigma,
    // K_ref > 0, and an explicit grid covering the log-moneyness axis
    if (!(tau.front() > 0.0) || !(sig.front() > 0.0) || !(K_ref > 0.0)) return false;
    if (explicit_grid && (lm.front() < explicit_grid->spec.x_min || lm.back() > explicit_grid->spec.x_max)) return false;
    // make_batch (:200-230): sigma outer, rate inner, S = K = K_ref, T = tau_max
    std::vector<PricingParams> batch;
    for (double s : sig)
        for (double r : rate) batch.push_back(PricingParams{K_ref, K_ref, tau.back(), r, q, s, type, std::nullopt, {}});
    BatchSolver bs;
    bs.n_threads = n_threads;
    bs.opts.snapshot_times = tau;
    double max_sst = 0.0;
    for (const auto& p : batch) max_sst = std::max(max_sst, p.volatility * std::sqrt(p.maturity));
    auto ensure_coverage = [&](GridAccuracyParams& a) {  // ensure_moneyness_coverage (:233-255)
        const double half = std::max(std::abs(lm.front()), std::abs(lm.back()));
        const double sst = std::max(max_sst, 1e-10);
        a.n_sigma = std::max(a.n_sigma, (half / sst) * 1.1);
    };
    std::vector<BatchItem> res;
    if (explicit_grid) {
        const GridSpec& g = explicit_grid->spec;
        const std::vector<double> xg = g.generate();
        double max_dx = 0.0;
        for (size_t i = 1; i < xg.size(); ++i) max_dx = std::max(max_dx, xg[i] - xg[i - 1]);
        const double width = g.x_max - g.x_min;
        const bool ok = width <= 5.8 && max_dx <= 0.05 && width >= 6.0 * max_sst;
        if (ok) {
            res = bs.solve_batch(batch, true, PDEGridConfig{g, explicit_grid->n_time});
        } else {
            info.fallback = true;
            GridAccuracyParams a;
            const size_t clamped = std::clamp<size_t>(g.n_points, 100, 1200);
            a.min_spatial_points = clamped;
            a.max_spatial_points = clamped;
            a.max_time_steps = explicit_grid->n_time;
            if (g.type == GridSpec::Type::SinhSpaced) a.alpha = g.concentration;
            const double max_abs_x = std::max(std::abs(g.x_min), std::abs(g.x_max));
            a.n_sigma = std::max(5.0, (max_abs_x / std::max(max_sst, 1e-10)) * 1.1);
            ensure_coverage(a);
            bs.accuracy = a;
            res = bs.solve_batch(batch, true);
        }
    } else {
        GridAccuracyParams a = acc_in;
        ensure_coverage(a);
        auto est = estimate_batch_pde_grid(batch, a);
        res = bs.solve_batch(batch, true, PDEGridConfig{est.first, est.second.n_steps});
    }
    info.pde_solves = bs.pde_solves;
    // extract_tensor (:395-480): tensor[m][tau][sigma][r] = natural spline of
    // snapshot j at ln m. Snapshot j is the j-th entry of the solve's sorted,
    // deduplicated snapshot list (at_time(j)), so when two maturities snap to
    // one step the later columns take the next recorded step, and a column
    // with no snapshot left is a spline failure.
    const size_t Nm = lm.size(), Nt = tau.size(), Ns = sig.size(), Nr = rate.size();
    const double nan = std::numeric_limits<double>::quiet_NaN();
    std::vector<double> t(Nm * Nt * Ns * Nr, nan);
    auto at = [&](size_t i, size_t j, size_t si, size_t ri) -> double& { return t[((i * Nt + j) * Ns + si) * Nr + ri]; };
    std::vector<char> pde_failed(Ns * Nr, 0);
    std::vector<std::vector<size_t>> spline_failed(Ns * Nr);
    for (size_t si = 0; si < Ns; ++si)
        for (size_t ri = 0; ri < Nr; ++ri) {
            const BatchItem& b = res[si * Nr + ri];
            if (b.status != kOk) { pde_failed[si * Nr + ri] = 1; ++info.failed_solves; continue; }
            const size_t n = b.sol.x.size();
            const size_t n_snap = b.sol.snapshot_indices.size();
            for (size_t j = 0; j < Nt; ++j) {
                CubicSpline sp;
                if (j >= n_snap || !sp.build(b.sol.x.data(), b.sol.snapshots.data() + j * n, n)) {
                    spline_failed[si * Nr + ri].push_back(j);
                    continue;
                }
                for (size_t i = 0; i < Nm; ++i) at(i, j, si, ri) = sp.eval(lm[i]);
            }
        }
    // max_failure_rate = 0 (bspline_builder.hpp:101; checked at bspline_builder.cpp:94-101)
    if (info.failed_solves) return false;
    // repair_failed_slices (:921-960): partial spline failures interpolate
    // along tau between the nearest valid columns (:802-861); slices with
    // every column failed copy the nearest valid (sigma, r) slice (:757-800, :863-905)
    std::vector<size_t> full;
    std::vector<char> valid(Ns * Nr, 1);
    for (size_t k = 0; k < Ns * Nr; ++k)
        if (pde_failed[k] || spline_failed[k].size() == Nt) { full.push_back(k); valid[k] = 0; }
    for (size_t k = 0; k < Ns * Nr; ++k) {
        const auto& fails = spline_failed[k];
        if (fails.empty() || fails.size() == Nt) continue;
        const size_t si = k / Nr, ri = k % Nr;
        info.repaired_points += fails.size();
        for (size_t j : fails) {
            long before = -1, after = -1;
            for (size_t jj = j; jj-- > 0;) if (!std::isnan(at(0, jj, si, ri))) { before = static_cast<long>(jj); break; }
            for (size_t jj = j + 1; jj < Nt; ++jj) if (!std::isnan(at(0, jj, si, ri))) { after = static_cast<long>(jj); break; }
            for (size_t i = 0; i < Nm; ++i) {
                if (before >= 0 && after >= 0) {
                    const double w = static_cast<double>(static_cast<long>(j) - before) / static_cast<double>(after - before);
                    at(i, j, si, ri) = (1.0 - w) * at(i, static_cast<size_t>(before), si, ri) + w * at(i, static_cast<size_t>(after), si, ri);
                } else {
                    at(i, j, si, ri) = at(i, static_cast<size_t>(before >= 0 ? before : after), si, ri);
                }
            }
        }
    }
    std::sort(full.begin(), full.end());
    for (size_t k : full) {
        const long s0 = static_cast<long>(k / Nr), r0 = static_cast<long>(k % Nr);
        long ns = -1, nr = -1;
        const long max_dist = static_cast<long>(Ns - 1 + Nr - 1);
        for (long d = 1; d <= max_dist && ns < 0; ++d)
            for (long ds = -d; ds <= d && ns < 0; ++ds) {
                const long dr = d - std::abs(ds);
                for (long sign : {-1L, 1L}) {
                    const long a = s0 + ds, b = r0 + sign * dr;
                    if (a >= 0 && a < static_cast<long>(Ns) && b >= 0 && b < static_cast<long>(Nr) && valid[a * Nr + b]) {
                        ns = a; nr = b; break;
                    }
                }
            }
        if (ns < 0) return false;
        for (size_t i = 0; i < Nm; ++i)
            for (size_t j = 0; j < Nt; ++j)
                at(i, j, static_cast<size_t>(s0), static_cast<size_t>(r0)) = at(i, j, static_cast<size_t>(ns), static_cast<size_t>(nr));
        valid[k] = 1;
        ++info.repaired_slices;
    }
    info.tensor = t;
    // EEP transform (BSplineTensorAccessor + eep_decompose)
    for (size_t i = 0; i < Nm; ++i)
        for (size_t j = 0; j < Nt; ++j)
            for (size_t si = 0; si < Ns; ++si)
                for (size_t ri = 0; ri < Nr; ++ri) {
                    double& v = t[((i * Nt + j) * Ns + si) * Nr + ri];
                    const double am = K_ref * v;
                    const double eu = european_price(std::exp(lm[i]) * K_ref, K_ref, tau[j], sig[si], rate[ri], q, type);
                    v = eep_floor(am - eu);
                }
    if (!fit_separable4(axes, t, 1e-6, &info.max_residual)) return false;
    out.spline.set(axes, std::move(t));
    out.K_ref = K_ref;
    out.q = q;
    out.type = type;
    return true;
}

// ---------------------------------------------------------------------------
// Interpolated implied volatility
// ---------------------------------------------------------------------------
struct IVConfig {
    size_t max_iter = 50;
    double tolerance = 1e-6;
    double sigma_min = 0.01, sigma_max = 3.0;
    double vega_threshold = 1e-4;
};
struct IVOut {
    int32_t status = 0;  // MangoStatus
    double iv = 0.0, final_error = 0.0;
    size_t iterations = 0;
};

// brent_find_root (root_finding.hpp:145-296); status 0 ok, 3 max iterations,
// 4 bracketing, 5 numerical.
template <typename F>
inline int32_t brent(F&& f, double a, double b, double tol, size_t max_iter, double& root, size_t& iters, double& err) {
    double fa = f(a), fb = f(b);
    iters = 0;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    if (!std::isfinite(fa) || !std::isfinite(fb)) { err = nan; return 5; }
    if (fa * fb > 0.0) { err = std::min(std::abs(fa), std::abs(fb)); return 4; }
    if (std::abs(fa) < tol) { root = a; err = std::abs(fa); return 0; }
    if (std::abs(fb) < tol) { root = b; err = std::abs(fb); return 0; }
    if (std::abs(fa) < std::abs(fb)) { std::swap(a, b); std::swap(fa, fb); }
    double c = a, fc = fa, d = 0.0;
    bool mflag = true;
    for (size_t it = 0; it < max_iter; ++it) {
        if (std::abs(fb) < tol || std::abs(b - a) < tol) { root = b; iters = it + 1; err = std::abs(fb); return 0; }
        double s;
        if (fa != fc && fb != fc)
            s = a * fb * fc / ((fa - fb) * (fa - fc)) + b * fa * fc / ((fb - fa) * (fb - fc)) + c * fa * fb / ((fc - fa) * (fc - fb));
        else
            s = b - fb * (b - a) / (fb - fa);
        const double bisect = (3.0 * a + b) / 4.0;
        const bool c1 = !((s > bisect && s < b) || (s < bisect && s > b));
        const bool c2 = mflag && std::abs(s - b) >= std::abs(b - c) / 2.0;
        const bool c3 = !mflag && std::abs(s - b) >= std::abs(c - d) / 2.0;
        const bool c4 = mflag && std::abs(b - c) < tol;
        const bool c5 = !mflag && std::abs(c - d) < tol;
        if (c1 || c2 || c3 || c4 || c5) { s = (a + b) / 2.0; mflag = true; }
        else mflag = false;
        const double fs = f(s);
        if (!std::isfinite(fs)) { iters = it + 1; err = nan; root = s; return 5; }
        d = c; c = b; fc = fb;
        if (fa * fs < 0.0) { b = s; fb = fs; } else { a = s; fa = fs; }
        if (std::abs(fa) < std::abs(fb)) { std::swap(a, b); std::swap(fa, fb); }
    }
    iters = max_iter;
    err = std::abs(fb);
    root = b;
    return 3;
}

// InterpolatedIVSolver::solve for a constant-rate query
inline IVOut solve_iv(const PriceTable& t, double S, double K, double tau, double r, double q, OptionType type,
                      double market, const IVConfig& cfg = {}) {
    IVOut o;
    auto fail = [&](int32_t st) { o.status = st; return o; };
    // validate_query (:382-397) then validate_iv_query (option_spec.cpp:52-88),
    // statuses as map_iv_error (mango_c_api.cpp:124-141) after
    // validation_error_to_iv_error (iv_result.hpp:55-100)
    if (type != t.type || std::abs(q - t.q) > 1e-10) return fail(1);
    if (!(S > 0.0) || !std::isfinite(S) || !(K > 0.0) || !std::isfinite(K) || !(tau > 0.0) || !std::isfinite(tau))
        return fail(1);
    if (!std::isfinite(r) || q < 0.0 || !std::isfinite(q)) return fail(2);
    if (!std::isfinite(market) || market <= 0.0) return fail(1);
    const double intr = intrinsic(S, K, type);
    const double upper = type == OptionType::CALL ? S : K;
    if (market < intr || market > upper) return fail(2);
    // adaptive_bounds (:399-423)
    const double tv = market - intr;
    const double tvp = tv / market;
    const double su = tvp > 0.5 ? 3.0 : (tvp > 0.2 ? 2.0 : 1.5);
    const auto& g = t.spline.grids;
    double smin = std::max(cfg.sigma_min, g[2].front());
    double smax = std::min({su, cfg.sigma_max, g[2].back()});
    if (smin >= smax) { smin = g[2].front(); smax = g[2].back(); }
    // is_in_bounds at both bracket ends
    const double x = std::log(S / K);
    auto inb = [&](double v) {
        return x >= g[0].front() && x <= g[0].back() && tau >= g[1].front() && tau <= g[1].back() &&
               v >= g[2].front() && v <= g[2].back() && r >= g[3].front() && r <= g[3].back();
    };
    if (!inb(smin) || !inb(smax)) return fail(1);
    if (cfg.vega_threshold > 0.0) {
        double mv = 0.0;
        for (double sv : {0.10, 0.25, 0.50}) mv = std::max(mv, std::abs(t.vega(S, K, tau, sv, r)));
        if (mv < cfg.vega_threshold) { o.final_error = mv; return fail(5); }
    }
    const double moneyness = S / K;
    auto f = [&](double s) { return t.price(moneyness * K, K, tau, s, r) - market; };
    double root = 0.0, err = 0.0;
    size_t it = 0;
    const int32_t st = brent(f, smin, smax, cfg.tolerance, cfg.max_iter, root, it, err);
    o.iterations = it;
    o.final_error = err;
    if (st != 0) return fail(st);  // 3 no convergence, 4 bracketing, 5 numerical
    o.iv = root;
    return o;
}

}  // namespace oracle
