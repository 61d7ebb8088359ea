// SPDX-License-Identifier: MIT
//
// Clamped cubic B-spline basis shared by the host (collocation matrices of
// the fit) and the device (table evaluation): clamped_knots_cubic,
// find_span_cubic and cubic_basis_nonuniform of the reference
// (src/math/bspline/bspline_basis.hpp:47-257).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace mango_amd {

// Clamped knots for data sites x: four copies of each end, and n - 4 interior
// knots placed at equal fractions of the site index range, each interpolated
// between its neighbouring sites and kept strictly inside that interval.
inline std::vector<double> clamped_knots(const std::vector<double>& x) {
    const int n = static_cast<int>(x.size());
    std::vector<double> t(static_cast<size_t>(n + 4));
    for (int k = 0; k < 4; ++k) t[static_cast<size_t>(k)] = x.front();
    for (int k = 0; k < n - 4; ++k) {
        const double pos = static_cast<double>(k + 1) / static_cast<double>(n - 3) * static_cast<double>(n - 1);
        int lo = static_cast<int>(std::floor(pos));
        if (lo >= n - 1) lo = n - 2;
        const double frac = pos - static_cast<double>(lo);
        const double left = x[static_cast<size_t>(lo)];
        const double right = x[static_cast<size_t>(lo + 1)];
        const double eps = std::max(128.0 * std::numeric_limits<double>::epsilon() * (right - left),
                                    std::numeric_limits<double>::epsilon() * std::max(std::abs(right), 1.0));
        const double knot = (1.0 - frac) * left + frac * right;
        t[static_cast<size_t>(4 + k)] = std::clamp(knot, left + eps, right - eps);
    }
    for (int k = 0; k < 4; ++k) t[static_cast<size_t>(n + k)] = x.back();
    return t;
}

// Span of x among n control points: the last knot index i in [3, n - 1] with
// t[i] <= x (x at or past the ends clamps).
__host__ __device__ inline int find_span(const double* __restrict__ t, int n, double x) {
    const int lo_span = 3;
    const int hi_span = n - 1 > lo_span ? n - 1 : lo_span;
    if (x <= t[lo_span]) return lo_span;
    if (x >= t[n]) return hi_span;
    int lo = lo_span, hi = n + 1;  // first index in [3, n] with t[idx] > x
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (t[mid] <= x) lo = mid + 1; else hi = mid;
    }
    int i = lo - 1;
    i = i < lo_span ? lo_span : i;
    return i > hi_span ? hi_span : i;
}

// Cox-de Boor: N[k] is the weight of control point i - k; deriv gives the
// first derivative. nk = number of knots.
__host__ __device__ inline void cubic_basis(const double* __restrict__ t, int nk, int i, double x, double N[4],
                                            bool deriv) {
    if (!deriv && fabs(x - t[nk - 1]) < 64.0 * 2.220446049250313e-16 * fmax(fabs(t[nk - 1]), 1.0)) {
        N[0] = 1.0; N[1] = N[2] = N[3] = 0.0;
        return;
    }
    double N0[4], N1[4], N2[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        N0[k] = (idx >= 0 && idx + 1 < nk && t[idx] <= x && x < t[idx + 1]) ? 1.0 : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        N1[k] = 0.0;
        if (idx >= 0 && idx + 2 < nk) {
            const double ld = t[idx + 1] - t[idx], rd = t[idx + 2] - t[idx + 1];
            const double l = ld > 0.0 ? (x - t[idx]) / ld * N0[k] : 0.0;
            const double r = (rd > 0.0 && k > 0) ? (t[idx + 2] - x) / rd * N0[k > 0 ? k - 1 : 0] : 0.0;
            N1[k] = l + r;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        N2[k] = 0.0;
        if (idx >= 0 && idx + 3 < nk) {
            const double ld = t[idx + 2] - t[idx], rd = t[idx + 3] - t[idx + 1];
            const double l = ld > 0.0 ? (x - t[idx]) / ld * N1[k] : 0.0;
            const double r = (rd > 0.0 && k > 0) ? (t[idx + 3] - x) / rd * N1[k > 0 ? k - 1 : 0] : 0.0;
            N2[k] = l + r;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        N[k] = 0.0;
        if (idx >= 0 && idx + 4 < nk) {
            const double ld = t[idx + 3] - t[idx], rd = t[idx + 4] - t[idx + 1];
            if (deriv) {
                const double l = ld > 0.0 ? (3.0 / ld) * N2[k] : 0.0;
                const double r = (rd > 0.0 && k > 0) ? (3.0 / rd) * N2[k > 0 ? k - 1 : 0] : 0.0;
                N[k] = l - r;
            } else {
                const double l = ld > 0.0 ? (x - t[idx]) / ld * N2[k] : 0.0;
                const double r = (rd > 0.0 && k > 0) ? (t[idx + 4] - x) / rd * N2[k > 0 ? k - 1 : 0] : 0.0;
                N[k] = l + r;
            }
        }
    }
}

// Second derivative of the cubic basis: degree-0/1 basis, first derivative
// of the quadratic basis, then the cubic's second derivative.
__host__ __device__ inline void cubic_basis_d2(const double* __restrict__ t, int nk, int i, double x, double N[4]) {
    double N0[4], N1[4], dN2[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        N0[k] = (idx >= 0 && idx + 1 < nk && t[idx] <= x && x < t[idx + 1]) ? 1.0 : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        N1[k] = 0.0;
        if (idx >= 0 && idx + 2 < nk) {
            const double ld = t[idx + 1] - t[idx], rd = t[idx + 2] - t[idx + 1];
            const double l = ld > 0.0 ? (x - t[idx]) / ld * N0[k] : 0.0;
            const double r = (rd > 0.0 && k > 0) ? (t[idx + 2] - x) / rd * N0[k > 0 ? k - 1 : 0] : 0.0;
            N1[k] = l + r;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        dN2[k] = 0.0;
        if (idx >= 0 && idx + 3 < nk) {
            const double ld = t[idx + 2] - t[idx], rd = t[idx + 3] - t[idx + 1];
            const double l = ld > 0.0 ? (2.0 / ld) * N1[k] : 0.0;
            const double r = (rd > 0.0 && k > 0) ? (2.0 / rd) * N1[k > 0 ? k - 1 : 0] : 0.0;
            dN2[k] = l - r;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int idx = i - k;
        N[k] = 0.0;
        if (idx >= 0 && idx + 4 < nk) {
            const double ld = t[idx + 3] - t[idx], rd = t[idx + 4] - t[idx + 1];
            const double l = ld > 0.0 ? (3.0 / ld) * dN2[k] : 0.0;
            const double r = (rd > 0.0 && k > 0) ? (3.0 / rd) * dN2[k > 0 ? k - 1 : 0] : 0.0;
            N[k] = l - r;
        }
    }
}

}  // namespace mango_amd
