# SPDX-License-Identifier: MIT
"""mango_option for MI355X: the reference's Python surface for the American
PDE path (src/python/mango_bindings.cpp:293-757), backed by gfx950 kernels
through the C ABI of include/mango_amd.h.

Names, argument meaning and error behaviour follow the pybind11 module:
`american_option_price(params, accuracy=None)` raises ValueError on invalid
parameters or a failed solve; `BatchAmericanOptionSolver.solve_batch(params,
use_shared_grid=False)` returns `(results, failed_count)` with
`(success, result|None, SolverError)` tuples.
"""
from __future__ import annotations

import ctypes as C
import enum
import math
from dataclasses import dataclass, field
from typing import Iterable, Sequence

import numpy as np

from . import _abi
from ._abi import (MANGO_CALL, MANGO_PUT, MANGO_OK, MANGO_GRID_MODE_ACCURACY, MANGO_GRID_MODE_EXPLICIT,
                   MANGO_GRID_MODE_ACCURACY_FIXED_TIME, MangoPricingParams, MangoGridConfig, MangoError,
                   MangoAmericanGreeks, MangoGridSpec, MangoGridAccuracy, MangoSinhCluster)
from .workloads import OptionBatch

__all__ = [
    "OptionType", "GridAccuracyProfile", "GridAccuracyParams", "make_grid_accuracy", "PricingParams",
    "Dividend", "TenorPoint", "YieldCurve", "AmericanOptionResult", "american_option_price", "BatchAmericanOptionSolver", "SolverError",
    "SolverErrorCode", "GridSpec", "grid_config", "price_batch", "estimate_pde_grid", "estimate_batch_pde_grid",
    "plan_grid_sizes",
]


class OptionType(enum.IntEnum):
    CALL = MANGO_CALL
    PUT = MANGO_PUT


class GridAccuracyProfile(enum.Enum):
    LOW = "low"
    MEDIUM = "medium"
    HIGH = "high"
    ULTRA = "ultra"


class SolverErrorCode(enum.IntEnum):
    ConvergenceFailure = 0
    LinearSolveFailure = 1
    InvalidConfiguration = 2
    Unknown = 3


@dataclass
class SolverError:
    code: SolverErrorCode = SolverErrorCode.Unknown
    iterations: int = 0
    residual: float = 0.0


@dataclass
class GridAccuracyParams:
    """GridAccuracyParams (src/option/grid_spec_types.hpp:23-53)."""
    n_sigma: float = 5.0
    alpha: float = field(default_factory=_abi.default_sinh_alpha)
    tol: float = 1e-2
    c_t: float = 0.75
    min_spatial_points: int = 100
    max_spatial_points: int = 1200
    max_time_steps: int = 5000

    def to_c(self) -> MangoGridAccuracy:
        return MangoGridAccuracy(self.n_sigma, self.alpha, self.tol, self.c_t, self.min_spatial_points,
                                 self.max_spatial_points, self.max_time_steps)


def make_grid_accuracy(profile: GridAccuracyProfile) -> GridAccuracyParams:
    """make_grid_accuracy (src/option/grid_spec_types.cpp:6-35)."""
    p = GridAccuracyParams()
    table = {
        GridAccuracyProfile.LOW: (5e-3, 150, 1500, 6000),
        GridAccuracyProfile.MEDIUM: (5e-5, 201, 2500, 12000),
        GridAccuracyProfile.HIGH: (1e-5, 301, 3500, 16000),
        GridAccuracyProfile.ULTRA: (5e-6, 401, 5000, 20000),
    }
    p.tol, p.min_spatial_points, p.max_spatial_points, p.max_time_steps = table[profile]
    return p


@dataclass
class Dividend:
    calendar_time: float = 0.0
    amount: float = 0.0


@dataclass
class TenorPoint:
    """TenorPoint (src/option/yield_curve.hpp:19-23): tenor and ln D(tenor)."""
    tenor: float = 0.0
    log_discount: float = 0.0


class YieldCurve:
    """YieldCurve (src/option/yield_curve.hpp:29-180; Python names of
    src/python/mango_bindings.cpp:344-375): log-linear interpolation of the
    discount factor between tenors, i.e. piecewise-constant forward rates.
    Passed as PricingParams.rate; the GPU solve reads its tenor points."""

    def __init__(self, points=None):
        self.points = [] if points is None else [(float(t), float(ld)) for t, ld in points]

    @staticmethod
    def flat(rate: float) -> "YieldCurve":
        return YieldCurve([(0.0, 0.0), (100.0, -rate * 100.0)])

    @staticmethod
    def from_points(points) -> "YieldCurve":
        pts = sorted(((p.tenor, p.log_discount) if isinstance(p, TenorPoint) else (float(p[0]), float(p[1]))
                      for p in points), key=lambda q: q[0])
        if not pts:
            raise ValueError("Empty points vector")
        if pts[0][0] != 0.0:
            raise ValueError("First point must have t=0")
        if abs(pts[0][1]) > 1e-10:
            raise ValueError("log_discount at t=0 must be 0")
        for i in range(1, len(pts)):
            if pts[i][0] <= pts[i - 1][0] + 1e-10:
                raise ValueError("Tenors must be strictly increasing")
        return YieldCurve(pts)

    @staticmethod
    def from_discounts(tenors, discounts) -> "YieldCurve":
        tenors, discounts = list(tenors), list(discounts)
        if len(tenors) != len(discounts):
            raise ValueError("Tenors and discounts must have same size")
        if not tenors:
            raise ValueError("Empty tenors vector")
        if any(d <= 0.0 for d in discounts):
            raise ValueError("Discount factors must be positive")
        return YieldCurve.from_points([(t, math.log(d)) for t, d in zip(tenors, discounts)])

    def _rate_between(self, i: int) -> float:
        if i + 1 >= len(self.points):
            return 0.0
        dt = self.points[i + 1][0] - self.points[i][0]
        return 0.0 if dt <= 0.0 else -(self.points[i + 1][1] - self.points[i][1]) / dt

    def _upper(self, t: float) -> int:
        import bisect
        return bisect.bisect_right([q[0] for q in self.points], t)

    def rate(self, t: float) -> float:
        if len(self.points) < 2:
            return 0.0
        if t <= 0.0:
            return self._rate_between(0)
        ub = self._upper(t)
        if ub == 0:
            return self._rate_between(0)
        if ub == len(self.points):
            return self._rate_between(len(self.points) - 2)
        return self._rate_between(ub - 1)

    def log_discount(self, t: float) -> float:
        pts = self.points
        if len(pts) < 2 or t <= 0.0:
            return 0.0
        ub = self._upper(t)
        if ub == 0:
            return 0.0
        if ub == len(pts):
            r = -(pts[-1][1] - pts[-2][1]) / (pts[-1][0] - pts[-2][0])
            return pts[-1][1] - r * (t - pts[-1][0])
        (t0, l0), (t1, l1) = pts[ub - 1], pts[ub]
        return l0 + (t - t0) / (t1 - t0) * (l1 - l0)

    def discount(self, t: float) -> float:
        return math.exp(self.log_discount(t))

    def zero_rate(self, t: float) -> float:
        return self.rate(0.0) if t <= 0.0 else -self.log_discount(t) / t

    def __repr__(self) -> str:
        return "<YieldCurve>"


@dataclass
class PricingParams:
    """PricingParams (src/option/option_spec.hpp:262-281): rate is a float or
    a YieldCurve; discrete_dividends a list of Dividend."""
    spot: float = 0.0
    strike: float = 0.0
    maturity: float = 0.0
    rate: object = 0.0
    dividend_yield: float = 0.0
    option_type: OptionType = OptionType.CALL
    volatility: float = 0.0
    discrete_dividends: list = field(default_factory=list)


@dataclass
class GridSpec:
    """Mirror of MangoGridSpec; clusters is a list of (center_x, alpha, weight)."""
    type: int = _abi.MANGO_GRID_SINH
    x_min: float = -3.0
    x_max: float = 3.0
    n_points: int = 101
    concentration: float = field(default_factory=_abi.default_sinh_alpha)
    clusters: list = field(default_factory=list)

    def to_c(self) -> MangoGridSpec:
        s = MangoGridSpec()
        s.type = self.type
        s.x_min = self.x_min
        s.x_max = self.x_max
        s.n_points = self.n_points
        s.concentration = self.concentration
        s.n_clusters = len(self.clusters)
        for i, (c, a, w) in enumerate(self.clusters):
            s.clusters[i] = MangoSinhCluster(c, a, w)
        return s

    @staticmethod
    def from_c(s: MangoGridSpec) -> "GridSpec":
        return GridSpec(s.type, s.x_min, s.x_max, s.n_points, s.concentration,
                        [(s.clusters[i].center_x, s.clusters[i].alpha, s.clusters[i].weight)
                         for i in range(s.n_clusters)])

    @staticmethod
    def uniform(x_min: float, x_max: float, n: int) -> "GridSpec":
        return GridSpec(_abi.MANGO_GRID_UNIFORM, x_min, x_max, n, 1.0, [])

    @staticmethod
    def sinh_spaced(x_min: float, x_max: float, n: int, concentration: float = 1.0) -> "GridSpec":
        return GridSpec(_abi.MANGO_GRID_SINH, x_min, x_max, n, concentration, [])

    @staticmethod
    def multi_sinh_spaced(x_min: float, x_max: float, n: int, clusters) -> "GridSpec":
        lib = _abi.load_library()
        arr = (MangoSinhCluster * max(1, len(clusters)))(*[MangoSinhCluster(*c) for c in clusters])
        out = MangoGridSpec()
        err = MangoError()
        st = lib.mango_grid_spec_multi_sinh(x_min, x_max, n, arr, len(clusters), C.byref(out), C.byref(err))
        if st != MANGO_OK:
            raise ValueError(err.message.decode())
        return GridSpec.from_c(out)

    def generate(self) -> np.ndarray:
        lib = _abi.load_library()
        x = np.empty(self.n_points)
        err = MangoError()
        st = lib.mango_grid_generate(C.byref(self.to_c()), x.ctypes.data_as(C.POINTER(C.c_double)), C.byref(err))
        if st != MANGO_OK:
            raise ValueError(err.message.decode())
        return x


def grid_config(mode: int = MANGO_GRID_MODE_ACCURACY, accuracy: GridAccuracyParams | None = None,
                spec: GridSpec | None = None, n_time: int = 1000, min_time_steps: int = 0) -> MangoGridConfig:
    """Build a MangoGridConfig: ACCURACY (GridAccuracyParams), EXPLICIT
    (PDEGridConfig{spec, n_time}) or ACCURACY_FIXED_TIME."""
    cfg = MangoGridConfig()
    _abi.load_library().mango_grid_config_default(C.byref(cfg))
    cfg.mode = mode
    if accuracy is not None:
        cfg.accuracy = accuracy.to_c()
    if spec is not None:
        cfg.spec = spec.to_c()
    cfg.n_time = n_time
    cfg.min_time_steps = min_time_steps
    return cfg


def _c_params(items) -> C.Array:
    if isinstance(items, OptionBatch):
        n = len(items)
        arr = (MangoPricingParams * max(1, n))()
        for i in range(n):
            p = arr[i]
            p.spot = float(items.spot[i]); p.strike = float(items.strike[i])
            p.maturity = float(items.maturity[i]); p.dividend_yield = float(items.dividend_yield[i])
            p.volatility = float(items.volatility[i]); p.rate_const = float(items.rate[i])
            p.option_type = int(items.option_type[i])
        return arr, n
    items = list(items)
    arr = (MangoPricingParams * max(1, len(items)))()
    for i, q in enumerate(items):
        p = arr[i]
        p.spot = q.spot; p.strike = q.strike; p.maturity = q.maturity; p.dividend_yield = q.dividend_yield
        p.volatility = q.volatility; p.option_type = int(q.option_type)
        if isinstance(q.rate, YieldCurve):
            p.rate_const = 0.0
            pts = (_abi.MangoTenorPoint * max(1, len(q.rate.points)))(
                *[_abi.MangoTenorPoint(t, ld) for t, ld in q.rate.points])
            p.tenor_points = C.cast(pts, C.POINTER(_abi.MangoTenorPoint))
            p.n_tenor_points = len(q.rate.points)
            setattr(arr, f"_curve{i}", pts)
        else:
            p.rate_const = q.rate
        if q.discrete_dividends:
            divs = (_abi.MangoDividend * len(q.discrete_dividends))(
                *[_abi.MangoDividend(d.calendar_time, d.amount) for d in q.discrete_dividends])
            p.dividends = C.cast(divs, C.POINTER(_abi.MangoDividend))
            p.n_dividends = len(q.discrete_dividends)
            setattr(arr, f"_keep{i}", divs)
    return arr, len(items)


def price_batch(options, use_shared_grid: bool = False, grid: MangoGridConfig | None = None,
                use_normalized: bool = True):
    """mango_price_american_batch_ex: returns (greeks[n,4], status[n], failed)."""
    lib = _abi.load_library()
    arr, n = _c_params(options)
    out = (MangoAmericanGreeks * max(1, n))()
    status = (C.c_int32 * max(1, n))()
    failed = C.c_uint64(0)
    err = MangoError()
    st = lib.mango_price_american_batch_ex(arr, n, int(use_shared_grid), int(use_normalized),
                                           C.byref(grid) if grid is not None else None, out, status,
                                           C.byref(failed), C.byref(err))
    if st != MANGO_OK:
        raise RuntimeError(f"mango_price_american_batch failed ({st}): {err.message.decode()}")
    g = np.frombuffer(out, dtype=np.float64, count=4 * n).reshape(n, 4).copy()
    s = np.frombuffer(status, dtype=np.int32, count=n).copy()
    return g, s, int(failed.value)


def price_batch_state(options, grid: MangoGridConfig | None = None, stride: int = 512):
    """mango_price_american_batch_state: per-option grids, final and previous
    solutions. Returns dict of arrays."""
    lib = _abi.load_library()
    arr, n = _c_params(options)
    x = np.zeros((n, stride)); u = np.zeros((n, stride)); up = np.zeros((n, stride))
    ns = np.zeros(n, dtype=np.uint64); dt = np.zeros(n)
    out = (MangoAmericanGreeks * max(1, n))()
    status = (C.c_int32 * max(1, n))()
    failed = C.c_uint64(0)
    err = MangoError()
    dp = C.POINTER(C.c_double)
    st = lib.mango_price_american_batch_state(arr, n, C.byref(grid) if grid is not None else None, stride,
                                              x.ctypes.data_as(dp), u.ctypes.data_as(dp), up.ctypes.data_as(dp),
                                              ns.ctypes.data_as(C.POINTER(C.c_uint64)), dt.ctypes.data_as(dp),
                                              out, status, C.byref(failed), C.byref(err))
    if st != MANGO_OK:
        raise RuntimeError(f"mango_price_american_batch_state failed ({st}): {err.message.decode()}")
    g = np.frombuffer(out, dtype=np.float64, count=4 * n).reshape(n, 4).copy()
    return {"x": x, "u": u, "u_prev": up, "n_space": ns.astype(np.int64), "dt": dt, "greeks": g,
            "status": np.frombuffer(status, dtype=np.int32, count=n).copy(), "failed": int(failed.value)}


def estimate_pde_grid(params: PricingParams, accuracy: GridAccuracyParams | None = None):
    """estimate_pde_grid (src/option/grid_spec_types.cpp:37-118) -> (GridSpec, n_time)."""
    lib = _abi.load_library()
    arr, _ = _c_params([params])
    spec = MangoGridSpec()
    nt = C.c_uint64(0)
    err = MangoError()
    acc = (accuracy or GridAccuracyParams()).to_c()
    st = lib.mango_estimate_pde_grid(arr, C.byref(acc), C.byref(spec), C.byref(nt), C.byref(err))
    if st != MANGO_OK:
        raise ValueError(err.message.decode())
    return GridSpec.from_c(spec), int(nt.value)


def plan_grid_sizes(batch: OptionBatch, grid: MangoGridConfig):
    """mango_plan_grid_sizes: per-option (n_space, n_time) the device entry
    points use for `grid` (0, 0 for invalid options)."""
    lib = _abi.load_library()
    rec = np.ascontiguousarray(batch.records())
    n = len(rec)
    ns = np.zeros(n, dtype=np.uint64)
    nt = np.zeros(n, dtype=np.uint64)
    err = MangoError()
    u64 = C.POINTER(C.c_uint64)
    st = lib.mango_plan_grid_sizes(rec.ctypes.data, n, C.byref(grid), ns.ctypes.data_as(u64),
                                   nt.ctypes.data_as(u64), C.byref(err))
    if st != MANGO_OK:
        raise ValueError(err.message.decode())
    return ns.astype(np.int64), nt.astype(np.int64)


def estimate_batch_pde_grid(params: Sequence[PricingParams], accuracy: GridAccuracyParams | None = None):
    lib = _abi.load_library()
    arr, n = _c_params(params)
    spec = MangoGridSpec()
    nt = C.c_uint64(0)
    err = MangoError()
    acc = (accuracy or GridAccuracyParams()).to_c()
    st = lib.mango_estimate_batch_pde_grid(arr, n, C.byref(acc), C.byref(spec), C.byref(nt), C.byref(err))
    if st != MANGO_OK:
        raise ValueError(err.message.decode())
    return GridSpec.from_c(spec), int(nt.value)


class AmericanOptionResult:
    """Result of one solve: at-spot Greeks and value_at(S) over the final
    solution (AmericanOptionResult, src/option/american_option_result.hpp:38-120)."""

    def __init__(self, handle: int | None, params: PricingParams, greeks=None):
        self._h = handle
        self._params = params
        self._g = greeks
        self.converged = True

    def __del__(self):
        if self._h:
            try:
                _abi.load_library().mango_american_result_free(self._h)
            except Exception:
                pass
            self._h = None

    @property
    def spot(self): return self._params.spot

    @property
    def strike(self): return self._params.strike

    def value(self) -> float:
        return self.value_at(self._params.spot)

    def value_at(self, spot: float) -> float:
        if self._h:
            out = C.c_double()
            err = MangoError()
            st = _abi.load_library().mango_american_value_at(self._h, spot, C.byref(out), C.byref(err))
            if st != MANGO_OK:
                raise ValueError(err.message.decode())
            return out.value
        if spot != self._params.spot:
            raise ValueError("batch results hold the at-spot quantities only")
        return float(self._g[0])

    def delta(self) -> float:
        return _abi.load_library().mango_american_delta(self._h) if self._h else float(self._g[1])

    def gamma(self) -> float:
        return _abi.load_library().mango_american_gamma(self._h) if self._h else float(self._g[2])

    def theta(self) -> float:
        return _abi.load_library().mango_american_theta(self._h) if self._h else float(self._g[3])


def _validation_message(p: PricingParams) -> str | None:
    if not (p.spot > 0 and math.isfinite(p.spot)):
        return f"Invalid pricing parameters: spot price must be positive (value={p.spot})"
    if not (p.strike > 0 and math.isfinite(p.strike)):
        return f"Invalid pricing parameters: strike must be positive (value={p.strike})"
    if not (p.maturity > 0 and math.isfinite(p.maturity)):
        return f"Invalid pricing parameters: maturity must be positive (value={p.maturity})"
    if not (p.volatility > 0 and math.isfinite(p.volatility)):
        return f"Invalid pricing parameters: volatility must be positive (value={p.volatility})"
    if not isinstance(p.rate, YieldCurve) and not math.isfinite(p.rate):
        return f"Invalid pricing parameters: invalid rate (value={p.rate})"
    if not (p.dividend_yield >= 0 and math.isfinite(p.dividend_yield)):
        return f"Invalid pricing parameters: invalid dividend yield (value={p.dividend_yield})"
    return None


def american_option_price(params: PricingParams, accuracy: GridAccuracyProfile | None = None) -> AmericanOptionResult:
    """american_option_price (src/python/mango_bindings.cpp:580-655)."""
    msg = _validation_message(params)
    if msg:
        raise ValueError(msg)
    lib = _abi.load_library()
    if accuracy is None:
        arr, _ = _c_params([params])
        h = C.c_void_p()
        err = MangoError()
        st = lib.mango_price_american(arr, C.byref(h), C.byref(err))
        if st != MANGO_OK:
            raise ValueError(f"American option solve failed (error code {st}): {err.message.decode()}")
        return AmericanOptionResult(h.value, params)
    cfg = grid_config(MANGO_GRID_MODE_ACCURACY, make_grid_accuracy(accuracy))
    g, s, _ = price_batch([params], False, cfg, use_normalized=False)
    if s[0] != MANGO_OK:
        raise ValueError(f"American option solve failed (error code {int(s[0])})")
    return AmericanOptionResult(None, params, g[0])


class BatchAmericanOptionSolver:
    """BatchAmericanOptionSolver (src/python/mango_bindings.cpp:681-757)."""

    def __init__(self):
        self._accuracy = GridAccuracyParams()
        self._use_normalized = True

    def set_grid_accuracy(self, profile: GridAccuracyProfile):
        self._accuracy = make_grid_accuracy(profile)
        return self

    def set_grid_accuracy_params(self, params: GridAccuracyParams):
        self._accuracy = params
        return self

    def set_use_normalized(self, enable: bool = True):
        self._use_normalized = enable
        return self

    def solve_batch(self, params: Sequence[PricingParams], use_shared_grid: bool = False):
        params = list(params)
        if not params:
            return [], 0
        cfg = grid_config(MANGO_GRID_MODE_ACCURACY, self._accuracy)
        g, s, failed = price_batch(params, use_shared_grid, cfg, self._use_normalized)
        results = []
        for i, p in enumerate(params):
            if s[i] == MANGO_OK:
                results.append((True, AmericanOptionResult(None, p, g[i]), SolverError()))
            else:
                code = SolverErrorCode.InvalidConfiguration if s[i] == _abi.MANGO_ERR_VALIDATION \
                    else SolverErrorCode.LinearSolveFailure
                results.append((False, None, SolverError(code)))
        return results, failed


from .table import (BSplineBackend, InterpolatedIVSolver, InterpolatedIVSolverConfig, IVError, IVGrid,  # noqa: E402
                    IVQuery, IVSolver, IVSolverConfig, IVSolverFactoryConfig, IVSuccess, PriceTable,
                    PriceTableBuilder, PriceTableConfig, RootFindingConfig, make_interpolated_iv_solver,
                    make_price_table)

__all__ += ["BSplineBackend", "InterpolatedIVSolver", "InterpolatedIVSolverConfig", "IVError", "IVGrid", "IVQuery",
            "IVSolver", "IVSolverConfig", "IVSolverFactoryConfig", "IVSuccess", "PriceTable", "PriceTableBuilder",
            "PriceTableConfig", "RootFindingConfig", "make_interpolated_iv_solver", "make_price_table"]
