# SPDX-License-Identifier: MIT
"""Price tables and the interpolated implied-volatility solver, on the GPU.

Mirrors the reference's Python surface (src/python/mango_bindings.cpp:
832-1170): `make_price_table(config)` and `make_interpolated_iv_solver(config)`
from an `IVSolverFactoryConfig` (B-spline backend), `PriceTable.price / vega /
make_iv_solver / solve_iv`, `InterpolatedIVSolver.solve / solve_batch` with
`(success, IVSuccess, IVError)` tuples. `PriceTableBuilder.from_vectors(...)
.build()` is the C++ builder entry (bspline_builder.cpp:544-594) with an
explicit PDE grid. Device-buffer forms (`price_device`, `solve_device`) take
torch tensors and run on the current stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from ._abi import (MANGO_OK, MANGO_PUT, MangoError, MangoGridConfig, MangoIvFactoryConfig, MangoIvQuery,
                   MangoIvSuccess, MangoInterpSolverConfig, MangoIvBatchSlot, MangoPricingParams,
                   MangoTableBuildInfo)

__all__ = ["InterpolatedIVSolverConfig", "IVGrid", "BSplineBackend", "IVSolverFactoryConfig", "PriceTableConfig",
           "IVQuery", "IVSuccess", "IVError", "PriceTable", "InterpolatedIVSolver", "make_price_table",
           "make_interpolated_iv_solver", "PriceTableBuilder", "IV_QUERY_DTYPE", "IV_RESULT_DTYPE"]

_dp = C.POINTER(C.c_double)

# MangoDeviceIvQuery / MangoDeviceIvResult as numpy records
IV_QUERY_DTYPE = np.dtype([("spot", "<f8"), ("strike", "<f8"), ("maturity", "<f8"), ("rate", "<f8"),
                           ("dividend_yield", "<f8"), ("market_price", "<f8"), ("option_type", "<i4"),
                           ("reserved", "<i4")])
IV_RESULT_DTYPE = np.dtype([("implied_vol", "<f8"), ("final_error", "<f8"), ("iterations", "<i4"),
                            ("status", "<i4")])


@dataclass
class InterpolatedIVSolverConfig:
    max_iter: int = 50
    tolerance: float = 1e-6
    sigma_min: float = 0.01
    sigma_max: float = 3.0
    vega_threshold: float = 1e-4

    def to_c(self) -> MangoInterpSolverConfig:
        return MangoInterpSolverConfig(int(self.max_iter), self.tolerance, self.sigma_min, self.sigma_max,
                                       self.vega_threshold)


@dataclass
class IVGrid:
    moneyness: list = field(default_factory=list)  # S / K
    vol: list = field(default_factory=list)
    rate: list = field(default_factory=list)


@dataclass
class BSplineBackend:
    maturity_grid: list = field(default_factory=list)


@dataclass
class IVSolverFactoryConfig:
    option_type: int = MANGO_PUT
    spot: float = 100.0
    dividend_yield: float = 0.0
    grid: IVGrid = field(default_factory=IVGrid)
    backend: BSplineBackend = field(default_factory=BSplineBackend)
    solver_config: InterpolatedIVSolverConfig = field(default_factory=InterpolatedIVSolverConfig)
    adaptive: object = None
    discrete_dividends: object = None


PriceTableConfig = IVSolverFactoryConfig


@dataclass
class IVQuery:
    spot: float = 0.0
    strike: float = 0.0
    maturity: float = 0.0
    rate: float = 0.0
    dividend_yield: float = 0.0
    option_type: int = MANGO_PUT
    market_price: float = 0.0
    discrete_dividends: list = field(default_factory=list)  # honored by the FDM IVSolver


@dataclass
class IVSuccess:
    implied_vol: float = 0.0
    iterations: int = 0
    final_error: float = 0.0
    vega: float | None = None
    used_rate_approximation: bool = False


@dataclass
class IVError:
    code: int = 0  # MangoStatus of the failure
    iterations: int = 0
    final_error: float = 0.0
    last_vol: float | None = None
    message: str = ""


def _raise(rc: int, err: MangoError, what: str):
    if rc != MANGO_OK:
        raise ValueError(f"{what} failed ({rc}): {err.message.decode(errors='replace')}")


def _darr(v):
    a = np.ascontiguousarray(v, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _set_rate(c, rate) -> None:
    """rate_const, or the tenor points of a YieldCurve (the array is kept on
    the struct object; the library collapses the curve to its zero rate at
    the maturity, like the reference's surfaces)."""
    from . import YieldCurve
    if isinstance(rate, YieldCurve):
        pts = (_abi.MangoTenorPoint * max(1, len(rate.points)))(*[_abi.MangoTenorPoint(t, ld) for t, ld in rate.points])
        c._tenors = pts
        c.rate_const = 0.0
        c.tenor_points = C.cast(pts, C.POINTER(_abi.MangoTenorPoint))
        c.n_tenor_points = len(rate.points)
    else:
        c.rate_const = rate


def _c_query(q: IVQuery) -> MangoIvQuery:
    c = MangoIvQuery()
    c.spot, c.strike, c.maturity = q.spot, q.strike, q.maturity
    c.dividend_yield, c.market_price = q.dividend_yield, q.market_price
    _set_rate(c, q.rate)
    c.option_type = int(q.option_type)
    return c


def _c_pricing(p) -> MangoPricingParams:
    c = MangoPricingParams()
    c.spot, c.strike, c.maturity = p.spot, p.strike, p.maturity
    c.dividend_yield, c.volatility = p.dividend_yield, p.volatility
    _set_rate(c, p.rate)
    c.option_type = int(p.option_type)
    return c


def _factory(config: IVSolverFactoryConfig):
    keep = []
    c = MangoIvFactoryConfig()
    c.option_type = int(config.option_type)
    c.spot = config.spot
    c.dividend_yield = config.dividend_yield
    for name, vals in (("moneyness", config.grid.moneyness), ("vol", config.grid.vol), ("rate", config.grid.rate),
                       ("maturity_grid", config.backend.maturity_grid)):
        a, p = _darr(vals)
        keep.append(a)
        setattr(c, name, p)
        setattr(c, "n_" + ("maturity" if name == "maturity_grid" else name), len(a))
    c.solver_config = config.solver_config.to_c()
    if config.adaptive is not None:
        ad = _abi.MangoAdaptiveGridParams()
        keep.append(ad)
        c.adaptive = C.pointer(ad)
    if config.discrete_dividends is not None:
        dd = _abi.MangoDiscreteDividendConfig()
        keep.append(dd)
        c.discrete_dividends = C.pointer(dd)
    return c, keep


class InterpolatedIVSolver:
    """InterpolatedIVSolver over a device-resident table."""

    def __init__(self, handle: int, table: "PriceTable | None" = None):
        self._h = C.c_void_p(handle)
        self._table = table  # keeps the table alive
        self._lib = _abi.load_library()

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.mango_interp_iv_solver_free(self._h)
            self._h = C.c_void_p()

    def solve(self, query: IVQuery):
        out = MangoIvSuccess()
        err = MangoError()
        rc = self._lib.mango_interp_iv_solve(self._h, C.byref(_c_query(query)), C.byref(out), C.byref(err))
        if rc == MANGO_OK:
            return True, IVSuccess(out.implied_vol, int(out.iterations), out.final_error,
                                   used_rate_approximation=bool(out.used_rate_approximation)), IVError()
        return False, IVSuccess(), IVError(code=rc, message=err.message.decode(errors="replace"))

    def solve_batch(self, queries):
        queries = list(queries)
        n = len(queries)
        cq = [_c_query(q) for q in queries]  # alive until the call returns (tenor arrays)
        arr = (MangoIvQuery * max(1, n))(*cq)
        slots = (MangoIvBatchSlot * max(1, n))()
        failed = C.c_uint64()
        err = MangoError()
        rc = self._lib.mango_interp_iv_solve_batch(self._h, arr, n, slots, C.byref(failed), C.byref(err))
        _raise(rc, err, "mango_interp_iv_solve_batch")
        res = []
        for i in range(n):
            s = slots[i]
            if s.status == MANGO_OK:
                res.append((True, IVSuccess(s.success.implied_vol, int(s.success.iterations),
                                            s.success.final_error,
                                            used_rate_approximation=bool(s.success.used_rate_approximation)),
                             IVError()))
            else:
                res.append((False, IVSuccess(), IVError(code=s.status)))
        return res, int(failed.value)

    def solve_arrays(self, spot, strike, maturity, rate, dividend_yield, market_price, option_type=MANGO_PUT):
        """Host arrays in, (implied_vol, iterations, final_error, status) numpy
        arrays out, through the device-buffer entry point."""
        import torch
        n = len(np.atleast_1d(spot))
        rec = np.zeros(n, dtype=IV_QUERY_DTYPE)
        for k, v in (("spot", spot), ("strike", strike), ("maturity", maturity), ("rate", rate),
                     ("dividend_yield", dividend_yield), ("market_price", market_price),
                     ("option_type", option_type)):
            rec[k] = v
        d_q = torch.from_numpy(rec.view(np.uint8).reshape(n, IV_QUERY_DTYPE.itemsize)).cuda()
        d_r = self.solve_device(d_q)
        torch.cuda.synchronize()
        out = d_r.cpu().numpy().reshape(-1).view(IV_RESULT_DTYPE)
        return out["implied_vol"].copy(), out["iterations"].copy(), out["final_error"].copy(), out["status"].copy()

    def solve_device(self, d_queries, d_out=None):
        """d_queries: uint8 CUDA tensor [n, 56] of MangoDeviceIvQuery records;
        returns (or fills) a uint8 tensor [n, 24] of MangoDeviceIvResult,
        asynchronously on torch's current stream."""
        import torch
        n = d_queries.shape[0]
        if d_out is None:
            d_out = torch.empty((n, IV_RESULT_DTYPE.itemsize), dtype=torch.uint8, device=d_queries.device)
        err = MangoError()
        stream = C.c_void_p(torch.cuda.current_stream(d_queries.device).cuda_stream)
        rc = self._lib.mango_interp_iv_solve_device(self._h, C.c_void_p(d_queries.data_ptr()), n,
                                                    C.c_void_p(d_out.data_ptr()), stream, C.byref(err))
        _raise(rc, err, "mango_interp_iv_solve_device")
        return d_out


class PriceTable:
    """A fitted EEP table resident in device memory (AnyPriceTable)."""

    def __init__(self, handle: int, info: MangoTableBuildInfo | None = None):
        self._h = C.c_void_p(handle)
        self._lib = _abi.load_library()
        self.build_info = info

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.mango_price_table_free(self._h)
            self._h = C.c_void_p()

    @property
    def option_type(self) -> int:
        return int(self._lib.mango_price_table_option_type(self._h))

    @property
    def dividend_yield(self) -> float:
        return float(self._lib.mango_price_table_dividend_yield(self._h))

    @property
    def surface_type(self) -> str:
        return "bspline_4d"

    def _validate(self, params):
        err = MangoError()
        rc = self._lib.mango_price_table_validate(self._h, C.byref(_c_pricing(params)), C.byref(err))
        _raise(rc, err, "price table validation")

    def price(self, params) -> float:
        self._validate(params)
        return float(self._lib.mango_price_table_price(self._h, C.byref(_c_pricing(params))))

    def vega(self, params) -> float:
        self._validate(params)
        return float(self._lib.mango_price_table_vega(self._h, C.byref(_c_pricing(params))))

    def _greek(self, name, params) -> float:
        out = C.c_double()
        err = MangoError()
        rc = getattr(self._lib, f"mango_price_table_{name}")(self._h, C.byref(_c_pricing(params)), C.byref(out),
                                                             C.byref(err))
        _raise(rc, err, name)
        return out.value

    def delta(self, params) -> float:
        return self._greek("delta", params)

    def gamma(self, params) -> float:
        return self._greek("gamma", params)

    def theta(self, params) -> float:
        return self._greek("theta", params)

    def rho(self, params) -> float:
        return self._greek("rho", params)

    def greeks_points(self, rows):
        """rows [n, 5] = (S, K, tau, sigma, r) -> [n, 4] delta, gamma, theta,
        rho (numpy), evaluated on the device without domain validation."""
        import torch
        rows = np.ascontiguousarray(rows, dtype=np.float64).reshape(-1, 5)
        d_rows = torch.from_numpy(rows).cuda()
        d_out = torch.empty((rows.shape[0], 4), dtype=torch.float64, device=d_rows.device)
        err = MangoError()
        stream = C.c_void_p(torch.cuda.current_stream(d_rows.device).cuda_stream)
        rc = self._lib.mango_price_table_greeks_device(self._h, C.c_void_p(d_rows.data_ptr()), rows.shape[0],
                                                       C.c_void_p(d_out.data_ptr()), stream, C.byref(err))
        _raise(rc, err, "mango_price_table_greeks_device")
        torch.cuda.synchronize()
        return d_out.cpu().numpy()

    def make_iv_solver(self, solver_config: InterpolatedIVSolverConfig | None = None) -> InterpolatedIVSolver:
        out = C.c_void_p()
        err = MangoError()
        cfg = (solver_config or InterpolatedIVSolverConfig()).to_c()
        rc = self._lib.mango_price_table_make_iv_solver(self._h, C.byref(cfg), C.byref(out), C.byref(err))
        _raise(rc, err, "make_iv_solver")
        return InterpolatedIVSolver(out.value, self)

    def solve_iv(self, query: IVQuery, solver_config: InterpolatedIVSolverConfig | None = None) -> IVSuccess:
        ok, res, e = self.make_iv_solver(solver_config).solve(query)
        if not ok:
            raise ValueError(f"interpolated IV failed ({e.code}): {e.message}")
        return res

    # ---- additions ----
    def shape(self):
        dims = (C.c_uint64 * 4)()
        kref = C.c_double()
        err = MangoError()
        _raise(self._lib.mango_price_table_shape(self._h, dims, C.byref(kref), C.byref(err)), err, "shape")
        return tuple(int(d) for d in dims), kref.value

    def export(self):
        """(axes list, coefficients [n_m, n_tau, n_vol, n_rate], prices V/K_ref
        before the EEP transform) copied to host memory."""
        dims, _ = self.shape()
        axes = np.empty(sum(dims))
        size = int(np.prod(dims))
        coeffs, prices = np.empty(size), np.empty(size)
        err = MangoError()
        rc = self._lib.mango_price_table_export(self._h, axes.ctypes.data_as(_dp), coeffs.ctypes.data_as(_dp),
                                                prices.ctypes.data_as(_dp), C.byref(err))
        _raise(rc, err, "export")
        out, k = [], 0
        for d in dims:
            out.append(axes[k:k + d].copy())
            k += d
        return out, coeffs.reshape(dims), prices.reshape(dims)

    def price_points(self, rows):
        """rows [n, 5] = (S, K, tau, sigma, r) -> (price, vega) numpy arrays,
        evaluated on the device."""
        import torch
        rows = np.ascontiguousarray(rows, dtype=np.float64).reshape(-1, 5)
        d_rows = torch.from_numpy(rows).cuda()
        pr, vg = self.price_device(d_rows)
        torch.cuda.synchronize()
        return pr.cpu().numpy(), vg.cpu().numpy()

    def price_device(self, d_rows, want_vega=True):
        """d_rows: float64 CUDA tensor [n, 5] = (S, K, tau, sigma, r); returns
        (price, vega) tensors, asynchronously on torch's current stream."""
        import torch
        n = d_rows.shape[0]
        pr = torch.empty(n, dtype=torch.float64, device=d_rows.device)
        vg = torch.empty(n, dtype=torch.float64, device=d_rows.device) if want_vega else None
        err = MangoError()
        stream = C.c_void_p(torch.cuda.current_stream(d_rows.device).cuda_stream)
        rc = self._lib.mango_price_table_price_device(self._h, C.c_void_p(d_rows.data_ptr()), n,
                                                      C.c_void_p(pr.data_ptr()),
                                                      C.c_void_p(vg.data_ptr() if vg is not None else 0), stream,
                                                      C.byref(err))
        _raise(rc, err, "mango_price_table_price_device")
        return pr, vg


def make_price_table(config: IVSolverFactoryConfig) -> PriceTable:
    c, keep = _factory(config)
    out = C.c_void_p()
    err = MangoError()
    rc = _abi.load_library().mango_make_price_table(C.byref(c), C.byref(out), C.byref(err))
    _raise(rc, err, "make_price_table")
    return PriceTable(out.value)


def make_interpolated_iv_solver(config: IVSolverFactoryConfig) -> InterpolatedIVSolver:
    c, keep = _factory(config)
    out = C.c_void_p()
    err = MangoError()
    rc = _abi.load_library().mango_make_interp_iv_solver(C.byref(c), C.byref(out), C.byref(err))
    _raise(rc, err, "make_interpolated_iv_solver")
    return InterpolatedIVSolver(out.value)


class PriceTableBuilder:
    """PriceTableBuilder::from_vectors(...).build(axes, EEP transform)."""

    def __init__(self, axes, K_ref, pde_grid: MangoGridConfig | None, option_type, dividend_yield,
                 max_failure_rate):
        self.axes = [np.ascontiguousarray(a, dtype=np.float64) for a in axes]
        self.K_ref, self.pde_grid, self.option_type = float(K_ref), pde_grid, int(option_type)
        self.dividend_yield, self.max_failure_rate = float(dividend_yield), float(max_failure_rate)

    @staticmethod
    def from_vectors(log_moneyness, maturity, volatility, rate, K_ref, pde_grid: MangoGridConfig | None = None,
                     option_type=MANGO_PUT, dividend_yield=0.0, max_failure_rate=0.0) -> "PriceTableBuilder":
        return PriceTableBuilder((log_moneyness, maturity, volatility, rate), K_ref, pde_grid, option_type,
                                 dividend_yield, max_failure_rate)

    def build(self) -> PriceTable:
        a = self.axes
        out = C.c_void_p()
        info = MangoTableBuildInfo()
        err = MangoError()
        rc = _abi.load_library().mango_price_table_build(
            a[0].ctypes.data_as(_dp), len(a[0]), a[1].ctypes.data_as(_dp), len(a[1]), a[2].ctypes.data_as(_dp),
            len(a[2]), a[3].ctypes.data_as(_dp), len(a[3]), self.K_ref, self.dividend_yield, self.option_type,
            C.byref(self.pde_grid) if self.pde_grid is not None else None, self.max_failure_rate, C.byref(out),
            C.byref(info), C.byref(err))
        _raise(rc, err, "mango_price_table_build")
        return PriceTable(out.value, info)


# ---------------------------------------------------------------------------
# FDM implied volatility (IVSolver; Python names of
# src/python/mango_bindings.cpp:420-521)
# ---------------------------------------------------------------------------
@dataclass
class RootFindingConfig:
    max_iter: int = 100
    tolerance: float = 1e-6
    jacobian_fd_epsilon: float = 1e-7
    brent_tol_abs: float = 1e-6


@dataclass
class IVSolverConfig:
    root_config: RootFindingConfig = field(default_factory=RootFindingConfig)
    batch_parallel_threshold: int = 4


class IVSolver:
    """IVSolver (src/option/iv_solver.hpp:70-198): Brent's method over full
    American PDE solves on the GPU (mango_solve_iv / mango_solve_iv_batch);
    yield curves and discrete dividends of the query are priced exactly."""

    def __init__(self, config: IVSolverConfig | None = None):
        self.config = config if config is not None else IVSolverConfig()
        self._lib = _abi.load_library()

    def _cfg(self):
        c = _abi.MangoIvConfig()
        c.max_iter = int(self.config.root_config.max_iter)
        c.brent_tol_abs = float(self.config.root_config.brent_tol_abs)
        return c

    @staticmethod
    def _c_fdm_query(q: IVQuery):
        c = _c_query(q)
        divs = q.discrete_dividends or []
        if divs:
            arr = (_abi.MangoDividend * len(divs))(*[_abi.MangoDividend(d.calendar_time, d.amount) for d in divs])
            c._divs = arr
            c.dividends = C.cast(arr, C.POINTER(_abi.MangoDividend))
            c.n_dividends = len(divs)
        return c

    def solve(self, query: IVQuery):
        """Returns (success, IVSuccess, IVError) like the reference binding."""
        out = _abi.MangoIvSuccess()
        err = MangoError()
        rc = self._lib.mango_solve_iv(C.byref(self._c_fdm_query(query)), C.byref(self._cfg()), C.byref(out),
                                      C.byref(err))
        if rc == MANGO_OK:
            return True, IVSuccess(out.implied_vol, int(out.iterations), out.final_error), IVError()
        return False, IVSuccess(), IVError(code=rc, message=err.message.decode(errors="replace"))

    def solve_batch(self, queries):
        """IVSolver::solve_batch: ([(success, IVSuccess, IVError)], failed_count)."""
        queries = list(queries)
        n = len(queries)
        if n == 0:
            return [], 0
        cq = [self._c_fdm_query(q) for q in queries]
        arr = (MangoIvQuery * n)(*cq)
        slots = (MangoIvBatchSlot * n)()
        failed = C.c_uint64()
        err = MangoError()
        rc = self._lib.mango_solve_iv_batch(arr, n, C.byref(self._cfg()), slots, C.byref(failed), C.byref(err))
        _raise(rc, err, "mango_solve_iv_batch")
        res = []
        for s in slots:
            if s.status == MANGO_OK:
                res.append((True, IVSuccess(s.success.implied_vol, int(s.success.iterations), s.success.final_error),
                            IVError()))
            else:
                res.append((False, IVSuccess(), IVError(code=s.status)))
        return res, int(failed.value)
