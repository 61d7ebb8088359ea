# SPDX-License-Identifier: MIT
"""PyTorch interop for the device solver: the American price as a
differentiable function of (sigma, r, q, K), with reverse-mode sensitivities
from the gfx950 adjoint kernel, and gradient-Newton implied-volatility
calibration on top of it.

The reference prices one option per call through pybind11
(src/python/mango_bindings.cpp:580-757) and finds implied volatilities with
Brent over full PDE solves (src/option/iv_solver.cpp:48-209; 5-8 solves per
query). Here a batch is priced and differentiated in one device call: the
price, delta, gamma, theta and dV/dsigma, dV/dr, dV/dq, dV/dK on each
option's grid (the grid the solve uses, held fixed for the derivative).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import MangoGridConfig

_REC_BYTES = 56  # MangoDeviceOption


def _torch():
    import torch
    return torch


def pack_options(spot, strike, maturity, volatility, rate, dividend, option_type):
    """MangoDeviceOption records (uint8 [n, 56]) on the device of `volatility`
    from float64 tensors (broadcast to a common length) and an int tensor of
    option types (0 call, 1 put)."""
    torch = _torch()
    cols = torch.broadcast_tensors(*(torch.as_tensor(c, dtype=torch.float64, device=volatility.device)
                                     for c in (spot, strike, maturity, volatility, rate, dividend)))
    n = cols[0].numel()
    rec = torch.zeros((n, _REC_BYTES), dtype=torch.uint8, device=volatility.device)
    rec[:, :48].view(torch.float64).copy_(torch.stack([c.reshape(n) for c in cols], dim=1))
    typ = torch.as_tensor(option_type, dtype=torch.int32, device=volatility.device).expand(n)
    rec[:, 48:52].view(torch.int32).copy_(typ.reshape(n, 1))
    return rec


def adjoint_device(records, grid: MangoGridConfig):
    """mango_price_american_adjoint_device on packed records: returns
    (out float64 [n, 8] = value, delta, gamma, theta, dV/dsigma, dV/dr,
    dV/dq, dV/dK; status int32 [n]), asynchronous on torch's current stream."""
    torch = _torch()
    lib = _abi.load_library()
    n = records.shape[0]
    out = torch.empty((n, 8), dtype=torch.float64, device=records.device)
    status = torch.empty(n, dtype=torch.int32, device=records.device)
    err = _abi.MangoError()
    stream = C.c_void_p(torch.cuda.current_stream(records.device).cuda_stream)
    rc = lib.mango_price_american_adjoint_device(records.data_ptr(), n, C.byref(grid), out.data_ptr(),
                                                 status.data_ptr(), stream, C.byref(err))
    if rc != 0:
        raise RuntimeError(f"mango_price_american_adjoint_device failed ({rc}): {err.message.decode()}")
    return out, status


def _default_grid():
    from . import GridAccuracyParams, grid_config
    acc = GridAccuracyParams(min_spatial_points=101, max_spatial_points=101)
    return grid_config(_abi.MANGO_GRID_MODE_ACCURACY_FIXED_TIME, acc, n_time=498)


def _price_function():
    torch = _torch()

    class AmericanPrice(torch.autograd.Function):
        """V = American price per option. Differentiable in sigma, r, q and K
        (spot and maturity are data); the gradients are the device adjoint's
        sensitivities on each option's fixed grid. Options that fail to price
        return NaN."""

        @staticmethod
        def forward(ctx, volatility, rate, dividend, strike, spot, maturity, option_type, grid):
            rec = pack_options(spot, strike, maturity, volatility, rate, dividend, option_type)
            out, status = adjoint_device(rec, grid)
            ctx.save_for_backward(out[:, 4:8])
            ctx.shapes = [t.shape if torch.is_tensor(t) else None for t in (volatility, rate, dividend, strike)]
            ctx.mark_non_differentiable(status)
            return out[:, 0].reshape(volatility.shape), status

        @staticmethod
        def backward(ctx, g_value, _g_status):
            (sens,) = ctx.saved_tensors
            g = g_value.reshape(-1, 1) * sens
            grads = []
            for j, shape in enumerate(ctx.shapes):
                if not ctx.needs_input_grad[j]:
                    grads.append(None)
                    continue
                gj = g[:, j]
                grads.append(gj.reshape(shape) if shape is not None and gj.numel() == int(np.prod(shape))
                             else gj.sum().reshape(shape))
            return (*grads, None, None, None, None)

    return AmericanPrice


_AMERICAN_PRICE = None


def american_price(volatility, rate, dividend, strike, spot, maturity, option_type, grid: MangoGridConfig | None = None):
    """Differentiable American prices (float64 tensor, shape of `volatility`)
    and per-option status. Gradients w.r.t. volatility, rate, dividend and
    strike come from the device adjoint (one reverse sweep per option)."""
    global _AMERICAN_PRICE
    if _AMERICAN_PRICE is None:
        _AMERICAN_PRICE = _price_function()
    return _AMERICAN_PRICE.apply(volatility, rate, dividend, strike, spot, maturity, option_type,
                                 grid if grid is not None else _default_grid())


def american_greeks(volatility, rate, dividend, strike, spot, maturity, option_type, grid: MangoGridConfig | None = None):
    """All outputs of one adjoint call as a dict of tensors: value, delta,
    gamma, theta, vega (dV/dsigma), rho (dV/dr), dv_ddividend, dv_dstrike,
    status."""
    rec = pack_options(spot, strike, maturity, volatility, rate, dividend, option_type)
    out, status = adjoint_device(rec, grid if grid is not None else _default_grid())
    names = ("value", "delta", "gamma", "theta", "vega", "rho", "dv_ddividend", "dv_dstrike")
    res = {k: out[:, i] for i, k in enumerate(names)}
    res["status"] = status
    return res


def implied_vol_newton(target, spot, strike, maturity, rate, dividend, option_type,
                       grid: MangoGridConfig | None = None, sigma0=0.3, tol=1e-9, max_iter=30,
                       sigma_bounds=(0.01, 3.0), min_vega=1e-10):
    """Gradient-Newton implied volatility, batched on the device: each
    iteration prices every unconverged option with its dV/dsigma in one
    adjoint call, then sigma <- clamp(sigma - (V - target) / vega).
    Converged when |V - target| <= tol * strike. Options whose vega falls
    below min_vega (price at intrinsic or at a bound) stop there.

    Returns (sigma, iterations, converged, pde_solves): sigma, converged
    and the per-option iteration counts as tensors, pde_solves the number of
    adjoint solves run."""
    torch = _torch()
    target = torch.as_tensor(target, dtype=torch.float64)
    dev = target.device
    n = target.numel()
    f64 = lambda v: torch.as_tensor(v, dtype=torch.float64, device=dev).expand(n).clone()
    spot, strike, maturity, rate, dividend = (f64(v) for v in (spot, strike, maturity, rate, dividend))
    typ = torch.as_tensor(option_type, dtype=torch.int32, device=dev).expand(n).clone()
    sigma = f64(sigma0)
    lo, hi = sigma_bounds
    active = torch.ones(n, dtype=torch.bool, device=dev)
    converged = torch.zeros(n, dtype=torch.bool, device=dev)
    iters = torch.zeros(n, dtype=torch.int32, device=dev)
    grid = grid if grid is not None else _default_grid()
    solves = 0
    for _ in range(max_iter):
        idx = torch.nonzero(active).flatten()
        if idx.numel() == 0:
            break
        iters[idx] += 1
        rec = pack_options(spot[idx], strike[idx], maturity[idx], sigma[idx], rate[idx], dividend[idx], typ[idx])
        out, status = adjoint_device(rec, grid)
        solves += idx.numel()
        v, vega = out[:, 0], out[:, 4]
        resid = v - target[idx]
        ok = status == 0
        done = ok & (resid.abs() <= tol * strike[idx])
        flat = ok & ~done & (vega.abs() < min_vega)
        step = torch.where(ok & ~done & ~flat, resid / torch.where(vega.abs() < min_vega, torch.ones_like(vega), vega),
                           torch.zeros_like(vega))
        sigma[idx] = torch.clamp(sigma[idx] - step, lo, hi)
        converged[idx] = done
        active[idx] = ok & ~done & ~flat
    return sigma, iters, converged, solves
