# SPDX-License-Identifier: MIT
"""GPU tests of the PyTorch layer: the differentiable American price
(torch.autograd.Function over the device adjoint) against the oracle's
reverse sweep, and gradient-Newton implied-vol calibration recovering hidden
volatilities from oracle-priced targets."""
import numpy as np
import pytest

import pyoracle as O
import mango_option as M
from mango_option import _abi, workloads as W

pytestmark = pytest.mark.gpu


def _cfg():
    acc = M.GridAccuracyParams(min_spatial_points=101, max_spatial_points=101)
    return M.grid_config(_abi.MANGO_GRID_MODE_ACCURACY_FIXED_TIME, acc, n_time=498)


def test_autograd_gradients_match_oracle_adjoint():
    print("test_autograd_gradients_match_oracle_adjoint: 48 puts + 16 calls", flush=True)
    import torch
    from mango_option import torch_ops as T
    puts = W.c5_puts(48, seed=5)
    chains, _ = W.c2_chains(1, 32, seed=1)
    calls = chains.take(np.arange(1, 32, 2))
    opts = W.OptionBatch(*(np.concatenate([getattr(puts, f), getattr(calls, f)]) for f in
                           ("spot", "strike", "maturity", "volatility", "rate", "dividend_yield", "option_type")))
    dev = torch.device("cuda")
    t = lambda a, g=False: torch.tensor(a, dtype=torch.float64, device=dev, requires_grad=g)
    sig, r, q, k = t(opts.volatility, True), t(opts.rate, True), t(opts.dividend_yield, True), t(opts.strike, True)
    v, st = T.american_price(sig, r, q, k, t(opts.spot), t(opts.maturity), torch.tensor(opts.option_type, device=dev),
                             _cfg())
    (v * 2.0).sum().backward()
    ref, ref_st = O.adjoint_batch(opts, _cfg(), n_threads=8)
    assert (st.cpu().numpy() == 0).all() and (ref_st == 0).all()
    assert np.max(np.abs(v.detach().cpu().numpy() - ref[:, 0])) <= 1e-9
    for grad, col in ((sig.grad, 2), (r.grad, 3), (q.grad, 4), (k.grad, 5)):
        g = grad.cpu().numpy() / 2.0
        assert np.all(np.abs(g - ref[:, col]) <= 1e-7 * np.abs(ref[:, col]) + 1e-9), col


def test_greeks_dict_and_scalar_broadcast():
    print("test_greeks_dict_and_scalar_broadcast", flush=True)
    import torch
    from mango_option import torch_ops as T
    opts = W.c5_puts(8, seed=5)
    dev = torch.device("cuda")
    sig = torch.tensor(opts.volatility, dtype=torch.float64, device=dev)
    rate = torch.tensor(0.03, dtype=torch.float64, device=dev, requires_grad=True)  # one rate for all
    v, _ = T.american_price(sig, rate, 0.0, torch.tensor(opts.strike, device=dev), 100.0,
                            torch.tensor(opts.maturity, device=dev), 1, _cfg())
    v.sum().backward()
    g = T.american_greeks(sig, 0.03, 0.0, torch.tensor(opts.strike, device=dev), 100.0,
                          torch.tensor(opts.maturity, device=dev), 1, _cfg())
    assert torch.allclose(g["value"], v.detach(), rtol=0, atol=0)
    assert abs(rate.grad.item() - g["rho"].sum().item()) <= 1e-9 * abs(g["rho"].sum().item())


def test_gradient_newton_recovers_hidden_vol():
    print("test_gradient_newton_recovers_hidden_vol: 64 puts, hidden sigma ~ U[0.1, 0.6]", flush=True)
    import torch
    from mango_option import torch_ops as T
    opts = W.c5_puts(64, seed=5)
    hidden = W.hidden_sigmas(64, seed=6)
    truth = W.OptionBatch(opts.spot, opts.strike, opts.maturity, hidden, opts.rate, opts.dividend_yield,
                          opts.option_type)
    target, st, _, _ = O.price_batch(truth, False, _cfg(), use_normalized=False, n_threads=8)
    assert (st == 0).all()
    dev = torch.device("cuda")
    sigma, iters, conv, solves = T.implied_vol_newton(
        torch.tensor(target[:, 0], device=dev), 100.0, torch.tensor(opts.strike, device=dev),
        torch.tensor(opts.maturity, device=dev), torch.tensor(opts.rate, device=dev),
        torch.tensor(opts.dividend_yield, device=dev), 1, _cfg(), tol=1e-11)
    ref_adj, _ = O.adjoint_batch(truth, _cfg(), n_threads=8)
    identifiable = ref_adj[:, 2] > 1e-2  # vega at the hidden sigma (prices at intrinsic carry no information)
    s = sigma.cpu().numpy()
    c = conv.cpu().numpy()
    assert c[identifiable].all()
    assert np.max(np.abs(s[identifiable] - hidden[identifiable])) <= 1e-7
    it = iters.cpu().numpy()
    assert it[identifiable].max() <= 10
    print(f"newton: max {it[identifiable].max()} / mean {it[identifiable].mean():.2f} iterations over "
          f"{identifiable.sum()} identifiable options, {solves} adjoint solves for {len(s)} options", flush=True)
