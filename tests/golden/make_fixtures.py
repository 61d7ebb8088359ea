# SPDX-License-Identifier: MIT
"""Writes the golden fixtures of tests/golden/ from the oracle (the CPU
restatement of the reference; the reference itself cannot be built here).

c1_64_puts.json : plain configuration, 64 puts seed 0, per-option 101-point
                  grid from estimate_pde_grid(min=max=101) x 498 steps:
                  inputs, grid specs and price/delta/gamma/theta.
thomas_small.json: small tridiagonal and projected systems with solutions.
c4_slice.npz    : price-table configuration, the (sigma = 0.8, r = 0.1) solve:
                  its estimated grid, step count, recorded snapshot steps and
                  rows (29 distinct steps for 30 maturities) and the 50 x 30
                  extracted prices V/K_ref of the table tensor.
Run: python tests/golden/make_fixtures.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "mango-option_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

import pyoracle as O  # noqa: E402
import mango_option as M  # noqa: E402
from mango_option import _abi, workloads as W  # noqa: E402


def c1():
    opts = W.c1_puts(64, seed=0)
    acc = M.GridAccuracyParams(min_spatial_points=101, max_spatial_points=101)
    cfg = M.grid_config(_abi.MANGO_GRID_MODE_ACCURACY_FIXED_TIME, acc, n_time=498)
    g, s, failed, _ = O.price_batch(opts, False, cfg, use_normalized=False)
    assert failed == 0
    arr, n = O.c_params(opts)
    specs = []
    for i in range(n):
        spec, _ = O.estimate_pde_grid(arr[i], acc.to_c())
        sp = M.GridSpec.from_c(spec)
        specs.append({"type": sp.type, "x_min": sp.x_min, "x_max": sp.x_max, "n_points": sp.n_points,
                      "clusters": sp.clusters})
    return {
        "description": "64 American puts, seed 0, 101x498 (per-option multi-sinh grid), oracle outputs",
        "n_time": 498,
        "inputs": {f: getattr(opts, f).tolist() for f in
                   ("spot", "strike", "maturity", "volatility", "rate", "dividend_yield")},
        "option_type": opts.option_type.tolist(),
        "grids": specs,
        "price": g[:, 0].tolist(), "delta": g[:, 1].tolist(), "gamma": g[:, 2].tolist(), "theta": g[:, 3].tolist(),
    }


def thomas_small():
    cases = []
    rng = np.random.default_rng(123)
    for n in (1, 2, 3, 6, 11):
        lo, up = rng.uniform(-1, 0, max(n - 1, 0)), rng.uniform(-1, 0, max(n - 1, 0))
        dg = 2.5 + rng.uniform(0, 1, n)
        rhs = rng.normal(size=n)
        psi = rng.normal(size=n) - 0.5
        ok, x = O.thomas(lo, dg, up, rhs)
        okp, xp = O.thomas_projected(lo, dg, up, rhs, psi)
        cases.append({"lower": lo.tolist(), "diag": dg.tolist(), "upper": up.tolist(), "rhs": rhs.tolist(),
                      "psi": psi.tolist(), "solution": x.tolist(), "projected": xp.tolist()})
    return {"description": "small diagonally dominant systems: Thomas and projected Thomas solutions",
            "cases": cases}


def c4_slice():
    ax = W.c4_axes()
    si, ri = 19, 9
    spec = M.GridSpec(n_points=201).to_c()
    t = O.table_build(ax, W.C4_K_REF, W.C4_DIVIDEND, True, spec=spec, n_time=1000, n_threads=8)
    # the builder's fallback: per-solve estimated grids, 201 points, <= 1000 steps
    max_sst = ax[2][-1] * np.sqrt(ax[1][-1])
    acc = M.GridAccuracyParams(min_spatial_points=201, max_spatial_points=201, max_time_steps=1000)
    acc.n_sigma = max(5.0, 3.0 / max_sst * 1.1, 0.6 / max_sst * 1.1)
    p = O.one_param(1.0, 1.0, ax[1][-1], ax[2][si], ax[3][ri], W.C4_DIVIDEND)
    g, nt = O.estimate_batch_pde_grid((_abi.MangoPricingParams * 1)(p), 1, acc.to_c())
    st, snaps, times = O.solve_snapshots(p, g, nt, ax[1])
    assert st == 0
    gs = M.GridSpec.from_c(g)
    return {"sigma_index": si, "rate_index": ri, "grid_type": gs.type, "x_min": gs.x_min, "x_max": gs.x_max,
            "n_points": gs.n_points, "clusters": np.array(gs.clusters, dtype=np.float64).reshape(-1, 3),
            "n_time": nt, "x": O.generate(g), "snapshot_times": times, "snapshots": snaps,
            "prices": t.tensor[:, :, si, ri].copy()}


if __name__ == "__main__":
    for name, fn in (("c1_64_puts.json", c1), ("thomas_small.json", thomas_small)):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(fn(), f, indent=1)
        print("wrote", name)
    np.savez_compressed(os.path.join(HERE, "c4_slice.npz"), **c4_slice())
    print("wrote c4_slice.npz")
