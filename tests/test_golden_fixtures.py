# SPDX-License-Identifier: MIT
"""The committed golden fixtures (tests/golden/) against the oracle (CPU)
and against the GPU library (gpu marker)."""
import json
import os

import numpy as np
import pytest

import pyoracle as O
import mango_option as M
from mango_option import _abi, workloads as W

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _c1():
    d = json.load(open(os.path.join(HERE, "c1_64_puts.json")))
    opts = W.OptionBatch(*(np.array(d["inputs"][f]) for f in
                           ("spot", "strike", "maturity", "volatility", "rate", "dividend_yield")),
                         np.array(d["option_type"], dtype=np.int32))
    want = np.stack([d["price"], d["delta"], d["gamma"], d["theta"]], axis=1)
    acc = M.GridAccuracyParams(min_spatial_points=101, max_spatial_points=101)
    cfg = M.grid_config(_abi.MANGO_GRID_MODE_ACCURACY_FIXED_TIME, acc, n_time=d["n_time"])
    return d, opts, want, cfg


def test_fixture_inputs_are_the_seeded_workload():
    d, opts, _, _ = _c1()
    ref = W.c1_puts(64, seed=0)
    assert np.array_equal(opts.strike, ref.strike) and np.array_equal(opts.volatility, ref.volatility)


def test_oracle_reproduces_c1_fixture():
    _, opts, want, cfg = _c1()
    g, s, failed, _ = O.price_batch(opts, False, cfg, use_normalized=False)
    assert failed == 0
    np.testing.assert_allclose(g, want, rtol=1e-13, atol=1e-14)


def test_oracle_reproduces_thomas_fixture():
    d = json.load(open(os.path.join(HERE, "thomas_small.json")))
    for c in d["cases"]:
        ok, x = O.thomas(c["lower"], c["diag"], c["upper"], c["rhs"])
        okp, xp = O.thomas_projected(c["lower"], c["diag"], c["upper"], c["rhs"], c["psi"])
        np.testing.assert_array_equal(x, c["solution"])
        np.testing.assert_array_equal(xp, c["projected"])
        n = len(c["diag"])
        A = np.diag(c["diag"]) + np.diag(c["upper"], 1) + np.diag(c["lower"], -1)
        np.testing.assert_allclose(A @ x, c["rhs"], atol=1e-12)


@pytest.mark.gpu
def test_gpu_matches_c1_fixture():
    _, opts, want, cfg = _c1()
    g, s, failed = M.price_batch(opts, False, cfg, use_normalized=False)
    assert failed == 0
    assert np.max(np.abs(g[:, 0] - want[:, 0])) <= 1e-9
    np.testing.assert_allclose(g[:, 1:], want[:, 1:], rtol=1e-7, atol=1e-10)


def _c4():
    d = np.load(os.path.join(HERE, "c4_slice.npz"), allow_pickle=False)
    return {k: d[k] for k in d.files}


def test_oracle_reproduces_c4_slice_fixture():
    d = _c4()
    ax = W.c4_axes()
    si, ri = int(d["sigma_index"]), int(d["rate_index"])
    spec = M.GridSpec(int(d["grid_type"]), float(d["x_min"]), float(d["x_max"]), int(d["n_points"]),
                      clusters=[tuple(c) for c in d["clusters"]]).to_c()
    if int(d["grid_type"]) == _abi.MANGO_GRID_SINH:
        spec.concentration = float(d["clusters"][0][1])
    np.testing.assert_array_equal(O.generate(spec), d["x"])
    p = O.one_param(1.0, 1.0, ax[1][-1], ax[2][si], ax[3][ri], W.C4_DIVIDEND)
    st, snaps, times = O.solve_snapshots(p, spec, int(d["n_time"]), ax[1])
    assert st == 0
    np.testing.assert_array_equal(times, d["snapshot_times"])
    np.testing.assert_allclose(snaps, d["snapshots"], rtol=0, atol=1e-14)
    # the tensor columns: snapshot j of the deduplicated list, the last one repeated
    for j in range(len(ax[1])):
        k = min(j, len(snaps) - 1)
        ok, v, _ = O.spline_eval(d["x"], snaps[k], ax[0])
        assert ok
        np.testing.assert_allclose(v, d["prices"][:, j], rtol=0, atol=1e-14)


@pytest.mark.gpu
def test_gpu_table_matches_c4_slice_fixture():
    d = _c4()
    ax = W.c4_axes()
    cfg = M.grid_config(_abi.MANGO_GRID_MODE_EXPLICIT, spec=M.GridSpec(n_points=201), n_time=1000)
    t = M.PriceTableBuilder.from_vectors(*ax, W.C4_K_REF, cfg, M.OptionType.PUT, W.C4_DIVIDEND).build()
    _, _, prices = t.export()
    si, ri = int(d["sigma_index"]), int(d["rate_index"])
    assert np.max(np.abs(prices[:, :, si, ri] - d["prices"])) <= 1e-11
