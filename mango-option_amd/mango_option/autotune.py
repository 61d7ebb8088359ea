# SPDX-License-Identifier: MIT
"""Autotuner for the solve kernel's launch shape.

The tunable is the variant (lanes per option K, grid nodes per lane M): K
lanes of one wavefront cooperate on one option (K = 64 is one option per
wavefront, K = 16 four options per wavefront); M sets the register footprint.
For each grid-size bin the tuner times every compiled variant that covers the
bin on a device-resident batch and keeps the fastest. Choices are written as
"max_n lanes nodes_per_lane" lines to tuning_gfx950.txt next to the library,
which the library reads at its first device call.

    python -m mango_option.autotune [--bins 51,101,201,401] [--options 16384]
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import time

import numpy as np

from . import _abi, workloads as W

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_gfx950.txt")


def compiled_variants():
    lib = _abi.load_library()
    buf = (_abi.MangoVariantRule * 64)()
    n = lib.mango_tuning_variants(buf, 64)
    return [(int(buf[i].lanes), int(buf[i].nodes_per_lane)) for i in range(min(n, 64))]


def _set_rules(rules):
    lib = _abi.load_library()
    arr = (_abi.MangoVariantRule * max(1, len(rules)))(*[_abi.MangoVariantRule(m, k, nd) for m, k, nd in rules])
    err = _abi.MangoError()
    if lib.mango_tuning_set(arr, len(rules), C.byref(err)) != 0:
        raise RuntimeError(err.message.decode())


def time_variant(n_space: int, variant, n_options: int, n_time: int, repeats: int) -> float:
    """Average solve-kernel milliseconds for one launch of n_options puts of
    the plain distribution at n_space x n_time with the given variant."""
    import torch
    import mango_option as M
    lib = _abi.load_library()
    _set_rules([(n_space, variant[0], variant[1])])
    opts = W.replicate(W.c1_puts(min(n_options, 2048), seed=0), n_options)
    dev = torch.device("cuda:0")
    d_opts = torch.from_numpy(opts.records().view(np.uint8)).to(dev)
    d_out = torch.empty((n_options, 4), dtype=torch.float64, device=dev)
    d_st = torch.empty(n_options, dtype=torch.int32, device=dev)
    acc = M.GridAccuracyParams(min_spatial_points=n_space, max_spatial_points=n_space)
    cfg = M.grid_config(_abi.MANGO_GRID_MODE_ACCURACY_FIXED_TIME, acc, n_time=n_time)
    sp = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    err = _abi.MangoError()

    def run():
        if lib.mango_price_american_batch_device(d_opts.data_ptr(), n_options, C.byref(cfg), d_out.data_ptr(),
                                                 d_st.data_ptr(), sp, C.byref(err)) != 0:
            raise RuntimeError(err.message.decode())

    run()
    torch.cuda.synchronize()
    if (d_st != 0).any():
        raise RuntimeError("solve failed during tuning")
    lib.mango_profile_enable(1)
    for _ in range(repeats):
        run()
    torch.cuda.synchronize()
    ms = C.c_double()
    cnt = C.c_uint64()
    lib.mango_profile_read(C.byref(ms), C.byref(cnt))
    lib.mango_profile_enable(0)
    return ms.value / max(1, cnt.value)


def tune(bins=(51, 101, 201, 401), n_options=16384, n_time=200, repeats=2, out_path=DEFAULT_OUT, verbose=True):
    variants = compiled_variants()
    rules = []
    report = []
    for n in bins:
        cands = [v for v in variants if v[0] * v[1] >= n]
        best = None
        for v in cands:
            ms = time_variant(n, v, n_options, n_time, repeats)
            report.append((n, v, ms))
            if verbose:
                print(f"n={n:4d} lanes={v[0]:2d} nodes={v[1]:2d}: {ms:8.3f} ms", flush=True)
            if best is None or ms < best[1]:
                best = (v, ms)
        rules.append((n, best[0][0], best[0][1]))
    _set_rules(rules)
    if out_path:
        with open(out_path, "w") as f:
            f.write("# max_n lanes nodes_per_lane  (written by mango_option.autotune on gfx950)\n")
            for m, k, nd in rules:
                f.write(f"{m} {k} {nd}\n")
    return rules, report


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--bins", default="51,101,201,401")
    ap.add_argument("--options", type=int, default=16384)
    ap.add_argument("--n-time", type=int, default=200)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args()
    t0 = time.time()
    rules, _ = tune(tuple(int(x) for x in a.bins.split(",")), a.options, a.n_time, out_path=a.out)
    print("rules:", rules, f"({time.time() - t0:.1f} s)")


if __name__ == "__main__":
    main()
