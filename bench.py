# SPDX-License-Identifier: MIT
"""Benchmark: batched American-option PDE solve on one MI355X.

Workload (the configuration the headline metric is quoted on, 101x498 FP64):
shared-grid option chains, 64 chains x 64 strikes = 4096 options, per chain
S=100, T~U[0.05,2], sigma~U[0.1,0.6], r~U[0,0.08], q~U[0,0.04], strikes
70..130, puts/calls alternating, one 101-point grid per chain (spatial union
of estimate_batch_pde_grid with 101 points) and 498 TR-BDF2 steps. One step =
one call of mango_price_american_chains_device over the 4096 options with
inputs already in HBM.

Also measured: the plain form (one PDE per option, per-option 101-point
multi-sinh grid x 498 steps) on 65,536 options of the plain-configuration
distribution, and the oracle's OpenMP batch on the host cores (CPU baseline).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mango-option_amd"))

PEAK_FP64_TFLOPS = 78.6
N_SPACE, N_TIME = 101, 498


def flops_per_option(n: int, nt: int) -> float:
    """Algorithmic FP64 work of one PDE solve: (n-2)[51(N-1) + 28 + 67]."""
    return (n - 2) * (51 * (nt - 1) + 95)


def chain_grid_config(M):
    from mango_option import _abi
    acc = M.GridAccuracyParams(min_spatial_points=N_SPACE, max_spatial_points=N_SPACE)
    return M.grid_config(_abi.MANGO_GRID_MODE_ACCURACY_FIXED_TIME, acc, n_time=N_TIME)


def count_distinct_pdes(opts, seg):
    n = 0
    for s in range(len(seg) - 1):
        keys = set()
        for i in range(seg[s], seg[s + 1]):
            keys.add((opts.volatility[i], opts.rate[i], opts.dividend_yield[i], opts.maturity[i],
                      int(opts.option_type[i])))
        n += len(keys)
    return n


def run_gpu(args):
    import torch
    import mango_option as M
    from mango_option import _abi, workloads as W

    lib = _abi.load_library()
    if args.variant:
        k, m = (int(v) for v in args.variant.split(","))
        rule = (_abi.MangoVariantRule * 1)(_abi.MangoVariantRule(N_SPACE, k, m))
        if lib.mango_tuning_set(rule, 1, None) != 0:
            raise SystemExit(f"variant {args.variant} is not compiled")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sp = C.c_void_p(stream.cuda_stream)
    err = _abi.MangoError()

    # ---- chain workload ----
    opts, chain = W.c2_chains(64, 64, seed=1)
    n = len(opts)
    seg = np.arange(0, n + 1, 64, dtype=np.int64)
    n_pde = count_distinct_pdes(opts, seg)
    d_opts = torch.from_numpy(opts.records().view(np.uint8)).to(dev)
    d_seg = torch.from_numpy(seg).to(dev)
    d_out = torch.empty((n, 4), dtype=torch.float64, device=dev)
    d_st = torch.empty(n, dtype=torch.int32, device=dev)
    cfg = chain_grid_config(M)

    def chain_step():
        rc = lib.mango_price_american_chains_device(d_opts.data_ptr(), n, d_seg.data_ptr(), len(seg) - 1,
                                                     C.byref(cfg), d_out.data_ptr(), d_st.data_ptr(), sp,
                                                     C.byref(err))
        if rc != 0:
            raise RuntimeError(err.message.decode())

    for _ in range(args.warmup):
        chain_step()
    torch.cuda.synchronize()
    lib.mango_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        chain_step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kms = C.c_double()
    kl = C.c_uint64()
    lib.mango_profile_read(C.byref(kms), C.byref(kl))
    lib.mango_profile_enable(0)
    status = d_st.cpu().numpy()
    greeks = d_out.cpu().numpy()
    if (status != 0).any():
        raise RuntimeError(f"{int((status != 0).sum())} chain options failed")
    chain = {
        "ms_per_step": 1e3 * elapsed / args.steps,
        "options_per_s": n * args.steps / elapsed,
        "kernel_ms": kms.value / max(1, kl.value),
        "launches": int(kl.value),
        "pde_solves": n_pde,
        "greeks": greeks,
    }

    # ---- plain form: one PDE per option, per-option grids ----
    plain = None
    if args.plain_options > 0:
        base = W.c1_puts(min(args.plain_options, 4096), seed=0)
        popts = W.replicate(base, args.plain_options)
        pn = len(popts)
        d_p = torch.from_numpy(popts.records().view(np.uint8)).to(dev)
        d_po = torch.empty((pn, 4), dtype=torch.float64, device=dev)
        d_ps = torch.empty(pn, dtype=torch.int32, device=dev)

        def plain_step():
            rc = lib.mango_price_american_batch_device(d_p.data_ptr(), pn, C.byref(cfg), d_po.data_ptr(),
                                                        d_ps.data_ptr(), sp, C.byref(err))
            if rc != 0:
                raise RuntimeError(err.message.decode())

        plain_step()
        torch.cuda.synchronize()
        lib.mango_profile_enable(1)
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            plain_step()
        torch.cuda.synchronize()
        pel = time.perf_counter() - t0
        lib.mango_profile_read(C.byref(kms), C.byref(kl))
        lib.mango_profile_enable(0)
        ps = d_ps.cpu().numpy()
        if (ps != 0).any():
            raise RuntimeError(f"{int((ps != 0).sum())} plain options failed")
        plain = {"options": pn, "ms_per_step": 1e3 * pel / args.plain_steps,
                 "options_per_s": pn * args.plain_steps / pel, "kernel_ms": kms.value / max(1, kl.value),
                 "greeks_head": d_po[:64].cpu().numpy()}
    return opts, seg, chain, plain


def parity_check(opts, seg, greeks, n_chains=2):
    """First chains against the oracle with the reference's routing."""
    import mango_option as M
    from mango_option import _abi
    import pyoracle as O
    cfg_acc = M.GridAccuracyParams(min_spatial_points=N_SPACE, max_spatial_points=N_SPACE).to_c()
    worst = 0.0
    for c in range(n_chains):
        idx = np.arange(seg[c], seg[c + 1])
        sub = opts.take(idx)
        arr, n = O.c_params(sub)
        spec, _ = O.estimate_batch_pde_grid(arr, n, cfg_acc)
        cfg = M.grid_config(_abi.MANGO_GRID_MODE_EXPLICIT, spec=M.GridSpec.from_c(spec), n_time=N_TIME)
        for typ in (0, 1):
            sel = np.where(sub.option_type == typ)[0]
            g, s, _, _ = O.price_batch(sub.take(sel), True, cfg, n_threads=os.cpu_count() or 1)
            worst = max(worst, float(np.max(np.abs(g[:, 0] - greeks[idx[sel], 0]))))
    return worst


def cpu_baseline(opts, seg, seconds: float):
    """The oracle (CPU restatement of the reference) pricing whole chains the
    reference's way: one solve_batch(group, use_shared_grid=true, grid) per
    (chain, type) with the chain's union grid; OpenMP across the regular
    route's options. Chains are timed until `seconds` elapse."""
    import mango_option as M
    from mango_option import _abi
    import pyoracle as O
    threads = min(os.cpu_count() or 1, 16)
    acc = M.GridAccuracyParams(min_spatial_points=N_SPACE, max_spatial_points=N_SPACE).to_c()
    done = 0
    solves = 0
    t0 = time.perf_counter()
    c = 0
    while c < len(seg) - 1:
        idx = np.arange(seg[c], seg[c + 1])
        sub = opts.take(idx)
        arr, n = O.c_params(sub)
        spec, _ = O.estimate_batch_pde_grid(arr, n, acc)
        cfg = M.grid_config(_abi.MANGO_GRID_MODE_EXPLICIT, spec=M.GridSpec.from_c(spec), n_time=N_TIME)
        for typ in (0, 1):
            sel = np.where(sub.option_type == typ)[0]
            _, _, _, s = O.price_batch(sub.take(sel), True, cfg, n_threads=threads)
            solves += s
        done += len(idx)
        c += 1
        if time.perf_counter() - t0 > seconds:
            break
    el = time.perf_counter() - t0
    return {"value": done / el, "unit": "options/s", "cores": threads, "kind": "port",
            "sample": f"{c} of {len(seg) - 1} chains ({done} options, {solves} PDE solves by the reference's "
                      f"routing) in {el:.1f} s, oracle OpenMP batch, {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-options", type=int, default=65536)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variant", default="", help="force lanes,nodes_per_lane of the solve kernel")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(options are independent; sharding is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))

    opts, seg, chain, plain = run_gpu(args)
    parity = parity_check(opts, seg, chain["greeks"])

    flops_launch = chain["pde_solves"] * flops_per_option(N_SPACE, N_TIME)
    achieved = flops_launch / (chain["kernel_ms"] * 1e-3) / 1e12
    line = {
        "metric": "American options/sec @101x498 FP64 (plain vs batched); adjoint Greeks/sec",
        "value": chain["options_per_s"],
        "unit": "options/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": chain["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded xoshiro256**), inputs resident in HBM",
        "config": {"workload": "shared-grid chains: 64 chains x 64 strikes = 4096 options, 101x498, "
                               "one grid per chain, one PDE per (chain, type)",
                   "options": int(len(opts)), "pde_solves": chain["pde_solves"], "n_space": N_SPACE,
                   "n_time": N_TIME},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "american_solve_kernel (FP64 vector ALU; no MFMA)",
                     "kernel_ms": chain["kernel_ms"], "flops_per_launch": flops_launch},
        "parity_max_abs_price_err_vs_oracle": parity,
    }
    if plain is not None:
        pf = plain["options"] * flops_per_option(N_SPACE, N_TIME)
        pach = pf / (plain["kernel_ms"] * 1e-3) / 1e12
        line["plain"] = {"workload": f"{plain['options']} puts of the plain configuration distribution, "
                                     f"per-option 101-point grids x 498 steps, one PDE per option",
                         "options_per_s": plain["options_per_s"], "ms_per_step": plain["ms_per_step"],
                         "kernel_ms": plain["kernel_ms"],
                         "roofline": {"achieved": pach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                      "frac": pach / PEAK_FP64_TFLOPS}}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(opts, seg, args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
