# SPDX-License-Identifier: MIT
"""Seeded synthetic workloads of the benchmark configurations.

PRNG: xoshiro256** seeded through splitmix64, uniform draws
u = (next() >> 11) * 2**-53, draws taken in a fixed order per option. The
same generator is used by the tests, the benchmark and the CPU baseline so
that every leg sees identical inputs.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_M64 = (1 << 64) - 1


def _rotl(x: int, k: int) -> int:
    return ((x << k) | (x >> (64 - k))) & _M64


class Xoshiro256ss:
    def __init__(self, seed: int):
        s = seed & _M64
        st = []
        for _ in range(4):  # splitmix64
            s = (s + 0x9E3779B97F4A7C15) & _M64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            st.append(z ^ (z >> 31))
        self.s = st

    def next(self) -> int:
        s = self.s
        result = (_rotl((s[1] * 5) & _M64, 7) * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


PUT = 1
CALL = 0

# numpy structured layout of MangoDeviceOption (56 bytes)
DEVICE_OPTION_DTYPE = np.dtype([("spot", "<f8"), ("strike", "<f8"), ("maturity", "<f8"), ("volatility", "<f8"),
                                ("rate", "<f8"), ("dividend_yield", "<f8"), ("option_type", "<i4"),
                                ("reserved", "<i4")])


@dataclass
class OptionBatch:
    """Struct-of-arrays option batch (constant rate, continuous yield)."""
    spot: np.ndarray
    strike: np.ndarray
    maturity: np.ndarray
    volatility: np.ndarray
    rate: np.ndarray
    dividend_yield: np.ndarray
    option_type: np.ndarray  # int32, 0 call / 1 put

    def __len__(self) -> int:
        return len(self.spot)

    def records(self) -> np.ndarray:
        rec = np.zeros(len(self), dtype=DEVICE_OPTION_DTYPE)
        for f in ("spot", "strike", "maturity", "volatility", "rate", "dividend_yield", "option_type"):
            rec[f] = getattr(self, f)
        return rec

    def take(self, idx) -> "OptionBatch":
        return OptionBatch(*(getattr(self, f)[idx] for f in
                             ("spot", "strike", "maturity", "volatility", "rate", "dividend_yield", "option_type")))


def c1_puts(n: int = 64, seed: int = 0) -> OptionBatch:
    """Plain configuration: American puts, S=100, K~U[80,120], T~U[0.1,2],
    sigma~U[0.1,0.5], r~U[0,0.08], q~U[0,0.04] (draw order K,T,sigma,r,q)."""
    g = Xoshiro256ss(seed)
    cols = np.empty((n, 5))
    for i in range(n):
        cols[i, 0] = 80.0 + 40.0 * g.uniform()
        cols[i, 1] = 0.1 + 1.9 * g.uniform()
        cols[i, 2] = 0.10 + 0.40 * g.uniform()
        cols[i, 3] = 0.08 * g.uniform()
        cols[i, 4] = 0.04 * g.uniform()
    return OptionBatch(np.full(n, 100.0), cols[:, 0].copy(), cols[:, 1].copy(), cols[:, 2].copy(),
                       cols[:, 3].copy(), cols[:, 4].copy(), np.full(n, PUT, dtype=np.int32))


def c2_chains(n_chains: int = 64, n_strikes: int = 64, seed: int = 1):
    """Shared-grid chains: per chain S=100, T~U[0.05,2], sigma~U[0.1,0.6],
    r~U[0,0.08], q~U[0,0.04]; strikes K_j = 70 + 60 j/(n_strikes-1), puts
    for even j and calls for odd j. Returns (batch, chain_id)."""
    g = Xoshiro256ss(seed)
    rows = []
    chain_id = []
    for c in range(n_chains):
        T = 0.05 + 1.95 * g.uniform()
        sig = 0.1 + 0.5 * g.uniform()
        r = 0.08 * g.uniform()
        q = 0.04 * g.uniform()
        for j in range(n_strikes):
            K = 70.0 + 60.0 * j / (n_strikes - 1)
            rows.append((100.0, K, T, sig, r, q, PUT if j % 2 == 0 else CALL))
            chain_id.append(c)
    a = np.array(rows)
    return (OptionBatch(a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy(), a[:, 3].copy(), a[:, 4].copy(),
                        a[:, 5].copy(), a[:, 6].astype(np.int32)), np.array(chain_id, dtype=np.int32))


def c3_uneven(n: int = 65536, seed: int = 2) -> OptionBatch:
    """Uneven batch: T = exp(U[ln(1/365), ln 3]), K/S~U[0.6,1.4] (S=100),
    sigma~U[0.05,0.8], r~U[0,0.1], q~U[0,0.05]; types alternate put/call."""
    g = Xoshiro256ss(seed)
    a = np.empty((n, 5))
    lo, hi = math.log(1.0 / 365.0), math.log(3.0)
    for i in range(n):
        a[i, 0] = math.exp(lo + (hi - lo) * g.uniform())
        a[i, 1] = 100.0 * (0.6 + 0.8 * g.uniform())
        a[i, 2] = 0.05 + 0.75 * g.uniform()
        a[i, 3] = 0.1 * g.uniform()
        a[i, 4] = 0.05 * g.uniform()
    types = np.where(np.arange(n) % 2 == 0, PUT, CALL).astype(np.int32)
    return OptionBatch(np.full(n, 100.0), a[:, 1].copy(), a[:, 0].copy(), a[:, 2].copy(), a[:, 3].copy(),
                       a[:, 4].copy(), types)


def replicate(batch: OptionBatch, n: int) -> OptionBatch:
    """Tile a batch to n options (throughput variants of a configuration)."""
    idx = np.arange(n) % len(batch)
    return batch.take(idx)


def c5_puts(n: int = 16384, seed: int = 5) -> OptionBatch:
    """Adjoint configuration: American puts drawn like the plain configuration."""
    return c1_puts(n, seed=seed)


def hidden_sigmas(n: int, seed: int = 6) -> np.ndarray:
    """Hidden volatilities sigma* ~ U[0.1, 0.6] for the calibration targets."""
    g = Xoshiro256ss(seed)
    return np.array([0.1 + 0.5 * g.uniform() for _ in range(n)])


# ---- price-table configuration ----

C4_K_REF = 100.0
C4_DIVIDEND = 0.02


def sqrt_uniform_grid(lo: float, hi: float, n: int) -> np.ndarray:
    """n points uniform in sqrt-space on [lo, hi] (denser near lo), as the
    reference's maturity-axis estimator lays them out
    (src/option/table/bspline/bspline_builder.hpp:219-228)."""
    s = math.sqrt(lo) + (math.sqrt(hi) - math.sqrt(lo)) * np.arange(n, dtype=np.float64) / (n - 1)
    return np.array([v * v for v in s])


def c4_axes():
    """Price-table axes: 50 log-moneyness points on [-0.6, 0.4], 30 maturities
    sqrt-uniform on [0.01, 2], 20 volatilities on [0.05, 0.8], 10 rates on
    [0, 0.1] (put, K_ref = 100, dividend yield 0.02)."""
    return (np.linspace(-0.6, 0.4, 50), sqrt_uniform_grid(0.01, 2.0, 30), np.linspace(0.05, 0.80, 20),
            np.linspace(0.0, 0.10, 10))


def splitmix64_uniform(seed: int, n: int, start: int = 0) -> np.ndarray:
    """Counter-based uniforms: u_k = (splitmix64 output k >> 11) * 2**-53 for
    k = start .. start+n-1 of the stream seeded with `seed` (vectorised; used
    where millions of draws are needed)."""
    k = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def c4_query_points(n: int = 1 << 20, seed: int = 4):
    """Interpolated-IV queries inside the table axes: columns (log-moneyness,
    maturity, sigma, rate) uniform on each axis range and a price noise factor
    1 + U[-0.01, 0.01]; draw k of query i is stream element 5 i + k."""
    lm, tau, sig, rate = c4_axes()
    u = splitmix64_uniform(seed, 5 * n).reshape(n, 5)
    pts = np.empty((n, 4))
    for a, g in enumerate((lm, tau, sig, rate)):
        pts[:, a] = g[0] + (g[-1] - g[0]) * u[:, a]
    return pts, 1.0 + 0.02 * (u[:, 4] - 0.5)
