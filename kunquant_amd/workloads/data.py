"""Synthetic OHLCV panels in TS layout [T, S] (seeded, numpy PCG64).

close: geometric random walk, daily log-returns ~ N(0, 0.02), close0 = 100
open:  previous close * exp(N(0, 0.005))
high/low: max/min(open, close) * (1 +- |N(0, 0.01)|)
volume ~ lognormal(13, 0.5); amount = volume * (high + low + close) / 3
Optionally a fraction of all cells is set to NaN (i.i.d.)."""
from typing import Dict

import numpy as np

FIELDS = ("open", "high", "low", "close", "volume", "amount")


def ohlcv(num_stocks: int, num_time: int, seed: int, dtype=np.float32, nan_frac: float = 0.0,
          price_scale=None, volume_scale=None) -> Dict[str, np.ndarray]:
    rng = np.random.Generator(np.random.PCG64(seed))
    S, T = num_stocks, num_time
    close0 = np.full(S, 100.0)
    if price_scale is not None:
        lo, hi = price_scale
        close0 = np.exp(rng.uniform(np.log(lo), np.log(hi), S))
    ret = rng.normal(0.0, 0.02, (T, S))
    close = close0 * np.exp(np.cumsum(ret, axis=0))
    prev = np.vstack([close0[None, :], close[:-1]])
    opn = prev * np.exp(rng.normal(0.0, 0.005, (T, S)))
    hi = np.maximum(opn, close) * (1 + np.abs(rng.normal(0.0, 0.01, (T, S))))
    lo = np.minimum(opn, close) * (1 - np.abs(rng.normal(0.0, 0.01, (T, S))))
    vol = rng.lognormal(13.0, 0.5, (T, S))
    if volume_scale is not None:
        a, b = volume_scale
        vol = vol / np.exp(13.0) * np.exp(rng.uniform(np.log(a), np.log(b), S))
    amount = vol * (hi + lo + close) / 3
    out = {"open": opn, "high": hi, "low": lo, "close": close, "volume": vol, "amount": amount}
    res = {}
    for k in FIELDS:
        a = out[k].astype(dtype)
        if nan_frac > 0:
            a[rng.random((T, S)) < nan_frac] = np.nan
        res[k] = np.ascontiguousarray(a)
    return res
