"""Batched uneven jobs: many independent factor-graph runs of different
universes, lengths and graphs, enqueued together on one GPU.

A job draws its stock count uniformly from [50, 6000], its day count from
[250, 2500] and its graph from a pool of seeded random factor graphs of
20, 30, 45 and 60 factors (catalog modules jobs_g0..jobs_g3, fp32, TS)."""
from typing import List, NamedTuple, Tuple

import numpy as np

JOB_GRAPHS: Tuple[str, ...] = ("jobs_g0", "jobs_g1", "jobs_g2", "jobs_g3")


class JobSpec(NamedTuple):
    graph: str  # catalog module name
    stocks: int
    days: int
    seed: int   # of the job's OHLCV panel


def draw_jobs(n: int = 64, seed: int = 4, stocks=(50, 6000), days=(250, 2500)) -> List[JobSpec]:
    rng = np.random.Generator(np.random.PCG64(seed))
    jobs = []
    for j in range(n):
        g = JOB_GRAPHS[int(rng.integers(len(JOB_GRAPHS)))]
        s = int(rng.integers(stocks[0], stocks[1] + 1))
        t = int(rng.integers(days[0], days[1] + 1))
        jobs.append(JobSpec(g, s, t, seed * 1000 + j))
    return jobs


def ohlcv_device(num_stocks: int, num_time: int, seed: int, device="cuda:0", nan_frac: float = 0.1):
    """OHLCV panel [T][S] float32 generated on the device with torch (the
    recipe of workloads.data.ohlcv, not its numbers): for benchmarks whose
    inputs are too large to generate on the host."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    S, T = num_stocks, num_time
    f64 = dict(dtype=torch.float64, device=device)

    def normal(std):
        return torch.randn((T, S), generator=g, **f64) * std

    close = 100.0 * torch.exp(torch.cumsum(normal(0.02), 0))
    prev = torch.cat([torch.full((1, S), 100.0, **f64), close[:-1]], 0)
    opn = prev * torch.exp(normal(0.005))
    hi = torch.maximum(opn, close) * (1 + normal(0.01).abs())
    lo = torch.minimum(opn, close) * (1 - normal(0.01).abs())
    vol = torch.exp(13.0 + normal(0.5))
    amount = vol * (hi + lo + close) / 3
    res = {}
    for k, a in (("open", opn), ("high", hi), ("low", lo), ("close", close), ("volume", vol), ("amount", amount)):
        a = a.to(torch.float32)
        if nan_frac > 0:
            a[torch.rand((T, S), generator=g, device=device) < nan_frac] = float("nan")
        res[k] = a.contiguous()
    return res
