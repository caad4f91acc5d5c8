"""Chunking autotuner: times batch runs of one (module, stocks, rows) under a
few row-chunking policies and keeps the fastest on the executor
(Executor.set_chunking_for). Results never depend on the policy; only the
launch geometry (rows per chunk, workgroups per kernel) changes."""
import time
from typing import Dict, Iterable, Optional, Tuple

from .runner import KunRunner as kr

# (target workgroups per kernel, chunk >= halo_factor x halo, chunk >= min_rows)
DEFAULT_CANDIDATES: Tuple[Tuple[int, int, int], ...] = (
    (4096, 3, 64), (16384, 3, 64), (32768, 3, 64), (65536, 3, 32), (32768, 2, 32), (131072, 2, 16))


def tune(executor: "kr.Executor", mod: "kr.Module", buffers: Dict[str, int], num_stocks: int, total_time: int,
         cur_time: int = 0, length: Optional[int] = None, candidates: Iterable[Tuple[int, int, int]] = None,
         reps: int = 3):
    """Try each candidate on device-resident `buffers` (name -> device
    pointer); returns (best policy, {policy: ms per run})."""
    length = total_time - cur_time if length is None else length
    times = {}
    for pol in (candidates or DEFAULT_CANDIDATES):
        executor.set_chunking_for(mod, num_stocks, length, *pol)
        kr.runGraphDevice(executor, mod, buffers, num_stocks, total_time, cur_time, length)  # warm-up
        executor.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            kr.runGraphDevice(executor, mod, buffers, num_stocks, total_time, cur_time, length)
        executor.synchronize()
        times[pol] = (time.perf_counter() - t0) / reps * 1e3
    best = min(times, key=times.get)
    executor.set_chunking_for(mod, num_stocks, length, *best)
    return best, times
