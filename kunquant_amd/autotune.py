"""Chunking autotuner: times batch runs of one (module, stocks, rows) under a
few row-chunking policies and keeps the fastest on the executor
(Executor.set_chunking_for). Results never depend on the policy; only the
launch geometry (rows per chunk, workgroups per kernel) changes.

Picks can persist in a JSON file (`cache=` or the KUN_TUNE_CACHE
environment variable), keyed by the sha256 of the module's library, the
module name and the run shape, so a later process applies them without
timing anything."""
import hashlib
import json
import os
import time
from typing import Dict, Iterable, Optional, Tuple

from .runner import KunRunner as kr

_SHA: Dict[Tuple[str, float, int], str] = {}


def _lib_sha(path: str) -> str:
    st = os.stat(path)
    key = (os.path.abspath(path), st.st_mtime, st.st_size)
    if key not in _SHA:
        h = hashlib.sha256()
        with open(path, "rb") as f:
            for chunk in iter(lambda: f.read(1 << 20), b""):
                h.update(chunk)
        _SHA[key] = h.hexdigest()
    return _SHA[key]


def cache_key(mod: "kr.Module", num_stocks: int, length: int, cur_time: int) -> str:
    return f"{_lib_sha(mod._lib.path)[:16]}/{mod.name}/{num_stocks}/{length}/{cur_time}"


def _read(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}

# (target workgroups per kernel, chunk >= halo_factor x halo, chunk >= min_rows)
DEFAULT_CANDIDATES: Tuple[Tuple[int, int, int], ...] = (
    (4096, 3, 64), (16384, 3, 64), (32768, 3, 64), (65536, 3, 32), (32768, 2, 32), (131072, 2, 16))


def tune(executor: "kr.Executor", mod: "kr.Module", buffers: Dict[str, int], num_stocks: int, total_time: int,
         cur_time: int = 0, length: Optional[int] = None, candidates: Iterable[Tuple[int, int, int]] = None,
         reps: int = 3, cache: Optional[str] = None):
    """Try each candidate on device-resident `buffers` (name -> device
    pointer); returns (best policy, {policy: ms per run}). With a cache file
    holding a pick for this library, module and shape, that pick is applied
    and returned with no timings."""
    length = total_time - cur_time if length is None else length
    path = cache if cache is not None else os.environ.get("KUN_TUNE_CACHE")
    key = cache_key(mod, num_stocks, length, cur_time) if path and mod.name else None
    if key:
        hit = _read(path).get(key)
        if hit:
            pol = tuple(int(x) for x in hit)
            executor.set_chunking_for(mod, num_stocks, length, *pol)
            return pol, {}
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
    if key:
        rec = _read(path)
        rec[key] = list(best)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(rec, f, indent=0, sort_keys=True)
        os.replace(tmp, path)
    return best, times
