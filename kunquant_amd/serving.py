"""Continuous-batching request loop for stream-mode serving.

Sessions of one STREAM module live in one StreamPool (rows and carried state
resident in HBM). A session opened while the pool has no contiguous room for
its lanes waits and is admitted when a release frees enough (admission by
state footprint). Ticks queue per session; every dispatch ticks each session
that has a queued tick and no tick in flight, in one pool cycle, while up to
`max_inflight` cycles run on the device: ticks that arrive while a cycle runs
are picked up by the next one without waiting for it. Ticks of one session
run in arrival order, one per cycle.

Under load a cycle takes at most `cycle_share` of the admitted sessions
(default: 1 / max_inflight of them, in order of readiness), so that cycles
overlap on the device, and at most `max_computing` of the cycles in flight
may still be running their phases when another starts (default: all of
them): with results copied to host inside the cycle, a limit below
max_inflight keeps the copies of some cycles running while others compute,
instead of cycles that start together computing together and copying
together."""
import math
import time
from collections import deque
from typing import Callable, Deque, Dict, Hashable, List, Optional, Tuple

from .runner import KunRunner as kr


class TickServer:
    def __init__(self, pool: "kr.StreamPool", max_inflight: int = 2, clock: Callable[[], float] = time.perf_counter,
                 on_done: Optional[Callable[[Hashable, int, float], None]] = None,
                 cycle_share: Optional[float] = None, max_computing: Optional[int] = None):
        self.pool = pool
        self.max_inflight = max_inflight
        self.cycle_share = (1.0 / max_inflight) if cycle_share is None else cycle_share
        self.max_computing = max_inflight if max_computing is None else max_computing
        self.clock = clock
        self.on_done = on_done      # (key, session id, arrival time) once a tick's results are in HBM
        self.waiting: Deque[Tuple[Hashable, int]] = deque()
        self.sid: Dict[Hashable, int] = {}
        self.queue: Dict[Hashable, Deque] = {}
        self.ready: Dict[Hashable, None] = {}   # admitted, tick queued, none in flight (insertion order)
        self.busy = set()
        self.inflight: Deque[Tuple[int, List[Tuple[Hashable, float]]]] = deque()
        self.latencies: List[float] = []
        self.cycles = 0
        self.ticks_in_cycles = 0
        self.submit_s = 0.0         # host time inside pool.submit

    # ---- sessions -------------------------------------------------------
    def open(self, key: Hashable, num_stocks: int) -> bool:
        """Queue a session for admission; True if admitted now."""
        self.queue.setdefault(key, deque())
        self.waiting.append((key, num_stocks))
        self._admit()
        return key in self.sid

    def _admit(self):
        while self.waiting:
            key, S = self.waiting[0]
            sid = self.pool.admit(S)
            if sid is None:
                return
            self.waiting.popleft()
            self.sid[key] = sid
            if self.queue[key]:
                self.ready[key] = None

    def close(self, key: Hashable):
        """Release a session once its queued ticks are done (drain first)."""
        while self.queue.get(key) or key in self.busy:
            self.step()
        sid = self.sid.pop(key)
        self.ready.pop(key, None)
        self.queue.pop(key, None)
        self.pool.release(sid)
        self._admit()

    # ---- ticks ----------------------------------------------------------
    def push(self, key: Hashable, rows, t_arrival: Optional[float] = None):
        """rows: {input name: row [S]} or StreamPool.Rows of the session;
        kept by reference until dispatched."""
        q = self.queue[key]
        q.append((self.clock() if t_arrival is None else t_arrival, rows))
        if key in self.sid and key not in self.busy:
            self.ready[key] = None

    def poll(self) -> int:
        """Retire finished cycles (oldest first); returns ticks completed."""
        n = 0
        while self.inflight and self.pool.done(self.inflight[0][0]):
            _, members = self.inflight.popleft()
            now = self.clock()
            for key, t0 in members:
                self.latencies.append(now - t0)
                self.busy.discard(key)
                if self.queue[key]:
                    self.ready[key] = None
                if self.on_done is not None:
                    self.on_done(key, self.sid[key], t0)
            n += len(members)
        return n

    def dispatch(self) -> bool:
        """Submit one cycle of every ready session, if a cycle slot is free."""
        if not self.ready or len(self.inflight) >= self.max_inflight:
            return False
        if self.max_computing < self.max_inflight and \
                sum(1 for c, _ in self.inflight if not self.pool.computed(c)) >= self.max_computing:
            return False
        ticks, members = [], []
        cap = max(1, math.ceil(len(self.sid) * self.cycle_share))
        for key in list(self.ready):
            if len(members) >= cap:
                break
            t0, rows = self.queue[key].popleft()
            ticks.append(rows if isinstance(rows, kr.StreamPool.Rows) else (self.sid[key], rows))
            members.append((key, t0))
            self.busy.add(key)
            del self.ready[key]
        t = self.clock()
        cyc = self.pool.submit(ticks)
        self.submit_s += self.clock() - t
        self.inflight.append((cyc, members))
        self.cycles += 1
        self.ticks_in_cycles += len(members)
        return True

    def step(self):
        self.poll()
        if self.dispatch() or not self.inflight:
            return
        computing = [c for c, _ in self.inflight if not self.pool.computed(c)]
        if self.ready and len(computing) >= self.max_computing and len(self.inflight) < self.max_inflight:
            # ticks wait for the computing cycle's phases only
            self.pool.computed(computing[-1], wait=True)
        else:
            # nothing to start: wait for the oldest cycle
            self.pool.done(self.inflight[0][0], wait=True)
        self.poll()

    def drain(self):
        while self.inflight or self.ready:
            self.step()
