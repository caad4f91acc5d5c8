"""Factor graphs of the benchmark workloads.

simple_graph(): x = (close - open) / ((high - low) + 0.001), WindowedStddev(x, 20),
WindowedSum(volume, 10) and the cross-sectional Rank of x.

random_graph(n, seed): n seeded random expression trees of depth 3-8 over
open/high/low/close/volume/amount with elementwise arithmetic,
compare + Select, Delay/Delta, WindowedSum/Avg/Stddev/Corr/Cov,
TsArgMax/TsRank and DecayLinear (windows from {5, 10, 20, 30, 60}); about
30% of the internal nodes are cross-sectional Rank or Scale."""
import random
from typing import List

from ..Op import Builder, ConstantOp, Input, OpBase, Output, Rank, Scale
from ..Stage import Function
from ..ops import (Abs, Add, BackRef, DecayLinear, Div, GreaterThan, LessThan, Mul, Select, Sub, TsArgMax,
                   TsRank, WindowedAvg, WindowedCorrelation, WindowedCovariance, WindowedStddev, WindowedSum)
from .data import FIELDS

WINDOWS = (5, 10, 20, 30, 60)


def simple_graph() -> Function:
    b = Builder()
    with b:
        close, opn, high, low, vol = (Input(n) for n in ("close", "open", "high", "low", "volume"))
        x = (close - opn) / ((high - low) + 0.001)
        Output(x, "x")
        Output(WindowedStddev(x, 20), "std20")
        Output(WindowedSum(vol, 10), "vsum10")
        Output(Rank(x), "rank_x")
    return Function(b.ops)


class _TreeGen:
    def __init__(self, seed: int, cs_frac: float = 0.3):
        self.rng = random.Random(seed)
        self.cs_frac = cs_frac
        self.inputs = {}
        self.counts = {"nodes": 0, "cs": 0, "windowed": 0}

    def leaf(self) -> OpBase:
        name = self.rng.choice(FIELDS)
        if name not in self.inputs:
            self.inputs[name] = Input(name)
        return self.inputs[name]

    def window(self) -> int:
        return self.rng.choice(WINDOWS)

    def tree(self, depth: int) -> OpBase:
        if depth <= 0:
            return self.leaf()
        r = self.rng
        self.counts["nodes"] += 1
        if r.random() < self.cs_frac:
            self.counts["cs"] += 1
            child = self.tree(depth - 1)
            return Rank(child) if r.random() < 0.6 else Scale(child)
        kind = r.random()
        if kind < 0.22:
            a = self.tree(depth - 1)
            b = self.tree(r.randint(0, depth - 1))
            if r.random() < 0.5:
                a, b = b, a
            op = r.choice((Add, Sub, Mul, Div))
            return op(a, b)
        if kind < 0.30:
            a = self.tree(depth - 1)
            c = r.choice((0.5, 2.0, 1.5, -1.0, 10.0))
            return r.choice((lambda x: x * c, lambda x: x + c, lambda x: Abs(x)))(a)
        if kind < 0.36:
            a = self.tree(depth - 1)
            b = self.tree(r.randint(0, depth - 1))
            cmp = GreaterThan(a, b) if r.random() < 0.5 else LessThan(a, b)
            return Select(cmp, a, b)
        if kind < 0.44:
            self.counts["windowed"] += 1
            a = self.tree(depth - 1)
            w = self.window()
            return BackRef(a, w) if r.random() < 0.5 else Sub(a, BackRef(a, w))
        if kind < 0.88:
            self.counts["windowed"] += 1
            a = self.tree(depth - 1)
            op = r.choice((WindowedSum, WindowedAvg, WindowedStddev, TsArgMax, TsRank, DecayLinear))
            return op(a, self.window())
        self.counts["windowed"] += 1
        a = self.tree(depth - 1)
        b = self.tree(r.randint(0, depth - 1))
        op = WindowedCorrelation if r.random() < 0.5 else WindowedCovariance
        return op(a, self.window(), b)


def random_graph(num_factors: int, seed: int, prefix: str = "f") -> Function:
    gen = _TreeGen(seed)
    b = Builder()
    with b:
        for i in range(num_factors):
            depth = gen.rng.randint(3, 8)
            Output(gen.tree(depth), f"{prefix}{i:03d}")
    f = Function(b.ops)
    f.workload_stats = dict(gen.counts)
    return f
