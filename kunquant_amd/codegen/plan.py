"""Kernel plan of a batch time-series stage.

The ops of a stage split into two classes:
  * recurrences -- FastWindowedSum (running Kahan add/subtract),
    ExpMovingAvg, WindowedLinearRegression sums, SkipListState, and the
    Accumulator webs: the value at row t depends on every row since the run
    start through floating-point rounding, so each stock walks its rows in
    order (the reference's semantics, Ops.hpp:300-375, 410-440);
  * window-pure ops -- elementwise ops, BackRef, window loops and their
    reductions: the value at row t is a fixed function of at most the last
    `w` rows of the op's inputs, so any range of rows can be computed
    independently once `w` rows of history are rebuilt.

Levels alternate: even levels hold window-pure ops and run as PAR kernels
(rows split into chunks, each chunk recomputing a halo of history before its
first stored row); odd levels hold recurrences and run as SEQ kernels (one
lane per stock walks all rows). An op takes the lowest level its inputs
allow; elementwise ops stay on the level of their inputs. A value read on a
higher level than the one computing it is materialised in HBM as a
[rows][stocks] array. Within a level, ops that share no value form
independent components, launched side by side along the grid's z axis.

Stages holding Accumulator webs (or anything the split cannot express) run
as a single SEQ level: every row in order, the level split into the
independent op groups of the stage (row-local values of the inputs are
recomputed by each group that reads them)."""
from typing import Dict, List, Optional, Set

from ..Op import ConstantOp, ForeachBackWindow, Input, OpBase, Output, ReductionOp

PAR, SEQ = 0, 1

REC_OPS = {"FastWindowedSum", "ExpMovingAvg", "WindowedLinearRegression", "SkipListState"}
BOUND_OPS = {"WindowedLinearRegressionRSqaureImpl", "WindowedLinearRegressionSlopeImpl",
             "WindowedLinearRegressionResiImpl", "SkipListQuantile", "SkipListRank", "SkipListMin",
             "SkipListMax", "SkipListArgMin"}
ACC_OPS = {"Accumulator", "SetAccumulator"}


def is_source(op: OpBase) -> bool:
    """Inputs and constants are available on every level."""
    return isinstance(op, (Input, ConstantOp))


class Component:
    """Ops of one independent group of a level, in program order (inputs
    and constants it reads are replicated into it)."""

    def __init__(self, ops: List[OpBase], ext: List[OpBase]):
        self.ops = ops
        self.ext = ext  # materialised values of lower levels read here
        # Outputs this group computes only for its own reads of their
        # history; another group stores them
        self.shadow: Set[OpBase] = set()


class Level:
    def __init__(self, kind: int, comps: List[Component]):
        self.kind = kind
        self.comps = comps


class StagePlan:
    def __init__(self, levels: List[Level], mat: Dict[OpBase, int], fallback: Optional[str]):
        self.levels = levels
        self.mat = mat  # materialised op -> slot in the stage's [slot][rows][stocks] arena
        self.fallback = fallback  # reason the stage runs as one SEQ level, or None


def _units(ops: List[OpBase]):
    """Top-level schedulable units: a window loop with its body and
    reductions, or a single op."""
    units = []
    unit_of: Dict[OpBase, int] = {}
    for op in ops:
        if isinstance(op, ReductionOp) or op.get_parent() is not None:
            continue
        if isinstance(op, ForeachBackWindow):
            members = [op] + [o for o in ops if o.get_parent() is op]
            members += [o for o in ops if isinstance(o, ReductionOp) and o.get_loop() is op]
        else:
            members = [op]
        for o in members:
            unit_of[o] = len(units)
        units.append(members)
    return units, unit_of


_NOT_REPLICABLE = REC_OPS | ACC_OPS | BOUND_OPS | {"IterValue"}


def _replicable(ops: List[OpBase]) -> Set[OpBase]:
    """Row-local values of the inputs (elementwise ops and lags over inputs
    and constants, nothing stateful; a windowed temporary of one, each group
    then keeping its own history of it): cheap to recompute, so each group
    that reads one gets its own copy instead of being joined to the others."""
    rep: Set[OpBase] = set()
    for op in ops:
        if is_source(op) or op.get_parent() is not None:
            continue
        if isinstance(op, (ForeachBackWindow, ReductionOp, Output)) or op.__class__.__name__ in _NOT_REPLICABLE:
            continue
        if all(is_source(i) or i in rep for i in op.inputs):
            rep.add(op)
    return rep


def _single_level(ops: List[OpBase], reason: str) -> StagePlan:
    """The whole stage as one sequential level, split into the independent
    op groups it holds: ops joined by a value, a loop and its body or
    reductions (an Accumulator and its SetAccumulator are joined by the
    value edge between them) share a group; inputs, constants and row-local
    values of them are replicated into every group reading them. Each group
    walks its rows on its own lanes, so a stage of many independent running
    statistics fills the GPU instead of one lane per stock."""
    rep = _replicable(ops)
    # an Output of a row-local value whose history other ops read: every
    # group reading it computes it (and keeps its history), one stores it
    read_out = {i for o in ops for i in o.inputs if isinstance(i, Output)}
    rep_out = {o for o in read_out if is_source(o.inputs[0]) or o.inputs[0] in rep}
    shared = lambda o: is_source(o) or o in rep or o in rep_out
    parent: Dict[OpBase, OpBase] = {}

    def find(x):
        while parent[x] is not x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    def union(a, b):
        ra, rb = find(a), find(b)
        if ra is not rb:
            parent[ra] = rb

    members = [op for op in ops if not shared(op)]
    for op in members:
        parent[op] = op
    for op in members:
        for inp in op.inputs:
            if not shared(inp):
                union(op, inp)
        if op.get_parent() is not None:
            union(op, op.get_parent())
        if isinstance(op, ReductionOp):
            union(op, op.get_loop())
    groups: Dict[OpBase, Set[OpBase]] = {}
    for op in members:
        groups.setdefault(find(op), set()).add(op)
    pos = {op: i for i, op in enumerate(ops)}
    comps = []
    for g in sorted(groups.values(), key=lambda g: min(pos[o] for o in g)):
        need = set(g)
        todo = [i for o in g for i in o.inputs if shared(i)]
        while todo:
            x = todo.pop()
            if x not in need:
                need.add(x)
                todo.extend(i for i in x.inputs if shared(i))
        comps.append(Component([op for op in ops if op in need], []))
    for o in sorted(rep_out, key=lambda o: pos[o]):
        holders = [c for c in comps if o in c.ops]
        if not holders:
            need, todo = {o}, list(o.inputs)
            while todo:
                x = todo.pop()
                if x not in need:
                    need.add(x)
                    todo.extend(x.inputs)
            comps.append(Component([op for op in ops if op in need], []))
            continue
        for c in holders[1:]:
            c.shadow.add(o)
    return StagePlan([Level(SEQ, comps)], {}, reason)


def plan_stage(ops: List[OpBase], force_seq: bool = False) -> StagePlan:
    if force_seq:
        return _single_level(ops, "forced")
    names = {op.__class__.__name__ for op in ops}
    if names & ACC_OPS:
        return _single_level(ops, "accumulator")
    for op in ops:
        p = op.get_parent()
        if p is not None and p.get_parent() is not None:
            return _single_level(ops, "nested loop")
    units, unit_of = _units(ops)
    if len(unit_of) != len(ops):
        return _single_level(ops, "unscheduled ops")
    lvl: Dict[int, Optional[int]] = {}
    for u, members in enumerate(units):
        head = members[0]
        if len(members) == 1 and is_source(head):
            lvl[u] = None
            continue
        m = 0
        for o in members:
            for inp in o.inputs:
                v = unit_of[inp]
                if v != u and lvl[v] is not None:
                    m = max(m, lvl[v])
        name = head.__class__.__name__
        if name in REC_OPS:
            lvl[u] = m if m % 2 == 1 else m + 1
        elif name in BOUND_OPS:
            state = lvl[unit_of[head.inputs[0]]]
            if state is None or m > state:
                return _single_level(ops, "state consumer above its state")
            lvl[u] = state
        elif isinstance(head, ForeachBackWindow) or name == "BackRef":
            lvl[u] = m if m % 2 == 0 else m + 1
        else:
            lvl[u] = m
    # Sequential levels walk every row on one lane per stock, so they keep
    # only what their recurrences need: a unit on a sequential level that no
    # recurrence (or state consumer) of the level depends on moves to the
    # next level, where rows run in parallel; an Output (or temp output)
    # moves with its input.
    def needs(u):
        return [unit_of[inp] for o in units[u] for inp in o.inputs if unit_of[inp] != u]

    needed = set()
    for u in reversed(range(len(units))):
        L = lvl[u]
        if L is None or L % 2 == 0:
            continue
        name = units[u][0].__class__.__name__
        if name in REC_OPS or name in BOUND_OPS or u in needed:
            needed.add(u)
            for v in needs(u):
                if lvl[v] == L:
                    needed.add(v)
    passthrough = ("WindowedTempOutput", "ReturnFirstValue")
    for u in range(len(units)):
        L = lvl[u]
        if L is None or L % 2 == 0 or u in needed or len(units[u]) != 1:
            continue
        head = units[u][0]
        if isinstance(head, Output) or head.__class__.__name__ in passthrough:
            if lvl[unit_of[head.inputs[0]]] == L:
                continue
        lvl[u] = L + 1
    level_of = {op: lvl[unit_of[op]] for op in ops}
    # materialised values: computed on one level, read on a higher one
    mat: Dict[OpBase, int] = {}
    for op in ops:
        lo = level_of[op]
        if lo is None:
            continue
        for inp in op.inputs:
            li = level_of[inp]
            if li is not None and li != lo and inp not in mat:
                assert li < lo
                mat[inp] = len(mat)
    # components per level (union-find over same-level data edges)
    parent: Dict[OpBase, OpBase] = {}

    def find(x):
        while parent[x] is not x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    def union(a, b):
        ra, rb = find(a), find(b)
        if ra is not rb:
            parent[rb] = ra

    for op in ops:
        if level_of[op] is not None:
            parent[op] = op
    for members in units:
        if level_of[members[0]] is None:
            continue
        for o in members[1:]:
            union(members[0], o)
        for o in members:
            for inp in o.inputs:
                if level_of[inp] == level_of[o]:
                    union(o, inp)
    order = {op: i for i, op in enumerate(ops)}
    levels: List[Level] = []
    for L in sorted({v for v in level_of.values() if v is not None}):
        groups: Dict[OpBase, List[OpBase]] = {}
        for op in ops:
            if level_of[op] == L:
                groups.setdefault(find(op), []).append(op)
        comps = []
        for members in groups.values():
            mset: Set[OpBase] = set(members)
            srcs: Set[OpBase] = set()
            ext: Set[OpBase] = set()
            for o in members:
                for inp in o.inputs:
                    if inp in mset:
                        continue
                    if level_of[inp] is None:
                        srcs.add(inp)
                    else:
                        ext.add(inp)
            allops = sorted(mset | srcs, key=order.__getitem__)
            comps.append(Component(allops, sorted(ext, key=order.__getitem__)))
        comps.sort(key=lambda c: order[c.ops[0]])
        levels.append(Level(PAR if L % 2 == 0 else SEQ, comps))
    return StagePlan(levels, mat, None)
