"""IR core of the factor compiler: op nodes, traits, the thread-local builder
and the structural hash used to name materialised intermediates.

Behavioural restatement of KunQuant/Op.py (Menooker/KunQuant):
  * builder scoping            -> Op.py:10-47
  * structural hash + naming   -> Op.py:62-66, 115-129, 155-157
  * IR printing                -> Op.py:57-61, 144-150
  * verification rules         -> Op.py:159-176, 330-343, 391-393, 409-412, 526-535
  * operator sugar             -> Op.py:178-259
The hash formula and the printed IR format must stay identical to the
reference because generated buffer names ("1302190f66243c466") and the
reference's golden IR strings depend on them.
"""
from collections import OrderedDict
from typing import Dict, List, Optional, Tuple, Union
import threading
import zlib

__all__ = [
    "Builder", "OpBase", "GraphSourceTrait", "ConstantOp", "UnaryElementwiseOp",
    "BinaryElementwiseOp", "CompositiveOp", "WindowedDataSourceOp", "Input",
    "SinkOpTrait", "Output", "WindowedTempOutput", "WindowedTrait",
    "ForeachBackWindow", "IterValue", "WindowLoopIndex", "BoolOpTrait",
    "StatefulOpTrait", "GloablStatefulOpTrait", "GlobalStatefulProducerTrait",
    "StateConsumerTrait", "ReductionOp", "CrossSectionalOp",
    "SimpleCrossSectionalOp", "Rank", "Scale", "AcceptSingleValueInputTrait",
    "traverse_replace_map", "print_attrs", "make_indents",
    "GenericCppCodegenTrait",
]


class _ScopeState(threading.local):
    """Per-thread 'current builder' state: the op list new ops are appended
    to and the window loop new ops are placed in."""

    def __init__(self):
        self.ops = None
        self.loop = None


_scope = _ScopeState()


class Builder:
    """Context manager collecting every op constructed inside it."""

    def __init__(self, loop=None) -> None:
        self.ops: List["OpBase"] = []
        self.loop = loop
        self._saved = None

    def set_loop(self, loop) -> None:
        self.loop = loop
        _scope.loop = loop

    def __enter__(self):
        self._saved = (_scope.ops, _scope.loop)
        _scope.ops = self.ops
        _scope.loop = self.loop

    def __exit__(self, *exc):
        _scope.ops, _scope.loop = self._saved


def make_indents(indent: int, spaces=2) -> str:
    return " " * (indent * spaces)


def print_attrs(attr) -> str:
    if not attr:
        return ""
    return "{" + ",".join(f"{k}:{v}" for k, v in attr.items()) + "}"


_HASH_MOD = 2 << 64


def _mix(x: int, y: int) -> int:
    # Same combination rule as the reference (Op.py:65-66) so that
    # hash-derived buffer names match.
    return (x * 23 + y + 1 + (x >> 16) + (y >> 16) * 7) % _HASH_MOD


def traverse_replace_map(op: "OpBase", replace_map: Dict["OpBase", "OpBase"]) -> "OpBase":
    """Follow a chain of replacements until a fixed point."""
    while True:
        nxt = replace_map.get(op, None)
        if not nxt or nxt == op:
            return op
        op = nxt


class AcceptSingleValueInputTrait:
    """Ops that accept one `single_value` (shape [S]) Input."""

    def get_single_value_input_id(self) -> int:
        raise NotImplementedError()


_NO_ARGS = {}


class OpBase:
    def __init__(self, inputs: List["OpBase"], attrs) -> None:
        for i in inputs:
            if not isinstance(i, OpBase):
                raise RuntimeError("Bad inputs, given " + str(type(i)))
        self.inputs = inputs
        self._parent_loop = _scope.loop
        self.attrs = OrderedDict(attrs) if attrs is not None else OrderedDict()
        if _scope.ops is not None:
            _scope.ops.append(self)

    # ---- graph editing -------------------------------------------------
    def replace_inputs(self, replace_map: Dict["OpBase", "OpBase"]):
        for idx, v in enumerate(self.inputs):
            self.inputs[idx] = traverse_replace_map(v, replace_map)
        if self._parent_loop in replace_map:
            self._parent_loop = traverse_replace_map(self._parent_loop, replace_map)

    def set_parent(self, loop: "ForeachBackWindow") -> None:
        if loop is not None and not isinstance(loop, ForeachBackWindow):
            raise RuntimeError("set_parent failed")
        self._parent_loop = loop

    def get_parent(self) -> "ForeachBackWindow":
        return self._parent_loop

    def attrs_str(self) -> str:
        return print_attrs(self.attrs)

    # ---- structural identity ------------------------------------------
    def get_args(self, identity: bool, **kwargs) -> List[Tuple["OpBase", dict]]:
        assert not kwargs
        return [(v, _NO_ARGS) for v in self.inputs]

    def fast_hash(self, cache: Optional[Dict["OpBase", int]] = None, **kwargs) -> int:
        if cache is not None and self in cache:
            return cache[self]
        attr_bytes = print_attrs(self.attrs).encode()
        h = zlib.adler32(self.__class__.__name__.encode())
        h = _mix(h, (zlib.adler32(attr_bytes) << 32) + zlib.crc32(attr_bytes))
        h = _mix(h, 114514)
        for arg, sub in self.get_args(True, **kwargs):
            h = _mix(h, arg.fast_hash(cache, **sub))
        h = _mix(h, 1818910)
        if cache is not None:
            cache[self] = h
        return h

    def hash_hex(self, cache: Optional[Dict["OpBase", int]] = None) -> str:
        return f"{self.fast_hash(cache):016x}"

    def items(self, **kwargs):
        """Flattened structural token stream of this op and its arguments."""
        yield self.__class__
        yield self.attrs
        yield "["
        for arg, sub in self.get_args(True, **kwargs):
            yield from arg.items(**sub)
        yield "]"

    def to_string(self, indent: int, identity: bool, **kwargs) -> str:
        pad = make_indents(indent)
        args = ",\n".join(a.to_string(indent + 1, identity, **sub)
                          for a, sub in self.get_args(identity, **kwargs))
        return f"{pad}{self.__class__.__name__}@{print_attrs(self.attrs)}(\n{args}\n{pad})"

    def __str__(self) -> str:
        return self.to_string(0, False)

    def fast_str(self) -> str:
        args = ",".join(str(id(v)) for v in self.inputs)
        return f"{self.__class__.__name__}@{id(self)}@{print_attrs(self.attrs)}({args})"

    # ---- verification --------------------------------------------------
    def verify(self, func) -> None:
        if isinstance(self, WindowedTrait):
            WindowedTrait.verify(self, func)
        allowed = self.get_single_value_input_id() if isinstance(self, AcceptSingleValueInputTrait) else -1
        for idx, inp in enumerate(self.inputs):
            if isinstance(inp, Input) and idx != allowed and inp.attrs.get("single_value", False):
                raise RuntimeError("verify() failed: using single value input: " + str(self))
            loop = inp.get_parent()
            if loop is not None and loop != self._parent_loop:
                raise RuntimeError("verify() failed: referencing cross loop results: " + str(self))

    # ---- operator sugar ------------------------------------------------
    def _with_scalar(self, other, binary_cls, const_cls, reverse: bool):
        if type(other) in (int, float):
            return const_cls(self, other, [("swap", True)]) if reverse else const_cls(self, other)
        if isinstance(other, OpBase):
            return binary_cls(other, self) if reverse else binary_cls(self, other)
        raise RuntimeError("Don't know how to build binary " + str(type(other)))

    def _with_const_node(self, other, binary_cls):
        if type(other) in (int, float):
            other = ConstantOp(other)
        if isinstance(other, OpBase):
            return binary_cls(self, other)
        raise RuntimeError("Don't know how to build binary " + str(type(other)))

    def __sub__(self, other):
        from .ops.ElewiseOp import Sub, SubConst
        return self._with_scalar(other, Sub, SubConst, False)

    def __rsub__(self, other):
        from .ops.ElewiseOp import Sub, SubConst
        return self._with_scalar(other, Sub, SubConst, True)

    def __add__(self, other):
        from .ops.ElewiseOp import Add, AddConst
        return self._with_scalar(other, Add, AddConst, False)

    def __radd__(self, other):
        from .ops.ElewiseOp import Add, AddConst
        return self._with_scalar(other, Add, AddConst, False)

    def __mul__(self, other):
        from .ops.ElewiseOp import Mul, MulConst
        return self._with_scalar(other, Mul, MulConst, False)

    def __rmul__(self, other):
        from .ops.ElewiseOp import Mul, MulConst
        return self._with_scalar(other, Mul, MulConst, True)

    def __truediv__(self, other):
        from .ops.ElewiseOp import Div, DivConst
        return self._with_scalar(other, Div, DivConst, False)

    def __rtruediv__(self, other):
        from .ops.ElewiseOp import Div, DivConst
        return self._with_scalar(other, Div, DivConst, True)

    def __lt__(self, other):
        from .ops.ElewiseOp import LessThan, LessThanConst
        return self._with_scalar(other, LessThan, LessThanConst, False)

    def __ge__(self, other):
        from .ops.ElewiseOp import GreaterEqual
        return self._with_const_node(other, GreaterEqual)

    def __gt__(self, other):
        from .ops.ElewiseOp import GreaterThan
        return self._with_const_node(other, GreaterThan)

    def __le__(self, other):
        from .ops.ElewiseOp import LessEqual
        return self._with_const_node(other, LessEqual)

    def __or__(self, other):
        from .ops.ElewiseOp import Or
        return self._with_const_node(other, Or)

    def __and__(self, other):
        from .ops.ElewiseOp import And
        return self._with_const_node(other, And)

    def __invert__(self):
        from .ops.ElewiseOp import Not
        return Not(self)

    # ops are graph nodes: identity semantics for hashing and equality
    __hash__ = object.__hash__

    def __eq__(self, other):
        return self is other

    def __ne__(self, other):
        return self is not other


class GraphSourceTrait:
    """Graph sources (inputs, constants): nodes without inputs."""


class ConstantOp(OpBase, GraphSourceTrait):
    def __init__(self, v) -> None:
        super().__init__([], [("value", v)])
        self.check()

    def check(self):
        v = self.attrs["value"]
        if v != v:
            raise RuntimeError("ConstantOp value cannot be float('nan'), use 'nan' string literal instead")

    def verify(self, func) -> None:
        self.check()
        return super().verify(func)


class UnaryElementwiseOp(OpBase):
    def __init__(self, lhs: OpBase, attrs=None) -> None:
        super().__init__([lhs], attrs)


class BinaryElementwiseOp(OpBase):
    def __init__(self, lhs: OpBase, rhs: OpBase, attrs=None) -> None:
        super().__init__([lhs, rhs], attrs)


class CompositiveOp(OpBase):
    """An op that is lowered into primitive ops by the decompose pass."""

    def decompose(self, options: dict) -> List[OpBase]:
        raise NotImplementedError()


class WindowedDataSourceOp(OpBase):
    """Ops whose history can be read by a windowed op."""


class Input(WindowedDataSourceOp, GraphSourceTrait):
    def __init__(self, name: str) -> None:
        super().__init__([], [("name", name)])

    def to_string(self, indent: int, identity: bool, **kwargs) -> str:
        return "{}input({})".format(make_indents(indent), self.attrs["name"])


class SinkOpTrait:
    """Graph sinks: kept even when nothing uses them."""


class Output(WindowedDataSourceOp, SinkOpTrait):
    def __init__(self, inp: OpBase, name: str = "") -> None:
        super().__init__([inp], [("name", name)])


class WindowedTempOutput(WindowedDataSourceOp):
    """Marks that a window of past values of the input must be kept."""

    def __init__(self, inp: OpBase, window: int) -> None:
        super().__init__([inp], [("window", window)])


class WindowedTrait:
    """Ops that read a window of past values of their inputs."""

    def verify(self, func) -> None:
        if not func.strict_window:
            return
        need = self.required_input_window()
        for inp in self.inputs:
            if not isinstance(inp, WindowedDataSourceOp):
                raise RuntimeError("bad input for this op. The input should be windowed: "
                                   + str(inp) + "\n This op = " + str(self))
            if not isinstance(inp, (Input, Output)) and inp.attrs["window"] < need:
                raise RuntimeError("bad input for this op. The window does not match: "
                                   + str(inp) + "\n This op = " + str(self))

    def required_input_window(self) -> int:
        return self.attrs["window"]


class ForeachBackWindow(OpBase, WindowedTrait):
    """Loop over the last `window` rows of one or more windowed inputs."""

    def __init__(self, inp, window: int, *args) -> None:
        super().__init__([inp] + list(args), [("window", window)])

    def get_args(self, identity: bool, **kwargs):
        if not kwargs:
            return super().get_args(identity)
        shown = kwargs["display"]
        assert shown in self.inputs
        return [(shown, _NO_ARGS)]

    def __enter__(self):
        assert _scope.loop is None
        _scope.loop = self
        return self

    def __exit__(self, *exc):
        assert _scope.loop is self
        _scope.loop = None


class IterValue(OpBase):
    """Value of one windowed input of a loop at the current iteration."""

    def __init__(self, loop: ForeachBackWindow, src: OpBase) -> None:
        super().__init__([loop, src], None)
        self.set_parent(loop)

    def get_args(self, identity: bool, **kwargs):
        assert len(self.inputs) == 2
        if not identity:
            return super().get_args(identity, **kwargs)
        return [(self.inputs[0], {"display": self.inputs[1]}), (self.inputs[1], _NO_ARGS)]

    def verify(self, func) -> None:
        super().verify(func)
        if not isinstance(self.inputs[0], ForeachBackWindow):
            raise RuntimeError("Bad IterValue: " + str(self))


class WindowLoopIndex(OpBase):
    """Iteration index of a loop: 0 for the oldest row, window-1 for the newest."""

    def __init__(self, forwindow: ForeachBackWindow) -> None:
        super().__init__([forwindow], [])
        self.set_parent(forwindow)

    def verify(self, func) -> None:
        super().verify(func)
        if not isinstance(self.inputs[0], ForeachBackWindow):
            raise RuntimeError("Bad WindowLoopIndex, input must be a forloop: " + str(self))


class BoolOpTrait:
    """Ops producing a boolean mask."""


class GenericCppCodegenTrait:
    """Naming hooks kept for API compatibility with user-defined stateful ops."""

    def get_func_or_class_template_name(self) -> str:
        return self.__class__.__name__


class StatefulOpTrait(GenericCppCodegenTrait):
    """Ops with internal state."""


class GloablStatefulOpTrait(StatefulOpTrait):
    """Ops whose state is carried from one time step to the next."""


class GlobalStatefulProducerTrait(GloablStatefulOpTrait):
    """Carried state that is read by StateConsumerTrait ops."""


class StateConsumerTrait:
    """Ops reading the state of a GlobalStatefulProducerTrait op."""


class ReductionOp(OpBase, StatefulOpTrait):
    """Reduces values produced inside a window loop; used after the loop."""

    def __init__(self, v: OpBase, init_val: OpBase = None, attrs=None) -> None:
        super().__init__([v] if init_val is None else [v, init_val], attrs)

    def get_loop(self) -> ForeachBackWindow:
        inp = self.inputs[0]
        if isinstance(inp, ForeachBackWindow):
            return inp
        return inp.get_parent()

    def verify(self, func) -> None:
        assert len(self.inputs) <= 2
        loop = self.get_loop()
        if loop is None or loop == self._parent_loop:
            raise RuntimeError(f"verify() failed: ReductionOp using non-loop result: {self}\n"
                               f"self._parent_loop = {self._parent_loop}\nloop = {loop}")
        if self._parent_loop != loop.get_parent():
            raise RuntimeError(f"verify() failed: ReductionOp not in parent of input: {self}\n"
                               f"self._parent_loop = {self._parent_loop}\n"
                               f"loop.get_parent() = {loop.get_parent()}")


class CrossSectionalOp(OpBase):
    pass


class SimpleCrossSectionalOp(CrossSectionalOp):
    def __init__(self, v: OpBase, attrs=None) -> None:
        super().__init__([v], attrs)


class Rank(SimpleCrossSectionalOp):
    """Cross-sectional percentile rank in [0, 1] (average method)."""


class Scale(SimpleCrossSectionalOp):
    """Cross-sectional x / sum(|x|)."""
