"""Elementwise primitives (class set of KunQuant/ops/ElewiseOp.py:1-120).

Lane semantics are implemented by the device templates
(kunquant_amd/csrc/kun_device.hpp) and by the oracle."""
from ..Op import UnaryElementwiseOp, BinaryElementwiseOp, OpBase, BoolOpTrait


class BinaryConstOp(UnaryElementwiseOp):
    """Binary op with one constant operand `value`; `swap` puts the
    constant on the left-hand side."""

    def __init__(self, inp: OpBase, v: float, swap: bool = False) -> None:
        attrs = [("value", v)]
        if swap:
            attrs.append(("swap", swap))
        super().__init__(inp, attrs)


class AddConst(BinaryConstOp): pass
class SubConst(BinaryConstOp): pass
class MulConst(BinaryConstOp): pass
class DivConst(BinaryConstOp): pass

class Mul(BinaryElementwiseOp): pass
class Add(BinaryElementwiseOp): pass
class Sub(BinaryElementwiseOp): pass
class Or(BinaryElementwiseOp, BoolOpTrait): pass
class And(BinaryElementwiseOp, BoolOpTrait): pass
class Max(BinaryElementwiseOp): pass
class Min(BinaryElementwiseOp): pass
class Div(BinaryElementwiseOp): pass

class CmpOp(BinaryElementwiseOp, BoolOpTrait): pass
class GreaterThan(CmpOp): pass
class GreaterEqual(CmpOp): pass
class LessThan(CmpOp): pass
class LessEqual(CmpOp): pass
class Equals(CmpOp): pass

class GreaterThanConst(BinaryConstOp, BoolOpTrait): pass
class LessThanConst(BinaryConstOp, BoolOpTrait): pass


class Sqrt(UnaryElementwiseOp):
    """Square root."""


class Log(UnaryElementwiseOp):
    """Natural logarithm."""


class Abs(UnaryElementwiseOp):
    """Absolute value."""


class Sign(UnaryElementwiseOp):
    """+1 / 0 / -1 (NaN stays NaN)."""


class Not(UnaryElementwiseOp, BoolOpTrait):
    """Logical not."""


class Exp(UnaryElementwiseOp):
    """Natural exponential."""


class SetInfOrNanToValue(UnaryElementwiseOp):
    """Replace +-inf and NaN by `value`."""

    def __init__(self, lhs: OpBase, value: float = 0.0) -> None:
        super().__init__(lhs, [("value", value)])


class Select(OpBase):
    """cond ? true_v : false_v, decided by the sign bit of cond."""

    def __init__(self, cond: OpBase, true_v: OpBase, false_v: OpBase) -> None:
        super().__init__([cond, true_v, false_v], None)
