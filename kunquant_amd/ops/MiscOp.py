"""Stateful and miscellaneous primitives.

Class set and verification rules of KunQuant/ops/MiscOp.py:
BackRef 5-13, FastWindowedSum 16-30, Accumulator 32-55, SetAccumulator 57-71,
ReturnFirstValue 73-79, ExpMovingAvg 81-116, WindowedLinearRegression 118-163,
SkipListState/consumers 165-199, GenericCrossSectionalOp 201-249.
"""
from typing import List, Union

from ..Op import (AcceptSingleValueInputTrait, BinaryElementwiseOp, CrossSectionalOp,
                  GloablStatefulOpTrait, GlobalStatefulProducerTrait, Input, OpBase,
                  SinkOpTrait, StateConsumerTrait, UnaryElementwiseOp, WindowedTrait)


class BackRef(OpBase, WindowedTrait):
    """The value `window` rows ago (NaN before the first row)."""

    def __init__(self, v: OpBase, window: int) -> None:
        super().__init__([v], [("window", window)])

    def required_input_window(self) -> int:
        return self.attrs["window"] + 1


class FastWindowedSum(OpBase, WindowedTrait, GloablStatefulOpTrait):
    """Rolling sum with Kahan-compensated add/remove (no window loop)."""

    def __init__(self, v: OpBase, window: int) -> None:
        super().__init__([v], [("window", window)])

    def required_input_window(self) -> int:
        return self.attrs["window"] + 1


class Accumulator(OpBase, GlobalStatefulProducerTrait):
    """A value carried across time steps; updated by one SetAccumulator."""

    def __init__(self, v: OpBase, name: str) -> None:
        super().__init__([v], [("name", name)])

    def verify(self, func) -> None:
        setter = None
        for user in func.op_to_id[self].uses:
            if isinstance(user, SetAccumulator):
                if setter is not None:
                    raise RuntimeError(f"Accumulator {self.attrs['name']} can only be used with one "
                                       f"SetAccumulator: {user}\nand  {setter}")
                setter = user
        if setter is None:
            raise RuntimeError(f"Accumulator {self.attrs['name']} is not used with any SetAccumulator")
        return super().verify(func)


class SetAccumulator(OpBase, StateConsumerTrait):
    """acc := mask ? value : acc; evaluates to the new accumulator value."""

    def __init__(self, accu: OpBase, mask: OpBase, value: OpBase) -> None:
        super().__init__([accu, mask, value], [])
        self.check()

    def check(self):
        if not isinstance(self.inputs[0], Accumulator):
            raise RuntimeError("SetAccumulator expects the first input to be an Accumulator op")

    def verify(self, func) -> None:
        self.check()
        return super().verify(func)


class ReturnFirstValue(OpBase):
    """Evaluates to its first input; the others are kept as dependencies."""

    def __init__(self, v: List[OpBase]) -> None:
        super().__init__(v, [])


class ExpMovingAvg(OpBase, GloablStatefulOpTrait, AcceptSingleValueInputTrait):
    """ewm(span=window, adjust=False, ignore_na=True).mean(); optional initial
    value from a single-value Input whose name starts with "__init"."""

    def __init__(self, v: OpBase, window: int, init_val: Union[Input, None] = None) -> None:
        args = [v]
        if init_val is not None:
            if (not isinstance(init_val, Input) or not init_val.attrs.get("single_value", False)
                    or not init_val.attrs.get("name", "").startswith("__init")):
                raise RuntimeError("EMA expects init_val to be Input op with single_value=True and "
                                   "name starting with __init")
            args.append(init_val)
        super().__init__(args, [("window", window)])

    def get_single_value_input_id(self) -> int:
        return 1


class WindowedLinearRegression(OpBase, WindowedTrait, GlobalStatefulProducerTrait):
    """Rolling regression sums of y against x = 1..window."""

    def __init__(self, v: OpBase, window: int) -> None:
        super().__init__([v], [("window", window)])

    def required_input_window(self) -> int:
        return self.attrs["window"] + 1


class WindowedLinearRegressionConsumerTrait(StateConsumerTrait):
    def verify(self, func) -> None:
        if not self.inputs or not isinstance(self.inputs[0], WindowedLinearRegression):
            raise RuntimeError("WindowedLinearRegressionImpl expects WindowedLinearRegression Op as input")
        return OpBase.verify(self, func)


class WindowedLinearRegressionImplUnaryBase(WindowedLinearRegressionConsumerTrait, UnaryElementwiseOp):
    pass


class WindowedLinearRegressionImplBinaryBase(WindowedLinearRegressionConsumerTrait, BinaryElementwiseOp):
    pass


class WindowedLinearRegressionRSqaureImpl(WindowedLinearRegressionImplUnaryBase):
    """R-squared of the rolling regression."""


class WindowedLinearRegressionSlopeImpl(WindowedLinearRegressionImplUnaryBase):
    """Slope of the rolling regression."""


class WindowedLinearRegressionResiImpl(WindowedLinearRegressionImplBinaryBase):
    """Residual of the newest value against the rolling regression."""


class SkipListState(OpBase, GlobalStatefulProducerTrait):
    """Sorted multiset of the last `window` non-NaN values."""

    def __init__(self, oldvalue: OpBase, value: OpBase, window: int) -> None:
        super().__init__([oldvalue, value], [("window", window)])


class SkipListConsumerOp(StateConsumerTrait):
    def verify(self, func) -> None:
        if not self.inputs or not isinstance(self.inputs[0], SkipListState):
            raise RuntimeError("SkipListConsumerOp expects SkipListState Op as input")
        return super().verify(func)


class SkipListQuantile(SkipListConsumerOp, OpBase):
    def __init__(self, v: OpBase, q: float) -> None:
        OpBase.__init__(self, [v], [("q", q)])


class SkipListRank(SkipListConsumerOp, BinaryElementwiseOp):
    pass


class SkipListMin(SkipListConsumerOp, UnaryElementwiseOp):
    pass


class SkipListMax(SkipListConsumerOp, UnaryElementwiseOp):
    pass


class SkipListArgMin(SkipListConsumerOp, OpBase):
    pass


class GenericCrossSectionalOp(CrossSectionalOp):
    """Cross-sectional op with a user-supplied body of C++ statements over one
    row (MiscOp.py:201-224 of the reference). The body is compiled as HIP
    device code and runs one device thread per row. Names in scope:
    `T` (float or double), `num_stocks`, `input_<n>[i]` (input n at stock
    i of the row), `output_<n>[i]`; `generate_head()` runs first, in the
    same scope. Device code only: no host library calls."""

    def generate_body(self) -> str:
        raise NotImplementedError("GenericCrossSectionalOp must be specialized")

    def generate_head(self) -> str:
        raise NotImplementedError("GenericCrossSectionalOp must be specialized")


class DiffWithWeightedSum(GenericCrossSectionalOp):
    """out[i] = v[i] - sum_j v[j] * w[j] over the row (MiscOp.py:227-249 of
    the reference): the sum runs over the stocks in order, products and
    additions rounded one at a time."""

    def __init__(self, v: OpBase, w: OpBase) -> None:
        super().__init__([v, w], None)

    def generate_head(self) -> str:
        return ""

    def generate_body(self) -> str:
        return """
        T acc = 0;
        for (size_t i = 0; i < num_stocks; ++i) {
            const T prod = input_0[i] * input_1[i];
            acc = acc + prod;
        }
        for (size_t i = 0; i < num_stocks; ++i) {
            output_0[i] = input_0[i] - acc;
        }
        """


class GenericPartition(OpBase, SinkOpTrait):
    pass
