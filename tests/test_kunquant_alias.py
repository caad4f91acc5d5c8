"""The `KunQuant` alias package (kunquant_amd/compat): code written against
the reference's package imports its names unchanged and gets this
package's modules -- graph building, compilation, the Alpha101 library and
the KunRunner binding -- in the style of the reference's own tests
(tests/test_runtime.py:1-20, tests/test_alpha101.py:1-12)."""
import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COMPAT = os.path.join(ROOT, "kunquant_amd", "compat")


@pytest.fixture()
def alias():
    sys.path.insert(0, COMPAT)
    try:
        yield importlib.import_module("KunQuant")
    finally:
        sys.path.remove(COMPAT)


def test_reference_style_imports(alias):
    from KunQuant.Driver import KunCompilerConfig, compileit
    from KunQuant.Op import Builder, Input, Output, Rank
    from KunQuant.ops import WindowedStddev, WindowedSum
    from KunQuant.Stage import Function
    from KunQuant.jit import cfake
    from KunQuant.predefined.Alpha101 import AllData, all_alpha
    from KunQuant.runner import KunRunner as kr
    import kunquant_amd.Op
    assert Builder is kunquant_amd.Op.Builder
    assert kr.runGraph is importlib.import_module("kunquant_amd.runner.KunRunner").runGraph
    assert callable(cfake.compileit) and len(all_alpha) == 82
    b = Builder()
    with b:
        close, opn = Input("close"), Input("open")
        Output(WindowedStddev(close - opn, 10), "std10")
        Output(Rank(WindowedSum(close, 5)), "rank5")
    ir = compileit(Function(b.ops), "alias_mod", **KunCompilerConfig(input_layout="TS", output_layout="TS").__dict__)
    assert ir.output_names() == ["std10", "rank5"]
    b = Builder()
    with b:
        d = AllData(low=Input("low"), high=Input("high"), close=Input("close"), open=Input("open"),
                    amount=Input("amount"), volume=Input("volume"))
        Output(all_alpha[-1](d), "alpha101")
    assert compileit(Function(b.ops), "alias101", input_layout="TS", output_layout="TS").output_names() == ["alpha101"]
