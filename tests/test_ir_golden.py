"""Golden IR strings of the reference compiler's own unit tests.

The graphs and the expected printed IR (including hash-derived buffer names)
are the fixtures of KunQuant tests/test.py:16-507 and tests/test2.py:33-119;
the front-end restatement must reproduce them exactly."""
import pytest

from kunquant_amd.Op import *
from kunquant_amd.Stage import Function
from kunquant_amd.ops import *
from kunquant_amd.passes import *


def expect(f, text):
    assert str(f) == text


def _avg_and_stddev():
    b = Builder()
    with b:
        a = Input("a")
        v1 = WindowedAvg(a, 10)
        v2 = WindowedStddev(a, 10)
        Output(v1, "ou1")
        Output(v2, "ou2")
    return Function(b.ops)


def test_simple():  # test.py:16-29
    b = Builder()
    with b:
        v1 = Mul(Input("a"), Input("b"))
        Output(AddConst(v1, 10))
    expect(Function(b.ops), '''v0 = Input@{name:a}()
v1 = Input@{name:b}()
v2 = Mul@(v0,v1)
v3 = AddConst@{value:10}(v2)
v4 = Output@{name:}(v3)''')


def test_decompose():  # test.py:41-64
    f = _avg_and_stddev()
    decompose(f)
    expect(f, '''v0 = Input@{name:a}()
v1 = WindowedTempOutput@{window:10}(v0)
v2 = ForeachBackWindow@{window:10}(v1)
v3 = IterValue@(v2,v1) in v2
v4 = ReduceAdd@(v3)
v5 = DivConst@{value:10}(v4)
v6 = WindowedTempOutput@{window:10}(v0)
v7 = ForeachBackWindow@{window:10}(v6)
v8 = IterValue@(v7,v6) in v7
v9 = ReduceAdd@(v8)
v10 = DivConst@{value:10}(v9)
v11 = WindowedTempOutput@{window:10}(v0)
v12 = ForeachBackWindow@{window:10}(v11)
v13 = IterValue@(v12,v11) in v12
v14 = Sub@(v13,v10) in v12
v15 = Mul@(v14,v14) in v12
v16 = ReduceAdd@(v15)
v17 = DivConst@{value:9}(v16)
v18 = Sqrt@(v17)
v19 = Output@{name:ou1}(v5)
v20 = Output@{name:ou2}(v18)''')


def test_fold_and_fast_sum():  # test.py:66-101
    f = _avg_and_stddev()
    decompose(f)
    expr_fold(f)
    expect(f, '''v0 = Input@{name:a}()
v1 = WindowedTempOutput@{window:10}(v0)
v2 = ForeachBackWindow@{window:10}(v1)
v3 = IterValue@(v2,v1) in v2
v4 = ReduceAdd@(v3)
v5 = DivConst@{value:10}(v4)
v6 = ForeachBackWindow@{window:10}(v1)
v7 = IterValue@(v6,v1) in v6
v8 = Sub@(v7,v5) in v6
v9 = Mul@(v8,v8) in v6
v10 = ReduceAdd@(v9)
v11 = DivConst@{value:9}(v10)
v12 = Sqrt@(v11)
v13 = Output@{name:ou1}(v5)
v14 = Output@{name:ou2}(v12)''')
    special_optimize(f)
    expect(f, '''v0 = Input@{name:a}()
v1 = WindowedTempOutput@{window:11}(v0)
v2 = FastWindowedSum@{window:10}(v1)
v3 = DivConst@{value:10}(v2)
v4 = ForeachBackWindow@{window:10}(v1)
v5 = IterValue@(v4,v1) in v4
v6 = Sub@(v5,v3) in v4
v7 = Mul@(v6,v6) in v4
v8 = ReduceAdd@(v7)
v9 = DivConst@{value:9}(v8)
v10 = Sqrt@(v9)
v11 = Output@{name:ou1}(v3)
v12 = Output@{name:ou2}(v10)''')


def test_gc():  # test.py:103-118
    b = Builder()
    with b:
        a, c = Input("a"), Input("b")
        v1 = Mul(a, c)
        unused = Mul(a, c)
        Mul(unused, unused)
        Output(AddConst(v1, 10))
    expect(Function(b.ops), '''v0 = Input@{name:a}()
v1 = Input@{name:b}()
v2 = Mul@(v0,v1)
v3 = AddConst@{value:10}(v2)
v4 = Output@{name:}(v3)''')


def test_tempwindow_elim():  # test.py:120-172
    b = Builder()
    with b:
        a = Input("a")
        Output(ReduceAdd(ForeachBackWindow(WindowedTempOutput(a, 10), 10)))
    f = Function(b.ops)
    temp_window_elim(f)
    expect(f, '''v0 = Input@{name:a}()
v1 = ForeachBackWindow@{window:10}(v0)
v2 = ReduceAdd@(v1)
v3 = Output@{name:}(v2)''')

    b = Builder()
    with b:
        x = Input("a")
        x = Mul(x, x)
        v1 = ReduceAdd(ForeachBackWindow(WindowedTempOutput(x, 10), 10))
        Output(x)
        Output(v1)
    f = Function(b.ops)
    temp_window_elim(f)
    expect(f, '''v0 = Input@{name:a}()
v1 = Mul@(v0,v0)
v2 = Output@{name:}(v1)
v3 = ForeachBackWindow@{window:10}(v2)
v4 = ReduceAdd@(v3)
v5 = Output@{name:}(v4)''')

    b = Builder()
    with b:
        x = Input("a")
        x = Mul(x, x)
        v1 = ReduceAdd(ForeachBackWindow(WindowedTempOutput(x, 10), 10))
        v2 = ReduceAdd(ForeachBackWindow(WindowedTempOutput(x, 15), 10))
        Output(v1)
        Output(v2)
    f = Function(b.ops)
    temp_window_elim(f)
    expect(f, '''v0 = Input@{name:a}()
v1 = Mul@(v0,v0)
v2 = WindowedTempOutput@{window:15}(v1)
v3 = ForeachBackWindow@{window:10}(v2)
v4 = ReduceAdd@(v3)
v5 = ForeachBackWindow@{window:10}(v2)
v6 = ReduceAdd@(v5)
v7 = Output@{name:}(v4)
v8 = Output@{name:}(v6)''')


def test_window_strict():  # test.py:174-186
    b = Builder()
    with b:
        x = Input("a")
        x = Mul(x, x)
        Output(WindowedSum(x, 10))
    with pytest.raises(Exception):
        Function(b.ops, True)


def test_opt_sum():  # test.py:188-207
    b = Builder()
    with b:
        a = Input("a")
        v1 = WindowedSum(a, 10)
        v2 = AddConst(v1, 10)
        v3 = WindowedSum(v2, 10)
        Output(v2)
        Output(v3)
    f = Function(b.ops)
    decompose(f)
    special_optimize(f)
    expect(f, '''v0 = Input@{name:a}()
v1 = WindowedTempOutput@{window:11}(v0)
v2 = FastWindowedSum@{window:10}(v1)
v3 = AddConst@{value:10}(v2)
v4 = WindowedTempOutput@{window:11}(v3)
v5 = FastWindowedSum@{window:10}(v4)
v6 = Output@{name:}(v3)
v7 = Output@{name:}(v5)''')


def test_fold_window_and_identity_print():  # test.py:210-310
    b = Builder()
    with b:
        a, c = Input("a"), Input("b")
        v1 = WindowedCorrelation(a, 10, c)
        v2 = WindowedStddev(a, 10)
        Output(v1, "ou1")
        out2 = Output(v2, "ou2")
    f = Function(b.ops)
    decompose(f)
    expr_fold(f)
    expect(f, '''v0 = Input@{name:a}()
v1 = Input@{name:b}()
v2 = WindowedTempOutput@{window:10}(v0)
v3 = ForeachBackWindow@{window:10}(v2)
v4 = IterValue@(v3,v2) in v3
v5 = ReduceAdd@(v4)
v6 = DivConst@{value:10}(v5)
v7 = WindowedTempOutput@{window:10}(v1)
v8 = ForeachBackWindow@{window:10}(v7)
v9 = IterValue@(v8,v7) in v8
v10 = ReduceAdd@(v9)
v11 = DivConst@{value:10}(v10)
v12 = ForeachBackWindow@{window:10}(v2,v7)
v13 = IterValue@(v12,v2) in v12
v14 = Sub@(v13,v6) in v12
v15 = IterValue@(v12,v7) in v12
v16 = Sub@(v15,v11) in v12
v17 = Mul@(v14,v14) in v12
v18 = Mul@(v16,v16) in v12
v19 = Mul@(v14,v16) in v12
v20 = ReduceAdd@(v19)
v21 = ReduceAdd@(v17)
v22 = Sqrt@(v21)
v23 = ReduceAdd@(v18)
v24 = Sqrt@(v23)
v25 = Mul@(v22,v24)
v26 = Div@(v20,v25)
v27 = DivConst@{value:9}(v21)
v28 = Sqrt@(v27)
v29 = Output@{name:ou1}(v26)
v30 = Output@{name:ou2}(v28)''')
    win = '''WindowedTempOutput@{window:10}(
{p}  input(a)
{p})'''
    def W(p):
        return win.replace("{p}", p)
    txt = out2.to_string(0, True)
    assert txt.startswith('''Output@{name:ou2}(
  Sqrt@(
    DivConst@{value:9}(
      ReduceAdd@(
        Mul@(
          Sub@(
            IterValue@(
              ForeachBackWindow@{window:10}(
                WindowedTempOutput@{window:10}(
                  input(a)
                )
              ),''')
    assert txt.count("input(a)") == 8 and "input(b)" not in txt
    assert txt.endswith('''          )
        )
      )
    )
  )
)''')


def test_div_cmp():  # test.py:312-332
    b = Builder()
    with b:
        a, c = Input("a"), Input("b")
        Output(a / c > 10, "ou1")
        Output(ConstantOp(1) < a / c, "ou1")
    f = Function(b.ops)
    special_optimize(f)
    expect(f, '''v0 = Input@{name:a}()
v1 = Input@{name:b}()
v2 = ConstantOp@{value:10}()
v3 = Mul@(v1,v2)
v4 = GreaterThan@(v0,v3)
v5 = Output@{name:ou1}(v4)
v6 = LessThan@(v1,v0)
v7 = Output@{name:ou1}(v6)''')


def test_duplicate_rank_in():  # test.py:335-353
    b = Builder()
    with b:
        a = Input("a") + 1
        r = Rank(a)
        s = Scale(a)
        Output(r, "ou1")
        Output(s, "ou1")
    f = Function(b.ops)
    decompose_rank(f)
    expect(f, '''v0 = Input@{name:a}()
v1 = AddConst@{value:1}(v0)
v2 = Output@{name:1302190f66243c466}(v1)
v3 = Rank@(v2)
v4 = Scale@(v2)
v5 = Output@{name:ou1}(v3)
v6 = Output@{name:ou1}(v4)''')


def test_duplicate_rank_out():  # test.py:355-367
    b = Builder()
    with b:
        r = Rank(Input("a"))
        Output(r, "ou1")
        Output(r, "ou1")
    f = Function(b.ops)
    move_dup_rank_output(f)
    expect(f, '''v0 = Input@{name:a}()
v1 = Rank@(v0)
v2 = Output@{name:ou1}(v1)
v3 = Output@{name:ou1}(v2)''')


def test_toposort():  # test.py:370-386
    b = Builder()
    with b:
        a, c = Input("a"), Input("b")
        loop = ForeachBackWindow(a, 10)
        b.loop = loop
        it = IterValue(loop, a)
        b.loop = None
        Output(ReduceRank(it, c))
    ops = b.ops
    ops[1], ops[2] = ops[2], ops[1]
    expect(Function(Function.topo_sort_ops(ops)), '''v0 = Input@{name:a}()
v1 = Input@{name:b}()
v2 = ForeachBackWindow@{window:10}(v0)
v3 = IterValue@(v2,v0) in v2
v4 = ReduceRank@(v3,v1)
v5 = Output@{name:}(v4)''')


def test_merge_loops():  # test.py:388-446
    b = Builder()
    with b:
        a, c = Input("a"), Input("b")
        reds = []
        for src, w, red in [(a, 10, ReduceAdd), (a, 20, ReduceMax), (a, 20, ReduceMin),
                            (a, 32, ReduceMin), (c, 32, ReduceMin)]:
            loop = ForeachBackWindow(src, w)
            b.loop = loop
            it = IterValue(loop, src)
            b.loop = None
            reds.append(red(it))
        Output(reds[0] + reds[1] + reds[2] + reds[3] + reds[4], "ou1")
    f = Function(b.ops)
    merge_loops(f)
    expect(f, '''v0 = Input@{name:a}()
v1 = ForeachBackWindow@{window:32,segment_end:20}(v0)
v2 = IterValue@(v1,v0) in v1
v3 = ReduceMin@(v2)
v4 = ForeachBackWindow@{window:20,copy_prev_body:True,segment_end:10}(v0)
v5 = IterValue@(v4,v0) in v4
v6 = ReduceMax@(v5)
v7 = ReduceMin@(v5)
v8 = ForeachBackWindow@{window:10,copy_prev_body:True}(v0)
v9 = IterValue@(v8,v0) in v8
v10 = ReduceAdd@(v9)
v11 = Input@{name:b}()
v12 = ForeachBackWindow@{window:32}(v11)
v13 = IterValue@(v12,v11) in v12
v14 = ReduceMin@(v13)
v15 = Add@(v10,v6)
v16 = Add@(v15,v7)
v17 = Add@(v16,v3)
v18 = Add@(v17,v14)
v19 = Output@{name:ou1}(v18)''')


def test_pow():  # test.py:449-491
    b = Builder()
    with b:
        a = Input("a")
        for e, n in [(5, "out5"), (5.5, "out55"), (0.5, "out05"), (-5, "outm5"),
                     (-5.5, "outm55"), (-0.5, "outm05")]:
            Output(Pow(a, ConstantOp(e)), n)
    f = Function(b.ops)
    decompose(f)
    expect(f, '''v0 = Input@{name:a}()
v1 = Mul@(v0,v0)
v2 = Mul@(v1,v1)
v3 = Mul@(v0,v2)
v4 = Output@{name:out5}(v3)
v5 = Mul@(v0,v0)
v6 = Mul@(v5,v5)
v7 = Mul@(v0,v6)
v8 = Sqrt@(v0)
v9 = Mul@(v8,v7)
v10 = Output@{name:out55}(v9)
v11 = Sqrt@(v0)
v12 = Output@{name:out05}(v11)
v13 = Mul@(v0,v0)
v14 = Mul@(v13,v13)
v15 = Mul@(v0,v14)
v16 = ConstantOp@{value:1}()
v17 = Div@(v16,v15)
v18 = Output@{name:outm5}(v17)
v19 = Mul@(v0,v0)
v20 = Mul@(v19,v19)
v21 = Mul@(v0,v20)
v22 = Sqrt@(v0)
v23 = Mul@(v22,v21)
v24 = ConstantOp@{value:1}()
v25 = Div@(v24,v23)
v26 = Output@{name:outm55}(v25)
v27 = Sqrt@(v0)
v28 = ConstantOp@{value:1}()
v29 = Div@(v28,v27)
v30 = Output@{name:outm05}(v29)''')


# ---- partitioning: tests/test2.py ------------------------------------------

def _opt(f):
    decompose(f)
    expr_fold(f)
    temp_window_elim(f)


def _check_partition(f, graph, funcs):
    mainf, impl = do_partition(f, 2)
    expect(mainf, graph)
    assert len(impl) == len(funcs)
    for fn, txt in zip(impl, funcs):
        expect(fn, txt)


def test_partition1():  # test2.py:33-48
    b = Builder()
    with b:
        Output(AddConst(Mul(Input("a"), Input("b")), 10), "out1")
    f = Function(b.ops)
    _opt(f)
    _check_partition(f, "v0 = GenericPartition@{name:out1}()", ['''name = out1
v0 = Input@{name:a}()
v1 = Input@{name:b}()
v2 = Mul@(v0,v1)
v3 = AddConst@{value:10}(v2)
v4 = Output@{name:out1}(v3)'''])


def test_partition_cyclic():  # test2.py:50-82
    b = Builder()
    with b:
        v1 = Mul(Input("a"), Input("b"))
        v2 = Rank(v1)
        v3 = Mul(v2, v1)
        Output(AddConst(v2, 1), "out1")
        Output(v3, "out2")
    f = Function(b.ops)
    _opt(f)
    _check_partition(f, '''v0 = GenericPartition@{name:18588255c1407f1d3}()
v1 = GenericPartition@{name:14dad88c7b817c3dc}(v0)
v2 = GenericPartition@{name:out1_out2}(v1,v0)''', [
        '''name = 14dad88c7b817c3dc
v0 = Input@{name:18588255c1407f1d3}()
v1 = Rank@(v0)
v2 = Output@{name:14dad88c7b817c3dc}(v1)''',
        '''name = 18588255c1407f1d3
v0 = Input@{name:a}()
v1 = Input@{name:b}()
v2 = Mul@(v0,v1)
v3 = Output@{name:18588255c1407f1d3}(v2)''',
        '''name = out1_out2
v0 = Input@{name:14dad88c7b817c3dc}()
v1 = AddConst@{value:1}(v0)
v2 = Output@{name:out1}(v1)
v3 = Input@{name:18588255c1407f1d3}()
v4 = Mul@(v0,v3)
v5 = Output@{name:out2}(v4)'''])


def test_partition_rank_out():  # test2.py:84-116
    b = Builder()
    with b:
        v1 = Mul(Input("a"), Input("b"))
        v2 = Rank(v1)
        v3 = Mul(v2, v1)
        Output(v2, "out1")
        Output(v3, "out2")
    f = Function(b.ops)
    _opt(f)
    _check_partition(f, '''v0 = GenericPartition@{name:18588255c1407f1d3}()
v1 = GenericPartition@{name:out1}(v0)
v2 = GenericPartition@{name:out2}(v1,v0)''', [
        '''name = out1
v0 = Input@{name:18588255c1407f1d3}()
v1 = Rank@(v0)
v2 = Output@{name:out1}(v1)''',
        '''name = 18588255c1407f1d3
v0 = Input@{name:a}()
v1 = Input@{name:b}()
v2 = Mul@(v0,v1)
v3 = Output@{name:18588255c1407f1d3}(v2)''',
        '''name = out2
v0 = Input@{name:out1}()
v1 = Input@{name:18588255c1407f1d3}()
v2 = Mul@(v0,v1)
v3 = Output@{name:out2}(v2)'''])
