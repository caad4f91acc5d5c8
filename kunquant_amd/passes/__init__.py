from .Decompose import decompose, decompose_rank, move_dup_rank_output
from .ExprFold import expr_fold
from .TempWindowElim import temp_window_elim
from .SpecialOpt import special_optimize
from .Partitioner import do_partition
from .InferWindow import infer_window, infer_input_window
from .MergeLoops import merge_loops
