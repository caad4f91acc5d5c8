from .CompOp import *
from .ElewiseOp import *
from .ReduceOp import *
from .MiscOp import *
from .CompOp import (WindowedCompositiveOp, WindowedReduce, _WindowedMinMaxBase,
                     _is_fast_stat, _decide_use_skip_list)
