"""Pass wrapper with KUN_DEBUG tracing (KunQuant/passes/Util.py:11-24)."""
import os


def _env_int(name: str) -> int:
    try:
        return int(os.environ.get(name, "0"))
    except (ValueError, TypeError):
        return 0


debug_mode = _env_int("KUN_DEBUG")
jit_debug_mode = _env_int("KUN_DEBUG_JIT")


def kun_pass(p):
    def run(f, opt={}):
        if debug_mode > 0:
            print("Running pass", p, "num original ops:", len(f.ops))
        p(f, opt)
        if debug_mode > 1:
            print("After the pass: ", f)
        if debug_mode > 0:
            print("new num ops:", len(f.ops))
    return run
