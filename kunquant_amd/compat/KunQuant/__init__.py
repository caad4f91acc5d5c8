"""Drop-in alias of the reference's package name: with the directory that
holds this package (kunquant_amd/compat) ahead of any other `KunQuant` on
sys.path, code written against the reference -- `from KunQuant.Op import
*`, `from KunQuant.jit import cfake`, `from KunQuant.runner import
KunRunner as kr`, `from KunQuant.predefined.Alpha101 import ...` -- runs
unchanged on the MI355X backend. Every submodule name resolves to the
kunquant_amd module of the same name (the same module objects, so types
and state are shared)."""
import importlib
import os
import sys

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if _PKG_ROOT not in sys.path:
    sys.path.append(_PKG_ROOT)

_SUBMODULES = ("Op", "Stage", "Driver", "ops", "ops.CompOp", "ops.ElewiseOp", "ops.MiscOp", "ops.ReduceOp",
               "passes", "jit", "jit.cfake", "runner", "runner.KunRunner", "predefined", "predefined.Alpha101",
               "predefined.Alpha158")

for _name in _SUBMODULES:
    _mod = importlib.import_module("kunquant_amd." + _name)
    sys.modules[__name__ + "." + _name] = _mod
    _parent, _, _leaf = (__name__ + "." + _name).rpartition(".")
    setattr(sys.modules[_parent], _leaf, _mod)
