"""MI355X-native backend for KunQuant-style factor pipelines.

Front-end (Op/Stage/ops/passes/Driver) restates the reference compiler; the
device runtime and kernels live in csrc/ and are reached through the C ABI
declared in include/kun_capi.h."""
import os as _os

PACKAGE_DIR = _os.path.dirname(_os.path.abspath(__file__))
