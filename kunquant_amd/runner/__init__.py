from . import KunRunner
