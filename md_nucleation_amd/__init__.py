"""MI355X-native (HIP, gfx950) per-step force and integration path of the
effective-domain nucleation MD method of tamadate/MD_nucleation.

The compute path is libmdn.so (C ABI, include/mdn.h); this package holds its
sources (csrc/), its build recipe (build.py), a ctypes binding (_lib.py) and
input-file writers (configs.py).
"""
from ._lib import (FP64, MIXED, U_LABELS, Engine, MdnError, State, SysDef, lib)  # noqa: F401
