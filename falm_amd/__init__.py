"""falm_amd — MI355X-native fractional-step LES step (FALM drop-in).

The product is the C-ABI library falm_amd/lib/libfalm_amd.so (HIP kernels
for gfx950, declared in include/falm_amd.h) and the C++ class API in
include/falm/.  This Python package only loads that library (runtime.py;
nothing is loaded until runtime.lib() is called), describes its structs
(_abi.py) and builds case set-ups (cases.py).
"""
from . import _abi  # noqa: F401
from . import cases  # noqa: F401
from . import runtime  # noqa: F401

__all__ = ["_abi", "cases", "runtime"]
