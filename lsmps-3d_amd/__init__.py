"""lsmps3d_amd — MI355X (gfx950) per-step LSMPS particle pipeline.

Python host side: ctypes bindings of the C-ABI in include/lsmps3d_amd.h
(`runtime`), the POD mirrors (`abi`) and deterministic case generators
(`cases`).  The compute path is the HIP library liblsmps3d_amd.so built from
csrc/ by build.py; nothing here computes on the CPU.
"""
from . import abi  # noqa: F401
from . import cases  # noqa: F401

__all__ = ["abi", "cases"]
