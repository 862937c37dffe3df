"""jetshow_amd: MI355X-native per-pixel radiative-transfer solver with the
interface of jetshow's Observation::run (see include/jetshow_amd.h).

Modules:
  abi     ctypes mirror of the C ABI structures and constants
  host    loads libjetshow_amd.so (HIP, gfx950) and wraps the C ABI
  models  descriptor builders for jet models and synthetic workloads
  build   compiles the HIP library in-tree
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
