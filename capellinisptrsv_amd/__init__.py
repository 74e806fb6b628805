"""MI355X-native (gfx950) sync-free sparse triangular solve.

C ABI: include/capellini_sptrsv.h, implemented by libcapellini_sptrsv.so
(HIP kernels + host code).  Python layers:
  capellinisptrsv_amd.api        numpy mirror of the reference's host interface
  capellinisptrsv_amd.torch_ops  device-resident plans, batches, autograd
"""
import importlib

from ._native import lib, NativeLibraryError, CapError  # noqa: F401

__version__ = "0.1.0"

_TORCH_NAMES = ("Plan", "Batch", "SpTRSV", "sptrsv", "TriangularSolve")


def __getattr__(name):
    if name in ("api", "torch_ops"):
        return importlib.import_module(f"{__name__}.{name}")
    if name in _TORCH_NAMES:
        return getattr(importlib.import_module(f"{__name__}.torch_ops"), name)
    raise AttributeError(name)
