"""vira_amd: MI355X (gfx950) ray-tracing hot path of vira.

Host-side mirror of vira's renderer API over the C-ABI in include/vira_gpu.h.
"""
from ._capi import ViraGPUError, load, lib_path  # noqa: F401

__all__ = ["ViraGPUError", "load", "lib_path"]
