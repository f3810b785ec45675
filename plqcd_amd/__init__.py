"""MI355X-native Wilson hopping matrix for PLQCD (gfx950 HIP kernels behind a C ABI).

The compute path is libplqcd_hip.so; this package only binds it (ctypes) and
mirrors the reference's host interface.  Importing the package does not load
the library; the first call does, and raises if it has not been built.

The solve server runs one HIP stream per engine and per transfer thread; HIP
reads GPU_MAX_HW_QUEUES (hardware queues per process, default 4) when it
initialises, so the package asks for 16 unless the environment says otherwise.
"""
import os as _os

_os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

from .capi import PlqcdError, lib  # noqa: E402,F401

__all__ = ["PlqcdError", "lib"]
