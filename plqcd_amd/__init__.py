"""MI355X-native Wilson hopping matrix for PLQCD (gfx950 HIP kernels behind a C ABI).

The compute path is libplqcd_hip.so; this package only binds it (ctypes) and
mirrors the reference's host interface.  Importing the package does not load
the library; the first call does, and raises if it has not been built.
"""
from .capi import PlqcdError, lib  # noqa: F401

__all__ = ["PlqcdError", "lib"]
