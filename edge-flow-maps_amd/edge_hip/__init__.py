"""edge_hip: host side of the MI355X EDGE flow-map time step.

libedge_hip.so (built from ../csrc) holds every kernel; this package binds it
with ctypes, provides the Taichi-field surface the reference's step loop uses
(Field, ti) and the Hermite-table cache.  The drop-in modules that mirror the
reference's API live beside it in ../sim and ../util.
"""
from . import _lib
from .field import Field, ti, touch, as_ptr, stream_handle  # noqa: F401

__all__ = ["Field", "ti", "touch", "as_ptr", "stream_handle", "lib", "available"]


def lib():
    return _lib.lib()


def available():
    """True when the HIP library loads and a GPU is visible."""
    import torch
    try:
        _lib.lib()
    except (RuntimeError, OSError):
        return False
    return torch.cuda.is_available()
