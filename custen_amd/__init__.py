"""MI355X-native cuSten: gfx950 stencil kernels behind the cuSten C API.

The native library (custen_amd/lib/libcusten.so) is required; `api` binds it
with ctypes.  Build it with `python -m custen_amd.build`.
"""
from . import api, synth  # noqa: F401

__all__ = ["api", "synth"]
