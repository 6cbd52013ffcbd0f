"""MI355X-native residual + dual-time LU-SGS hot path of ricklee209/CGLI_Fu_Flow.

Compute runs only in the gfx950 HIP library ``_lib/libcgli.so`` (C ABI in
``include/cgli.h``).  ``native`` binds it through ctypes (contexts, fields,
the hot-path routines, the channel driver and the learned guess's network),
``build`` compiles it, and ``reports`` writes the larger-configuration
reports.  The reference's own routine names and parameter lists are the C++
header ``include/cgli_reference_api.hpp``.
"""
from . import native  # noqa: F401

__all__ = ["native"]
