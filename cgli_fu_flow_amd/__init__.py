"""MI355X-native residual + dual-time LU-SGS hot path of ricklee209/CGLI_Fu_Flow.

Compute runs only in the gfx950 HIP library ``_lib/libcgli.so`` (C ABI in
``include/cgli.h``); this package binds it (``native``), mirrors the
reference's routine names (``reference_api``) and drives the channel solver
(``channel``).
"""
from . import native  # noqa: F401

__all__ = ["native"]
