"""MI355X (gfx950) implementation of the mammos-mumag per-field-step hot path.

Host code is Python; the arithmetic runs in hand-written HIP kernels reached through
the C ABI of ``_lib/libmumag_hip.so`` (declared in ``include/mumag_hip.h``).
"""
__version__ = "0.1.0"

from .model import Model, boundary_setup  # noqa: E402,F401
