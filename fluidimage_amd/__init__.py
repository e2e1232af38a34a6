"""fluidimage_amd: the PIV correlation path of fluidimage on MI355X (gfx950).

Compute runs in libpiv_hip.so (hand-written HIP, C ABI in include/piv_hip.h)
called through ctypes; torch provides device memory and streams.
"""

__version__ = "0.1.0"
