"""MI355X-native dielectric VIE FFT-GMRES (drop-in for nec2c's dielectric_fft.h path).

The product is the C-ABI library nec2c_amd/lib/libvie_hip.so (headers under
include/); ``nec2c_amd.vie`` mirrors its interface in Python over ctypes.
"""
from . import vie  # noqa: F401

__all__ = ["vie"]
