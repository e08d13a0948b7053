"""MI355X-native BRiDCT: orthonormal float32 2D DCT-II / DCT-III plans on gfx950.

Native pieces (built in-tree by ``make -C bridct_amd``):
  libbridct_amd.so     the C plan API of include/bridct.h and include/bridct_gpu.h
  bridct_numpy.*.so    the NumPy ``Plan`` extension (drop-in for the reference's
                       python/bridct_numpy.c)
``bridct_amd.capi`` is a ctypes binding of the C ABI for device-pointer callers.
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
BRIDCT_VERSION = '0.2.0-dev.7'


def numpy_module():
    """Imports and returns the NumPy extension module ``bridct_numpy``."""
    from . import bridct_numpy   # noqa: F401  (built extension, fails loudly if absent)
    return bridct_numpy
