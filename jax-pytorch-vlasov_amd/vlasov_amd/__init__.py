"""MI355X-native semi-Lagrangian Vlasov-Poisson solver.

``vlasov_amd.vlasov1D1V`` and ``vlasov_amd.vlasov2D2V`` mirror the reference
PyTorch modules of yasahi-hpc/JAX-PyTorch-Vlasov; the computation runs in
hand-written gfx950 HIP kernels (``libvlasov_amd.so``, C ABI in
``include/vlasov_amd.h``) loaded through ctypes.
"""
from .build import LIB_PATH  # noqa: F401

__version__ = "0.1.0"
