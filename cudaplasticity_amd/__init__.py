"""MI355X-native (gfx950) plasticity hot path: the incremental elasto-plastic
FEM loop of SATOH-SD/CudaPlasticity behind its Mesh / Material /
PlasticitySolver interface.  The compute lives in libcp_hip.so (HIP kernels +
C-ABI, include/cp_abi.h); this package holds its ctypes binding and a Python
mirror of the reference's C++ classes."""
from . import abi  # noqa: F401
