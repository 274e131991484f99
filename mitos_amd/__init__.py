"""MI355X (gfx950) implementation of the Mitos initial-condition hot path.

The compute runs in libmitos_hip.so (hand-written HIP kernels behind the C ABI
of include/mitos_hip.h).  This package holds its ctypes front end
(``mitos_amd.ic``), the build script and the torch autograd wrapper of the
adjoint.
"""
from ._lib import MITOS_FP32, MITOS_FP64, MitosError, load  # noqa: F401

__all__ = ["MITOS_FP32", "MITOS_FP64", "MitosError", "load"]
