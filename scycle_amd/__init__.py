"""scycle_amd: MI355X (gfx950) implementation of SCycle's quasi-dynamic and
fully dynamic step hot path behind a C ABI (include/scycle_hip.h).

The compute lives in libscycle_hip.so (HIP kernels + host C++ driver); this
package only binds it (ctypes) and mirrors the reference's SbpOps / fault /
mediator interfaces for Python callers and tests.
"""
from ._lib import LIB_PATH, ScError, load  # noqa: F401
from .device import DeviceArray  # noqa: F401
from .sbp_ops import SbpOps, DIRICHLET, NEUMANN  # noqa: F401
from .qd import PowerLawMediator, QdMediator  # noqa: F401
from .fd import FdMediator  # noqa: F401
from .qdfd import QdFdMediator  # noqa: F401

__all__ = ["LIB_PATH", "ScError", "load", "DeviceArray", "SbpOps", "DIRICHLET", "NEUMANN",
           "QdMediator", "PowerLawMediator", "FdMediator", "QdFdMediator"]
