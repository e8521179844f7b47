"""MI355X-native Chorin projection step of NsFemOptim (P2/P1 Navier-Stokes).

The compute path is libnsfem_hip.so (HIP for gfx950) behind the C ABI in
include/nsfem_hip.h; this package holds its Python host mirror.  Importing
the package loads the library and fails if it has not been built."""

from . import _lib

_lib.load()

from . import cu, fem, mesher  # noqa: E402
from ._lib import F32, F64, NsfemError, NsfemInvalidArgument, NsfemNonFinite  # noqa: E402

__all__ = ["cu", "fem", "mesher", "F32", "F64", "NsfemError", "NsfemInvalidArgument", "NsfemNonFinite"]
