"""Host-side mirror of the reference's 2D operator interface over the gfx950
C ABI (include/pfh.h). Importing this package loads lib/libpfh.so lazily;
any kernel call without it raises instead of falling back."""
from . import abi, params  # noqa: F401
