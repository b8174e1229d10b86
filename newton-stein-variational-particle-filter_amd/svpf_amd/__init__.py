"""MI355X-native Stein variational particle filter (gfx950 HIP kernels behind a C ABI)."""
from . import config, data  # noqa: F401
from ._lib import LIB_PATH, load  # noqa: F401

__version__ = "0.1.0"
