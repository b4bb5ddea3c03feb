"""Device side of the DG transport path: ctypes binding of librtx.so, the
mesh flattener and the matrix-free operator / Krylov solver built on it."""
from . import _lib  # noqa: F401
from ._lib import RtxError, load  # noqa: F401
