"""Loads the product library (build/libcuacd.so next to this file).

The library is built in-tree by ``__graft_entry__.build()``; a missing build is
an error, never a silent fallback.
"""
import ctypes
import os

from ._ffi import Api

_HERE = os.path.dirname(os.path.abspath(__file__))
_API = None


class CuacdError(RuntimeError):
    """An error reported by the library; the message matches the reference's Error text."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def library_path():
    return os.path.join(_HERE, "build", "libcuacd.so")


def load_product():
    """The product API (prefix cuacd_); raises when the HIP library has not been built."""
    global _API
    if _API is None:
        path = library_path()
        if not os.path.exists(path):
            raise CuacdError(-1, f"HIP library not built: {path} is missing (run __graft_entry__.build())")
        _API = Api(ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL), "cuacd_", host=True)
    return _API


class AbortError(CuacdError):
    """The call saw its abort flag set (src/decompose.ts:238-240 AbortError)."""


def check(api, code):
    if code == 5:
        raise AbortError(code, api.error())
    if code != 0:
        raise CuacdError(code, api.error())
