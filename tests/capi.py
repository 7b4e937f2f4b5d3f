"""ctypes access to the C ABI declared in include/strns2cards_hip.h.

The header is parsed for its ``s2c_*`` declarations, so the tests call
exactly the entry points a foreign caller would bind.  Tensors are passed as
device pointers; ``stream`` is torch's current stream on the tensor's device.
"""
from __future__ import annotations

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "strns2cards_hip.h")
LIB = os.path.join(ROOT, "strns2cards_amd", "libstrns2cards_hip.so")

_DECL = re.compile(r"^\s*(int|const char\*)\s+(s2c_\w+)\s*\(([^;]*?)\)\s*;", re.M | re.S)


def declarations() -> dict[str, tuple[str, list[str]]]:
    """name -> (return type, [parameter declarations])."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for ret, name, params in _DECL.findall(text):
        ps = [p.strip() for p in params.replace("\n", " ").split(",") if p.strip()]
        if ps == ["void"]:
            ps = []
        out[name] = (ret, ps)
    return out


def _ctype(param: str):
    if "*" in param:
        return ctypes.c_void_p
    if param.startswith("float"):
        return ctypes.c_float
    if param.startswith("double"):
        return ctypes.c_double
    if param.startswith("int"):
        return ctypes.c_int
    raise ValueError(f"unsupported parameter {param!r}")


_lib = None


def lib():
    global _lib
    if _lib is None:
        import torch  # noqa: F401  (the HIP runtime torch links is the one used)
        l = ctypes.CDLL(LIB)
        for name, (ret, params) in declarations().items():
            fn = getattr(l, name)
            fn.restype = ctypes.c_int if ret == "int" else ctypes.c_char_p
            fn.argtypes = [_ctype(p) for p in params]
        _lib = l
    return _lib


class CapiError(RuntimeError):
    def __init__(self, name: str, code: int, msg: str):
        super().__init__(f"{name} -> {code}: {msg}")
        self.code = code


def call(name: str, *args, stream=None) -> int:
    """Calls ``name`` with tensors / ints / floats; the trailing ``stream``
    argument is supplied automatically.  Raises CapiError on failure."""
    import torch

    conv = []
    for a in args:
        if isinstance(a, torch.Tensor):
            conv.append(ctypes.c_void_p(a.data_ptr()))
        elif isinstance(a, ctypes._SimpleCData) or isinstance(a, ctypes.Array):
            conv.append(a)
        elif a is None:
            conv.append(ctypes.c_void_p(0))
        else:
            conv.append(a)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    rc = getattr(lib(), name)(*conv, ctypes.c_void_p(stream))
    if rc != 0:
        raise CapiError(name, rc, lib().s2c_last_error().decode())
    return rc


def call_count(name: str, *args) -> int:
    """For entry points that return a count through a trailing int*."""
    n = ctypes.c_int(-7)
    call(name, *args, ctypes.byref(n))
    return n.value
