"""strns2cards_amd — the pystrns2cards hot path on AMD MI355X (gfx950).

The public API mirrors the reference package ``pystrns2cards``
(kenji-tojo/strns2cards, pystrns2cards/__init__.py:1-14): every function of
``dda``, ``rasterize``, ``antialias``, ``bitset_utils`` and ``curve_utils`` is
re-exported at the top level.  All GPU work runs in hand-written HIP kernels
in ``libstrns2cards_hip.so`` (C ABI: include/strns2cards_hip.h), reached
through the ``_core`` extension; there is no CPU fallback, and importing the
package fails if the native libraries are not built
(``python strns2cards_amd/build.py``).
"""
import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))
LIBRARY_PATH = _os.path.join(_HERE, "libstrns2cards_hip.so")

try:
    import torch as _torch  # noqa: F401  (loads the HIP runtime torch links)
    from . import _core
except ImportError as _e:  # pragma: no cover - exercised only on broken installs
    raise ImportError(
        "strns2cards_amd: native extension not built or not loadable "
        f"({_e}); run `python strns2cards_amd/build.py`") from _e


def _check_build() -> None:
    """Refuses libraries built from other sources than those beside them
    (when the sources are present): a stale build would run silently."""
    import ctypes

    from .build import CSRC, source_hash

    if not _os.path.isdir(CSRC):
        return
    want = source_hash()
    lib = ctypes.CDLL(LIBRARY_PATH)
    lib.s2c_build_id.restype = ctypes.c_char_p
    got = {"libstrns2cards_hip.so": lib.s2c_build_id().decode(), "_core": _core.build_id}
    stale = {k: v for k, v in got.items() if v != want}
    if stale:
        raise ImportError(f"strns2cards_amd: {', '.join(stale)} built from other sources "
                          f"(build id {stale}, sources {want}); run "
                          "`python strns2cards_amd/build.py`")


_check_build()

from .adjoint import *  # noqa: E402,F401,F403
from .antialias import *  # noqa: E402,F401,F403
from .bitset_utils import *  # noqa: E402,F401,F403
from .curve_utils import *  # noqa: E402,F401,F403
from .dda import *  # noqa: E402,F401,F403
from .rasterize import *  # noqa: E402,F401,F403
from .strand_io import *  # noqa: E402,F401,F403
from . import optimizer  # noqa: E402,F401


def print_version() -> None:
    """Prints the native library version (reference: __init__.py:16-17)."""
    _core.print_version()
