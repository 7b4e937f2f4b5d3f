"""The reference's import name, ``pystrns2cards``, over strns2cards_amd.

Code written against kenji-tojo/strns2cards imports ``pystrns2cards as s2c``
(1_cluster.py:3, 2_texture.py:3, 3_fitting.py:4, texture_utils.py) and uses
the names its package re-exports (pystrns2cards/__init__.py:4-204) and the
submodules ``optimizer`` and ``_core``.  This package gives the same names,
all implemented by strns2cards_amd (HIP kernels behind the C ABI in
include/strns2cards_hip.h).  ``import pystrns2cards.dda`` and the like resolve
to the strns2cards_amd modules.

Not provided, as in strns2cards_amd: the sklearn clustering helpers
``cluster_points_kmeans`` / ``compute_strand_features`` (pystrns2cards/
cluster.py) and the camera module (pystrns2cards/camera.py).
"""
import importlib as _importlib
import sys as _sys

from strns2cards_amd import *  # noqa: F401,F403
from strns2cards_amd import _core, optimizer, print_version  # noqa: F401

# As in the reference, the package attribute ``rasterize`` is the function
# (the star import shadows the submodule); the submodules themselves are
# reachable by their dotted import paths.
for _name in ("_core", "antialias", "bitset_utils", "curve_utils", "dda", "optimizer", "rasterize"):
    _sys.modules[f"{__name__}.{_name}"] = _importlib.import_module(f"strns2cards_amd.{_name}")
del _name
