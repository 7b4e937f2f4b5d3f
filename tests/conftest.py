import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X / gfx950)")
    config.addinivalue_line("markers", "slow: full-size workloads")


def _carries(path, build_id):
    """True when the built file embeds `build_id` (read as bytes: loading it
    here would keep this copy in the process even after a rebuild)."""
    try:
        with open(path, "rb") as f:
            return build_id.encode() in f.read()
    except OSError:
        return False


def _ensure_built():
    """Builds the oracle when it is missing and the native package when it is
    missing or was built from other sources (the package refuses such a
    build at import)."""
    import oracle

    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    lib = os.path.join(ROOT, "strns2cards_amd", "libstrns2cards_hip.so")
    import glob
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "_s2c_build", os.path.join(ROOT, "strns2cards_amd", "build.py"))
    build = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(build)
    core = glob.glob(os.path.join(ROOT, "strns2cards_amd", "_core*.so"))
    want = build.source_hash()
    if not (_carries(lib, want) and core and _carries(core[0], want)):
        build.build_all()


_ensure_built()


@pytest.fixture(scope="session")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def report(msg: str) -> None:
    """Progress line that names the case before a GPU step."""
    print(msg, flush=True)


# fp32 unit roundoff
U32 = 2.0 ** -24


def elementwise_close(got, want, absum, rtol=1e-5, k=64, what=""):
    """Asserts, for every element, |got - want| <= rtol * |want| + k * u *
    absum, with u the fp32 unit roundoff and absum the element's sum of
    |term| (from the oracle): the relative bar, plus the rounding that a
    different summation order (float atomics) or an exponential differing in
    its last ulp can add to a sum of those terms.  Prints the largest excess
    over the relative bar in units of u * absum."""
    import numpy as np

    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    absum = np.asarray(absum, np.float64)
    assert got.shape == want.shape == absum.shape
    err = np.abs(got - want)
    tol = rtol * np.abs(want) + k * U32 * absum
    excess = (err - rtol * np.abs(want)) / np.maximum(U32 * absum, 1e-300)
    worst = float(excess.max(initial=0.0))
    print(f"{what}: elementwise |err| <= {rtol:g}|want| + {k} u sum|term|; "
          f"largest excess {worst:.3g} u sum|term| over {got.size} elements", flush=True)
    bad = np.argwhere(err > tol)
    assert bad.size == 0, (f"{what}: {len(bad)} elements outside the bound, first {bad[0].tolist()}: "
                           f"got {got[tuple(bad[0])]!r} want {want[tuple(bad[0])]!r} "
                           f"absum {absum[tuple(bad[0])]!r}")
