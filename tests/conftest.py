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


def _ensure_built():
    """Builds the oracle and the native package only when nothing is built."""
    import oracle

    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    lib = os.path.join(ROOT, "strns2cards_amd", "libstrns2cards_hip.so")
    import glob

    core = glob.glob(os.path.join(ROOT, "strns2cards_amd", "_core*.so"))
    if not os.path.exists(lib) or not core:
        import importlib.util

        spec = importlib.util.spec_from_file_location(
            "_s2c_build", os.path.join(ROOT, "strns2cards_amd", "build.py"))
        build = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(build)
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
