import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_engine():
    """Import the liggghts-inl_amd package (hyphenated directory) by path."""
    name = "liggghts_inl_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(ROOT, "liggghts-inl_amd", "__init__.py"),
        submodule_search_locations=[os.path.join(ROOT, "liggghts-inl_amd")])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def eng():
    return load_engine()


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    return oracle


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False
