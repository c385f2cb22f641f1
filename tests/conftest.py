"""Shared pytest setup: the `gpu` marker and import paths."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no HIP device is visible")
    return torch


@pytest.fixture(scope="session", autouse=True)
def _autotune_cache_dir(tmp_path_factory):
    """Device-resident models of the C++ layer tune their step on
    initialize() and cache the choice in OPENPFC_GPU_AUTOTUNE_CACHE_DIR
    (default: the working directory); the test session keeps that cache in a
    temporary directory, shared by the programs it starts."""
    if "OPENPFC_GPU_AUTOTUNE_CACHE_DIR" not in os.environ:
        os.environ["OPENPFC_GPU_AUTOTUNE_CACHE_DIR"] = str(tmp_path_factory.mktemp("autotune"))
    yield
