"""Shared test setup: the `gpu` marker, import paths and the two modules under test.

`hydro` is the product package (scalable-fpga-hydrodynamics_amd, ctypes over the HIP
library); `oracle` is the CPU restatement used as the checker.
"""
import importlib
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

PKG = "scalable-fpga-hydrodynamics_amd"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU oracle check (opt in with -m slow)")


def load_hydro():
    return importlib.import_module(PKG)


def load_oracle():
    from oracle import oracle as O  # noqa: WPS433

    return O


@pytest.fixture(scope="session")
def oracle():
    return load_oracle()


@pytest.fixture(scope="session")
def hydro():
    return load_hydro()


@pytest.fixture(scope="session")
def gpu(hydro):
    """The product package with a visible GPU; fails (never skips) when the library
    is missing, so a GPU run cannot pass on a silent fallback."""
    import torch

    assert torch.cuda.is_available(), "gpu test run without a visible GPU"
    hydro.lib()
    torch.cuda.init()
    return hydro
