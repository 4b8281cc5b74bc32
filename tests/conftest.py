"""Shared test setup: the `gpu` marker, package loading and GPU/oracle fixtures."""
from __future__ import annotations

import importlib.util
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
TESTS = ROOT / "tests"
GOLDEN = TESTS / "golden"
for p in (str(TESTS), str(ROOT), str(ROOT / "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def load_package():
    if "cr_amd" not in sys.modules:
        pkg = ROOT / "cpu-renderer_amd"
        spec = importlib.util.spec_from_file_location("cr_amd", pkg / "__init__.py",
                                                      submodule_search_locations=[str(pkg)])
        mod = importlib.util.module_from_spec(spec)
        sys.modules["cr_amd"] = mod
        spec.loader.exec_module(mod)
    return sys.modules["cr_amd"]


load_package()
from cr_amd import abi, scenes  # noqa: E402,F401


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 device and the built libcrgpu.so")


@pytest.fixture(scope="session")
def gpu():
    """A crgpu context on cuda:0 (HIP device 0).  Fails loudly if the library or device is missing."""
    from cr_amd import crgpu
    with crgpu.Context(0) as ctx:
        yield ctx


@pytest.fixture(scope="session")
def oracle():
    import probe
    return probe.Probe("orc")


def load_golden(name: str):
    import numpy as np
    return np.load(GOLDEN / name, allow_pickle=False)
