"""Shared test set-up: loads the package (its directory name is not a Python
identifier) as module ``cgm_amd`` and the CPU oracle as ``oracle``."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "computational-genomics-mojo_amd")


def load_package(name: str = "cgm_amd"):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # noqa: E402
    return oracle


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def cgm():
    return load_package()


@pytest.fixture(scope="session")
def orc():
    return load_oracle()


def gpu_count() -> int:
    """Number of visible GPUs, via the library (0 when none)."""
    import ctypes
    m = load_package()
    n = ctypes.c_int32(0)
    try:
        m._lib.load().cgm_device_count(ctypes.byref(n))
    except OSError:
        return 0
    return int(n.value)


@pytest.fixture(scope="session")
def dev(cgm):
    if gpu_count() < 1:
        pytest.fail("GPU test run without a visible GPU")
    d = cgm.GenomicsDevice(0)
    info = d.info()
    print(f"\n[device] {info}", flush=True)
    yield d
    d.close()
