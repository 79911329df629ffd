"""Shared fixtures.  `gpu`-marked tests need an MI355X; everything else runs
on the CPU (oracle pins, host logic, library loading and error paths)."""
import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run with -m gpu)")


def _load_package():
    if "batch_cc_amd" in sys.modules:
        return sys.modules["batch_cc_amd"]
    pkg_dir = os.path.join(REPO, "batch-cc_amd")
    spec = importlib.util.spec_from_file_location("batch_cc_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["batch_cc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


bcc_pkg = _load_package()


@pytest.fixture(scope="session")
def bcc():
    return bcc_pkg


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib.Oracle()


@pytest.fixture(scope="session")
def gpu(bcc):
    n = bcc.device_count()
    if n < 1:
        pytest.fail("no HIP device visible: gpu tests must run on a GPU box")
    return n
