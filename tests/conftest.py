import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


def load_package():
    """Import the acc-gp_amd package (its directory name is not an identifier)."""
    if "acc_gp_amd" in sys.modules:
        return sys.modules["acc_gp_amd"]
    pkg_dir = os.path.join(ROOT, "acc-gp_amd")
    spec = importlib.util.spec_from_file_location("acc_gp_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["acc_gp_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def gp():
    return load_package()


@pytest.fixture(scope="session")
def og():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ctx(gp):
    c = gp.Context(0)
    yield c
    c.close()
