import importlib.util
import pathlib
import sys

import pytest

REPO = pathlib.Path(__file__).resolve().parents[1]
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def _load_package():
    if "gs_native_amd" in sys.modules:
        return sys.modules["gs_native_amd"]
    package_dir = REPO / "gs-native_amd"
    spec = importlib.util.spec_from_file_location("gs_native_amd", package_dir / "__init__.py",
                                                  submodule_search_locations=[str(package_dir)])
    module = importlib.util.module_from_spec(spec)
    sys.modules["gs_native_amd"] = module
    spec.loader.exec_module(module)
    return module


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gsn():
    return _load_package()


@pytest.fixture(scope="session")
def scenes():
    _load_package()
    import gs_native_amd.scenes as module
    return module


@pytest.fixture(scope="session")
def oracle():
    import oracle as module
    return module


@pytest.fixture(scope="session")
def context(gsn):
    if gsn.device_count() < 1:
        pytest.fail("a GPU test ran without a visible GPU")
    print(f"device: {gsn.device_name()}", flush=True)
    ctx = gsn.Context()
    yield ctx
    ctx.close()
