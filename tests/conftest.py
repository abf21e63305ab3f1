import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(REPO, "dayflower-path-tracer_amd")
for p in (PKG_DIR, REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def _ensure_built():
    """Builds a library when it is missing or was built from other sources than the tree holds now (each build
    stamps the SHA-256 of its sources; a built .so travels with the tree, so an edited source must never be
    tested through the binary of an older one)."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("dayflower_build", os.path.join(PKG_DIR, "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    if not mod.is_current():
        mod.build(jobs=4)
    from oracle import oracle_ffi

    if not oracle_ffi.is_current():
        oracle_ffi.build()


_ensure_built()

_scene_cache = {}


@pytest.fixture(scope="session")
def scenes():
    """Workload scenes built once per session: scenes(config_id)."""
    import dayflower_amd as dfa

    def get(cid):
        if cid not in _scene_cache:
            _scene_cache[cid] = dfa.build_config(cid)
        return _scene_cache[cid]

    return get


@pytest.fixture(scope="session")
def gpu():
    import dayflower_amd as dfa

    if dfa.device_count() < 1:
        pytest.fail("no HIP device visible to a gpu-marked test")
    return 0
