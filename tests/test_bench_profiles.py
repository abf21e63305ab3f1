"""The bench line's profile lookups (CPU): the committed rocprof summaries hold every kernel instantiation the
bench attributes a stage to, the per-dispatch means combine by dispatch count, and a multi-GPU request is
refused."""

import importlib.util
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(REPO, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("stage", ["trace", "escape", "shade"])
def test_committed_profiles_cover_the_stage_kernels(bench, stage):
    issue = bench.issue_from_profiles(3, bench.KERNEL_SYMBOLS[stage])
    assert issue is not None, "issue counters missing for %s" % bench.KERNEL_SYMBOLS[stage]
    assert issue["SQ_INSTS_VALU"] > 0
    assert 0.0 < issue["active_lane_fraction"] <= 1.0
    traffic = bench.traffic_from_profiles(3, bench.KERNEL_SYMBOLS[stage])
    assert traffic is not None and traffic > 0


def test_stage_means_weight_by_dispatches(bench):
    a = {"SQ_INSTS_VALU": 100.0, "hbm_bytes_per_dispatch": 10.0, "dispatches": 3}
    b = {"SQ_INSTS_VALU": 20.0, "hbm_bytes_per_dispatch": 50.0, "dispatches": 1}
    r = bench._per_dispatch([a, b], ("SQ_INSTS_VALU", "hbm_bytes_per_dispatch"))
    assert r["SQ_INSTS_VALU"] == pytest.approx(80.0)
    assert r["hbm_bytes_per_dispatch"] == pytest.approx(20.0)
    # a stage with an instantiation the summary does not hold has no figure rather than a partial one
    assert bench._per_dispatch([a, None], ("SQ_INSTS_VALU",)) is None


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode != 0
    assert "gpus" in p.stderr.lower()
