"""The bench line's `roofline.traffic` comes from the committed rocprof summaries: every kernel the
bench can name as dominant has a traffic line there, and each summary records the command that made
it.  CPU only (reads text files)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("key", sorted(bench.TRAFFIC_PROFILES))
def test_every_traffic_profile_entry_resolves(key):
    value, path = bench.committed_traffic(*key)
    assert value is not None and value > 0, key
    text = open(os.path.join(ROOT, path)).read()
    assert "rocprofv3 --pmc FETCH_SIZE" in text and "rocprofv3 --pmc WRITE_SIZE" in text


def test_unprofiled_kernels_report_none():
    assert bench.committed_traffic("vortex2d", "face", "k_update") == (None, None)
