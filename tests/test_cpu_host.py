"""The C++ host API's benchmark report (benchmark_frames.csv,
benchmark_summary.json schema 3, benchmark_manifest.json) against the known
answers of the reference's report test: nearest-rank aggregates (avg 12,
p50 10, p95 14), CSV/JSON escaping, manifest counts, FNV-1a64("hello") =
a430d84680aabd0b, RendererPool size validation."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "host", "test_report")


@pytest.fixture(scope="module")
def report_binary(rt):
    if not os.path.exists(BIN):
        from ray_tracing_amd import build as rt_build
        rt_build.build_host(verbose=False)
    return BIN


def test_report_known_answers(report_binary, tmp_path):
    r = subprocess.run([report_binary, str(tmp_path / "out")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK (0 failures)" in r.stdout
    assert (tmp_path / "out" / "benchmark_summary.json").exists()


API_BIN = os.path.join(ROOT, "tests", "host", "test_host_api")


def test_host_api_validation_and_call_shapes(rt):
    """Request validation with the reference's messages as std::runtime_error
    (render_request_validation.cpp:15-66), RendererPool's size check, the
    reference's RendererPool / OptixRenderer signatures (static_asserts) and
    the profile's C ABI round trip; no GPU call."""
    if not os.path.exists(API_BIN):
        from ray_tracing_amd import build as rt_build
        rt_build.build_host(verbose=False)
    r = subprocess.run([API_BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK (0 failures)" in r.stdout
