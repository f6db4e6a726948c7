"""The render_realtime CLI against the reference's own CLI checks: the
invocations and assertions of cmake/VerifyRenderRealtimeProfiling.cmake and
cmake/VerifyRenderRealtimeFourCameraProfiling.cmake (registered as CTests in
CMakeLists.txt:981-1018), restated as pytest: stdout fields and patterns,
the schema-3 summary JSON, the per-camera record order, the FNV-1a64
artifact manifest, PNG images (or none with --skip-image-write), denoise
timing per profile and the luminance spread between the four cameras of
final_room."""
import json
import os
import re
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_cli(rt, out, *args):
    exe = os.path.join(os.path.dirname(rt.__file__), "render_realtime")
    if not os.path.exists(exe):
        pytest.fail("render_realtime is not built")
    os.makedirs(out, exist_ok=True)
    r = subprocess.run([exe, *args, "--output-dir", str(out)], capture_output=True, text=True,
                       timeout=240)
    say(r.stdout[-400:])
    assert r.returncode == 0, r.stderr
    return r.stdout


def pngs(out):
    return [f for f in os.listdir(out) if f.endswith(".png")]


def test_profiling_cli_single_camera(rt, tmp_path):
    """VerifyRenderRealtimeProfiling.cmake:10-138 (default scene, 1 camera)."""
    say("[gpu] render_realtime single camera, realtime profile")
    out = tmp_path / "single"
    so = run_cli(rt, out, "--camera-count", "1", "--frames", "2", "--warmup-frames", "1",
                 "--seed", "7", "--profile", "realtime")
    for pat in ("download_ms=", "pipeline_ms=", "host_overhead_ms=", "warmup=0 sample_stream=7",
                "summary profile=realtime warmup_frames=1 frames=2 seed=7", "p99_frame_ms=",
                "peak_gpu_memory_mib="):
        assert pat in so, pat
    assert "host_overhead_ms=-" not in so
    images = pngs(out)
    assert images, "the default run writes PNG images"
    with open(out / images[0], "rb") as f:
        assert f.read(8) == b"\x89PNG\r\n\x1a\n"
    csv = open(out / "benchmark_frames.csv").read()
    assert ("frame_index,sample_stream,camera_count,profile,width,height,samples_per_pixel,"
            "max_bounces,denoise_enabled") in csv
    assert "0,8,1,realtime,640,480" in csv
    js = open(out / "benchmark_summary.json").read()
    for field in ('"download_ms"', '"pipeline_ms"', '"render_work_ms"', '"frames"', '"per-camera"',
                  '"schema_version": 3', '"warmup_frames": 1', '"random_seed": 7', '"sample_stream": 8',
                  '"p99"', '"provenance"', '"source_revision"', '"source_scope"', '"source_state_sha256"',
                  '"environment"', '"gpu_name"', '"nvidia_driver_version"',
                  '"cuda_driver_api_version_text"', '"gpu_memory"', '"peak_used_bytes"',
                  '"peak_delta_bytes"', '"optix_version_text"'):
        assert field in js, field
    assert re.search(r'"build_configuration" *: *"[^"]+"', js)
    assert re.search(r'"source_revision" *: *"[^"]+"', js)
    assert re.search(r'"source_state_sha256" *: *"[0-9a-f]+"', js)
    man = open(out / "benchmark_manifest.json").read()
    for field in ('"report_schema_version": 3', '"measured_frame_count": 2',
                  '"per_camera_record_count": 2', '"filename": "benchmark_frames.csv"',
                  '"filename": "benchmark_summary.json"', '"digest_algorithm": "fnv1a64"'):
        assert field in man, field
    # the skip-write variant (VerifyRenderRealtimeProfiling.cmake:141-170)
    skip = tmp_path / "skip"
    so = run_cli(rt, skip, "--camera-count", "1", "--frames", "2", "--warmup-frames", "1",
                 "--seed", "7", "--profile", "realtime", "--skip-image-write")
    assert "image_write_ms=0.000" in so
    assert not pngs(skip)


def four_camera(rt, out, scene, profile, frames, expect_denoise, min_spread_milli=0):
    """VerifyRenderRealtimeFourCameraProfiling.cmake:25-238."""
    so = run_cli(rt, out, "--scene", scene, "--camera-count", "4", "--frames", str(frames),
                 "--warmup-frames", "1", "--seed", "11", "--profile", profile, "--skip-image-write")
    assert "denoise_ms=" in so and "pipeline_ms=" in so and "p99_frame_ms=" in so
    assert "host_overhead_ms=-" not in so
    assert "warmup=0 sample_stream=11 cameras=4" in so
    if not expect_denoise:
        lines = re.findall(r"frame=[0-9]+ sample_stream=[0-9]+ cameras=4 [^\r\n]*denoise_ms=0\.000", so)
        assert len(lines) == frames
    js = open(out / "benchmark_summary.json").read()
    for field in ('"pipeline_ms"', '"render_work_ms"', '"schema_version": 3', '"warmup_frames": 1',
                  '"random_seed": 11', '"sample_stream": 12', '"p99"', '"provenance"',
                  '"source_scope"', '"source_state_sha256"', '"environment"',
                  '"nvidia_driver_version"', '"gpu_memory"', '"peak_used_bytes"',
                  '"peak_delta_bytes"'):
        assert field in js, field
    man = open(out / "benchmark_manifest.json").read()
    assert '"filename": "benchmark_frames.csv"' in man and '"digest_algorithm": "fnv1a64"' in man
    flat = re.sub(r"[\r\n\t ]+", " ", js)
    meta = re.search(r'"metadata" *: *\{[^{}]*\}', flat)
    assert meta, "metadata object"
    meta = meta.group(0)
    assert re.search(r'"camera_count" *: *([0-9]+)', meta).group(1) == "4"
    assert re.search(r'"profile" *: *"([^"]+)"', meta).group(1) == profile
    assert re.search(r'"scene" *: *"([^"]+)"', meta).group(1) == scene
    assert re.search(r'"denoise_enabled" *: *(true|false)', meta).group(1) == \
        ("true" if expect_denoise else "false")
    recs = re.findall(r'\{[^{}]*"frame_index" *: *[0-9]+[^{}]*"camera_index" *: *[0-9]+[^{}]*\}', flat)
    assert len(recs) == frames * 4
    for i, rec in enumerate(recs):
        assert int(re.search(r'"frame_index" *: *([0-9]+)', rec).group(1)) == i // 4
        assert int(re.search(r'"camera_index" *: *([0-9]+)', rec).group(1)) == i % 4
    if min_spread_milli > 0:
        milli = []
        for rec in recs:
            v = re.search(r'"average_luminance" *: *([-0-9.eE+]+)', rec).group(1)
            m = re.fullmatch(r"([0-9]+)(\.([0-9]+))?", v)
            assert m, f"unsupported average_luminance format: {v}"
            milli.append(int(m.group(1)) * 1000 + int(((m.group(3) or "") + "000")[:3]))
        say(f"  luminance milli {milli}")
        assert max(milli) - min(milli) >= min_spread_milli
    return so


def test_four_camera_profiling_cli(rt, tmp_path):
    say("[gpu] render_realtime smoke, 4 cameras, realtime profile")
    four_camera(rt, tmp_path / "rt", "smoke", "realtime", 2, True)


def test_four_camera_quality_cli(rt, tmp_path):
    say("[gpu] render_realtime smoke, 4 cameras, quality profile")
    four_camera(rt, tmp_path / "q", "smoke", "quality", 2, False)


def test_final_room_quality_cli(rt, tmp_path):
    say("[gpu] render_realtime final_room, 4 cameras, quality profile")
    four_camera(rt, tmp_path / "fr", "final_room", "quality", 1, False, min_spread_milli=10)


def test_cli_renders_yaml_scene(rt, tmp_path):
    """render_realtime --scene <id> --scene-dir <dir>: a YAML scene file
    flattened by the package's ingest and rendered through the pool."""
    say("[gpu] render_realtime on the cornell_box YAML scene")
    scenes = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")
    out = tmp_path / "yaml"
    run_cli(rt, out, "--scene", "cornell_box", "--scene-dir", scenes, "--profile", "scene",
            "--camera-count", "1", "--frames", "2", "--warmup-frames", "1")
    summary = json.load(open(out / "benchmark_summary.json"))
    assert summary["metadata"]["scene"] == "cornell_box"
    assert len(pngs(out)) >= 1
