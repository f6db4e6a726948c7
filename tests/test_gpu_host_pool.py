"""The reference-shaped host API on the GPU, and the per-slot random-number
identity it is built on.

- tests/host/test_pool_gpu.cpp restates the checks of the reference's
  tests/test_renderer_pool.cpp and tests/test_optix_prepared_scene.cpp
  (pool vs lone renderer per camera, device frames, diagnostics, request
  validation, prepared-scene lifecycle, sample streams) through the C++
  RendererPool / Renderer;
- rt_render_frame_streams: a camera rendered alone in slot 0 under camera
  key k and sample stream s equals camera k of a whole-rig frame whose slot
  k draws stream s, bit for bit, over two frames with ReSTIR DI temporal and
  spatial reuse, reservoirs included; keys outside [0, 4) are refused.
"""
import os
import subprocess
import sys

import ctypes as C
import numpy as np
import pytest

from test_gpu_restir import scaled_rig

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
POOL_BIN = os.path.join(ROOT, "tests", "host", "test_pool_gpu")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_pool_and_renderer_program(rt):
    say("[gpu] tests/host/test_pool_gpu")
    assert os.path.exists(POOL_BIN), "build the host targets first (ray_tracing_amd.build.build_host)"
    r = subprocess.run([POOL_BIN], capture_output=True, text=True, timeout=300)
    say(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK (0 failures)" in r.stdout


def one_camera_rig(rt, rig, k):
    one = rt.CameraRig()
    one.active_count = 1
    C.memmove(C.byref(one.cameras[0]), C.byref(rig.cameras[k]), C.sizeof(rig.cameras[k]))
    return one


def test_slot_keys_and_streams_reproduce_whole_rig_frames(rt, oracle, device):
    g = rt.Generated("instanced_mesh_restir", 0)
    sc = rt.Scene(device, g.desc)
    sc.build_bvh()
    prof = g.profile
    prof.samples_per_pixel, prof.enable_denoise = 1, 1
    whole = rt.Renderer(device, sc)
    frames = []
    for f in range(2):
        rig = scaled_rig(oracle, g, f)
        streams = [5 + c + f for c in range(4)]
        whole.render(rig, prof, streams=streams)
        frames.append([(whole.download(c), whole.reservoirs(c)) for c in range(rig.active_count)])
    for k in range(4):
        say(f"[gpu] camera {k} alone in slot 0 under key {k}")
        lone = rt.Renderer(device, sc)
        for f in range(2):
            rig = scaled_rig(oracle, g, f)
            lone.render(one_camera_rig(rt, rig, k), prof, streams=[5 + k + f] * 4, keys=[k, 1, 2, 3])
            got, res = lone.download(0), lone.reservoirs(0)
            ref, ref_res = frames[f][k]
            for key in ("beauty", "normal", "albedo", "depth"):
                assert got[key].tobytes() == ref[key].tobytes(), f"frame {f} camera {k} {key}"
            assert res.tobytes() == ref_res.tobytes(), f"frame {f} camera {k} reservoirs"
    with pytest.raises(rt.RtError):
        whole.render(scaled_rig(oracle, g, 0), prof, streams=[0] * 4, keys=[0, 1, 2, 4])
    d = whole.launch_diagnostics()
    assert d == {"allocation_count": 1, "upload_count": 2}
