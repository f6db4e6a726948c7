"""The camera-model debug image (rt_render_direction_debug, after
direction_debug_renderer.cpp:51-66 and programs.cu:1807-1826) on the
reference's contract rig: equal to the oracle's image byte for byte for the
distorted pinhole32 and the equi62_lut1d camera (the oracle's image is
pinned to the forward camera models in test_oracle_camera.py); a camera
index outside the rig is refused."""
import ctypes as C
import sys

import numpy as np
import pytest

from test_oracle_camera import contract_rig

pytestmark = pytest.mark.gpu


def test_direction_debug_equals_oracle(rt, oracle, device):
    print("[gpu] direction debug, contract rig", flush=True)
    sys.stdout.flush()
    cams, _ = contract_rig()
    rig = rt.CameraRig()
    rig.active_count = 2
    for i, c in enumerate(cams):
        C.memmove(C.byref(rig.cameras[i]), C.byref(c), C.sizeof(c))
    for i, c in enumerate(cams):
        gpu = rt.direction_debug(device, rig, i)
        ref = np.zeros((c.height, c.width, 4), np.uint8)
        oracle.lib().rto_direction_debug(C.byref(c), ref.ctypes.data)
        assert gpu.shape == (48, 64, 4)
        assert gpu.tobytes() == ref.tobytes(), f"camera {i}"
    with pytest.raises(rt.RtError):
        rt.direction_debug(device, rig, 2)
