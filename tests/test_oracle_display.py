"""The oracle's viewer tone map (rto_display_u8 / rto_tone_map_rgba8, after
src/realtime/viewer/cuda_gl_presenter.cu:23-38) pinned by

1. the reference's known answers (tests/test_display_transfer.cpp:4-12);
2. the reference's host twin rt::linear_to_display_u8
   (src/realtime/display_transfer.h:9-15) compiled in place (oracle/_ref),
   on every float around each of the 255 rounding thresholds and a dense
   sweep of [-2, 8];
3. an independent numpy restatement (correctly rounded float32 sqrt and
   product, ties rounded away from zero in float64, where they are exact);
4. invariants: monotone in the input, exactly one step per display level,
   NaN and negative inputs black, +inf at 255, alpha always 255.
"""
import numpy as np
import pytest

import oracle_ctypes as O
from helpers import sweep, threshold_neighbourhood  # noqa: F401

needs_ref = pytest.mark.skipif(O.ref_lib() is None,
                               reason="oracle/_ref not built (reference tree absent)")


def oracle_u8(v: np.ndarray) -> np.ndarray:
    v = np.asarray(v, np.float32).ravel()
    px = np.zeros((v.size, 4), np.float32)
    px[:, 0] = v
    return O.tone_map_rgba8(px)[:, 0]


def numpy_u8(v: np.ndarray) -> np.ndarray:
    v = np.asarray(v, np.float32)
    with np.errstate(invalid="ignore"):
        d = np.sqrt(np.fmax(v, np.float32(0.0)))
    d = np.minimum(d, np.float32(0.999))
    x = (d * np.float32(255.0)).astype(np.float32).astype(np.float64)
    return np.floor(x + 0.5).astype(np.uint8)      # x >= 0: half away from zero


def test_reference_known_answers():
    """test_display_transfer.cpp:4-12."""
    L = O.lib()
    assert L.rto_display_u8(0.0) == 0             # zero stays black
    assert L.rto_display_u8(-1.0) == 0            # negative radiance clamps to black
    assert L.rto_display_u8(0.04) == 51           # gamma 2: sqrt(0.04) = 0.2 -> 51
    assert L.rto_display_u8(4.0) == 255           # 0.999 * 255 = 254.745 -> 255


def test_presenter_test_pixels():
    """The source pixels of tests/test_cuda_gl_presenter.cpp:44-49, whose
    expectations are the host u8 conversion per channel and alpha 255; the
    values follow from sqrt(v) * 255 (0.25 -> 127.5, a tie, rounds up)."""
    src = np.array([[0.0, 0.25, 1.0, 0.0], [4.0, -1.0, 0.01, 0.0],
                    [0.04, 0.16, 0.36, 0.0], [0.64, 0.81, 0.99, 0.0]], np.float32)
    out = O.tone_map_rgba8(src)
    assert out.tolist() == [[0, 128, 255, 255], [255, 0, 26, 255],
                            [51, 102, 153, 255], [204, 230, 254, 255]]


@needs_ref
def test_matches_reference_host_conversion():
    R = O.ref_lib()
    v = np.concatenate([threshold_neighbourhood(8), np.linspace(-1.0, 5.0, 20001, dtype=np.float32)])
    ref = np.array([R.ref_linear_to_display_u8(float(x)) for x in v], np.uint8)
    assert np.array_equal(oracle_u8(v), ref)
    assert abs(R.ref_linear_to_display(0.04) - 0.2) < 1e-6
    assert R.ref_linear_to_display(4.0) == pytest.approx(0.999, abs=1e-6)


def test_matches_independent_restatement():
    v = sweep()
    assert np.array_equal(oracle_u8(v), numpy_u8(v))


def test_invariants():
    v = np.sort(np.concatenate([threshold_neighbourhood(), np.linspace(0, 1.2, 400001, dtype=np.float32)]))
    u = oracle_u8(v).astype(np.int32)
    steps = np.diff(u)
    assert (steps >= 0).all() and (steps <= 1).all()
    assert u[0] == 0 and u[-1] == 255 and int(steps.sum()) == 255
    # level k is reached where sqrt(v) * 255 crosses k - 1/2
    first = np.searchsorted(u, np.arange(1, 256), side="left")
    x = np.sqrt(v[first].astype(np.float64)) * 255.0
    assert (np.abs(x - (np.arange(1, 256) - 0.5)) < 1e-4).all()
    special = np.array([np.nan, -np.nan, np.inf, -np.inf, -1e30, -0.0], np.float32)
    assert oracle_u8(special).tolist() == [0, 0, 255, 0, 0, 0]
    px = np.random.default_rng(1).uniform(-1, 3, (1000, 4)).astype(np.float32)
    assert (O.tone_map_rgba8(px)[:, 3] == 255).all()
