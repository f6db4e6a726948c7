"""The oracle's homogeneous media (oracle_render.c sample_media /
medium_boundary, after src/realtime/gpu/programs.cu:1356-1419,1711-1753)
pinned by closed forms:

- the fixed-sequence logarithm of the free-flight distance: within 2 ulp of
  the FP64 logarithm over [1e-6, 1], exact at powers of two;
- Beer-Lambert: a pure absorber (isotropic, albedo 0) in front of a
  constant sky leaves sky x exp(-density x chord), with the chord of a box
  (slab intersection, also rotated and translated in its own frame) and of
  a sphere (2 sqrt(r^2 - b^2));
- single scattering at depth 1 (next-event estimation only; shadow rays
  pass through media, as in the reference): a medium of albedo a under a
  uniform sky shows sky x (T + a (1 - T)), T the transmittance along the
  pixel ray, which fixes the event probability, the isotropic phase
  function's density and NEE's normalization together.
"""
import math

import numpy as np
import pytest

import oracle_ctypes as O
from helpers import QUAD_DT, medium
from test_oracle_integrator import assert_z, material, mc, profile

SKY = (0.5, 0.6, 0.7)


def test_log_fixed_is_within_two_ulp():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(1e-6, 1.0, 40000), np.exp(rng.uniform(math.log(1e-6), 0.0, 40000)),
                        [1.0, 0.5, 0.25, 2.0 ** -10, 1e-6, 0.70710678, 0.7071068, 0.99999994]]).astype(np.float32)
    f = O.lib().rto_log_fixed
    got = np.array([f(float(v)) for v in x], np.float32)
    ref = np.log(x.astype(np.float64))
    ulp = np.abs(np.spacing(ref.astype(np.float32)).astype(np.float64))
    err = np.abs(got.astype(np.float64) - ref) / ulp
    assert err.max() <= 2.0, float(err.max())
    for k in range(0, 20):
        assert f(2.0 ** -k) == np.float32(-k * math.log(2.0))


def camera(w=16, h=12, f=14.0, z=4.0):
    """Pinhole at (0, 0, z) looking down -z."""
    c = O.pinhole_camera(f, f, w / 2, h / 2, width=w, height=h)
    c.origin[:] = (0.0, 0.0, z)
    c.basis_x[:] = (1.0, 0.0, 0.0)
    c.basis_y[:] = (0.0, -1.0, 0.0)
    c.basis_z[:] = (0.0, 0.0, -1.0)
    return c


def pixel_rays(cam):
    rays = O.primary_rays(cam, 0, 0, 0, 0)
    return rays["o"].astype(np.float64), rays["d"].astype(np.float64)


def box_chord(o, d, lo, hi, R=np.eye(3), t=np.zeros(3)):
    lo_, d_ = (o - t) @ np.asarray(R).T, d @ np.asarray(R).T
    with np.errstate(divide="ignore", invalid="ignore"):
        a0 = (np.asarray(lo) - lo_) / d_
        a1 = (np.asarray(hi) - lo_) / d_
    t0 = np.maximum(np.minimum(a0, a1).max(-1), 0.0)
    t1 = np.maximum(a0, a1).min(-1)
    return np.clip(t1 - t0, 0.0, None) * np.linalg.norm(d, axis=-1)


def sphere_chord(o, d, c, r):
    dn = d / np.linalg.norm(d, axis=-1, keepdims=True)
    v = np.asarray(c) - o
    b2 = (v * v).sum(-1) - (v * dn).sum(-1) ** 2
    return 2.0 * np.sqrt(np.clip(r * r - b2, 0.0, None))


def scene(media, albedo):
    mats = material(4, albedo=(albedo,) * 3)
    # a small quad far behind the camera keeps the primitive list non-empty
    q = np.zeros(1, QUAD_DT)
    q[0] = ((50, 50, 50), (0.1, 0, 0), (0, 0.1, 0), 0)
    return O.scene_from_arrays(quads=q, materials=mats, background=SKY, media=media)


R45 = np.array([[math.cos(0.6), 0, math.sin(0.6)], [0, 1, 0], [-math.sin(0.6), 0, math.cos(0.6)]])

CASES = {
    "box": (lambda: medium("box", 0, 1.3, lo=(-1.5, -1.2, -0.4), hi=(1.5, 1.2, 0.6)),
            lambda o, d: box_chord(o, d, (-1.5, -1.2, -0.4), (1.5, 1.2, 0.6))),
    "rotated_box": (lambda: medium("box", 0, 0.9, lo=(-1.0, -1.0, -0.5), hi=(1.0, 1.0, 0.5),
                                   rotation=R45, translation=(0.2, -0.1, 0.3)),
                    lambda o, d: box_chord(o, d, (-1.0, -1.0, -0.5), (1.0, 1.0, 0.5), R45,
                                           np.array([0.2, -0.1, 0.3]))),
    "sphere": (lambda: medium("sphere", 0, 1.1, centre=(0.1, 0.0, -0.2), radius=1.2),
               lambda o, d: sphere_chord(o, d, (0.1, 0.0, -0.2), 1.2)),
}


@pytest.mark.parametrize("kind", list(CASES))
def test_beer_lambert_through_an_absorber(kind):
    make, chord = CASES[kind]
    m = make()
    s, keep = scene(m, 0.0)
    cam = camera()
    o, d = pixel_rays(cam)
    T = np.exp(-float(m["density"][0]) * chord(o, d)).reshape(cam.height, cam.width)
    assert 0.02 < T.min() < 0.9 and T.max() == 1.0     # the medium covers part of the view
    ref = T[..., None] * np.array(SKY)
    mean, se = mc(s, cam, profile(32, 3))
    assert_z(mean, se, ref, label=f"{kind} absorber")


@pytest.mark.parametrize("kind", list(CASES))
def test_single_scattering_under_a_uniform_sky(kind):
    make, chord = CASES[kind]
    a = 0.7
    m = make()
    m["material_index"] = 0
    s, keep = scene(m, a)
    cam = camera()
    o, d = pixel_rays(cam)
    T = np.exp(-float(m["density"][0]) * chord(o, d)).reshape(cam.height, cam.width)
    ref = (T + a * (1.0 - T))[..., None] * np.array(SKY)
    mean, se = mc(s, cam, profile(32, 1))
    assert_z(mean, se, ref, label=f"{kind} single scattering")
