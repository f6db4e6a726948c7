"""The oracle's camera models (rto_unproject / rto_project / rto_make_equi_lut,
after src/realtime/gpu/programs.cu:308-532 and
src/realtime/camera_models.cpp:64-318) pinned by

1. the reference's known answers (tests/test_camera_models.cpp:92-220): the
   centre rays, identity-equi theta = 0.3, the LUT ceiling sample, the
   out-of-range fallback, off-axis round trips, the default equi focal
   length 0.5 w / (0.5 hfov);
2. the analytic forward models, restated here from their published form
   (OpenCV's Brown-Conrady k1 k2 k3 p1 p2 distortion for pinhole32; the
   Kannala-Brandt theta polynomial plus the same tangential terms for
   equi62): the oracle's Newton undistortion and LUT inversion must land on
   pixels that the forward model maps back to where they started.

The oracle returns FP32 camera-space directions (as the GPU does), so the
reference's FP64 tolerances (1e-12) become FP32 ones (about 2e-7 per
component) wherever the reference's own tolerance is tighter than that.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ctypes as O

F32 = 2.5e-7   # a few ulps of a unit vector component


def normalized(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def brown_conrady(xy, k1, k2, k3, p1, p2):
    """OpenCV's distortion of normalized coordinates (radial k1..k3, tangential p1 p2)."""
    x, y = xy[..., 0], xy[..., 1]
    r2 = x * x + y * y
    radial = 1 + k1 * r2 + k2 * r2 ** 2 + k3 * r2 ** 3
    xd = x * radial + 2 * p1 * x * y + p2 * (r2 + 2 * x * x)
    yd = y * radial + p1 * (r2 + 2 * y * y) + 2 * p2 * x * y
    return np.stack([xd, yd], -1)


def equi_forward(d, fx, fy, cx, cy, radial, tangential):
    """Kannala-Brandt equidistant projection with a 6-term theta polynomial
    and Brown-Conrady tangential terms on the distorted radius."""
    d = np.asarray(d, np.float64)
    uv = d[:2] / d[2]
    r = np.linalg.norm(uv)
    theta = math.atan(r)
    thetad = theta * (1 + sum(k * theta ** (2 * (i + 1)) for i, k in enumerate(radial)))
    xr = uv * (thetad / r if r > 1e-12 else 1.0)
    p1, p2 = tangential
    x, y = xr
    r2 = x * x + y * y
    xd = x + 2 * p1 * x * y + p2 * (r2 + 2 * x * x)
    yd = y + p1 * (r2 + 2 * y * y) + 2 * p2 * x * y
    return np.array([fx * xd + cx, fy * yd + cy])


PINHOLE = dict(fx=320.0, fy=330.0, cx=160.0, cy=120.0, k1=0.07, k2=-0.03, k3=0.01, p1=0.002,
               p2=-0.0015, width=320, height=240)
PINHOLE_HARD = dict(fx=300.0, fy=315.0, cx=158.0, cy=121.0, k1=0.8, k2=-0.7, k3=0.25, p1=0.03,
                    p2=-0.025, width=320, height=240)
EQUI_RADIAL = (0.02, -0.005, 0.001, -0.0002, 0.00005, -0.00001)
EQUI_TANGENTIAL = (0.0012, -0.0009)


def pinhole(params):
    return O.pinhole_camera(**params)


def test_pinhole_centre_unprojects_to_optical_axis():
    """test_camera_models.cpp:92-93."""
    assert O.unproject(pinhole(PINHOLE), 160.0, 120.0).tolist() == [0.0, 0.0, 1.0]


@pytest.mark.parametrize("params, direction", [
    (PINHOLE, (0.23, -0.17, 1.0)),        # test_camera_models.cpp:125-126
    (PINHOLE_HARD, (1.28, -0.96, 1.0)),   # test_camera_models.cpp:156-158
])
def test_pinhole_off_axis_round_trip(params, direction):
    cam = pinhole(params)
    px = brown_conrady(np.array(direction[:2]) / direction[2], params["k1"], params["k2"],
                       params["k3"], params["p1"], params["p2"])
    pixel = (params["fx"] * px[0] + params["cx"], params["fy"] * px[1] + params["cy"])
    d = O.unproject(cam, *pixel)
    assert np.abs(d - normalized(direction)).max() < 4 * F32
    # the oracle's projection agrees with the forward model (FP32 pixels)
    err = np.abs(O.project(cam, normalized(direction)) - np.array(pixel))
    assert np.all(err < 1e-4 + 4e-7 * np.abs(pixel))


@pytest.mark.parametrize("params", [PINHOLE, PINHOLE_HARD])
def test_pinhole_undistortion_inverts_brown_conrady(params):
    """Every pixel the Newton iteration converges on maps back through the
    forward model to itself (a wrong sign or swapped p1/p2 in the oracle's
    distortion moves pixels by far more than the FP32 tolerance)."""
    cam = pinhole(params)
    xs = np.linspace(4.0, params["width"] - 4.0, 13)
    ys = np.linspace(4.0, params["height"] - 4.0, 11)
    worst = 0.0
    count = 0
    for py in ys:
        for px in xs:
            d = O.unproject(cam, px, py).astype(np.float64)
            xy = d[:2] / d[2]
            xd = brown_conrady(xy, params["k1"], params["k2"], params["k3"], params["p1"],
                               params["p2"])
            back = np.array([params["fx"] * xd[0] + params["cx"], params["fy"] * xd[1] + params["cy"]])
            worst = max(worst, float(np.abs(back - (px, py)).max()))
            count += 1
    assert count == 143
    assert worst < 3e-4, f"pixel drift {worst}"


def test_pinhole_projection_matches_forward_model():
    cam = pinhole(PINHOLE)
    rng = np.random.default_rng(3)
    for _ in range(200):
        d = normalized(np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.35, 0.35), 1.0]))
        xd = brown_conrady(d[:2] / d[2], *(PINHOLE[k] for k in ("k1", "k2", "k3", "p1", "p2")))
        expect = (PINHOLE["fx"] * xd[0] + PINHOLE["cx"], PINHOLE["fy"] * xd[1] + PINHOLE["cy"])
        assert np.abs(O.project(cam, d) - expect).max() < 1e-4
    # behind-the-camera guard of project_pinhole32 (programs.cu:375-378)
    assert O.project(cam, (0.3, 0.2, 0.0)).tolist() == [-1.0, -1.0]


def equi(radial=EQUI_RADIAL, tangential=EQUI_TANGENTIAL, fx=280.0, fy=282.0):
    return O.equi_camera(640, 480, fx, fy, 320.0, 240.0, radial, tangential)


def test_equi_centre_unprojects_to_optical_axis():
    """test_camera_models.cpp:166-167."""
    assert O.unproject(equi(), 320.0, 240.0).tolist() == [0.0, 0.0, 1.0]


def test_equi_identity_theta_is_the_pixel_radius():
    """test_camera_models.cpp:193-196: with no distortion a pixel at radius
    f*theta from the centre looks theta off axis."""
    cam = equi((0.0,) * 6, (0.0, 0.0), 280.0, 280.0)
    theta = 0.3
    d = O.unproject(cam, 320.0 + 280.0 * theta, 240.0)
    assert np.abs(d - (math.sin(theta), 0.0, math.cos(theta))).max() < F32
    for theta in (0.05, 0.6, 1.0):
        d = O.unproject(cam, 320.0, 240.0 + 280.0 * theta)
        assert np.abs(d - (0.0, math.sin(theta), math.cos(theta))).max() < F32


def test_equi_lut_ceiling_sample():
    """test_camera_models.cpp:204-216: the last LUT entry is used exactly at
    the table's end."""
    cam = O.equi_camera(1, 1, 1.0, 1.0, 0.0, 0.0, build_lut=False)
    cam.lut_step = 0.25
    for i in range(1024):
        cam.lut[i] = 0.0
    cam.lut[1023] = 0.3
    d = O.unproject(cam, 0.25 * 1023, 0.0)
    assert np.abs(d - (math.sin(0.3), 0.0, math.cos(0.3))).max() < F32


def test_equi_out_of_range_fallback():
    """test_camera_models.cpp:218-220: beyond the LUT the ray is the
    normalized pinhole ray of the raw coordinates."""
    cam = equi((0.0,) * 6, (0.0, 0.0), 280.0, 280.0)
    d = O.unproject(cam, 320.0 + 280.0 * 5.0, 240.0 - 280.0 * 1.5)
    assert np.abs(d - normalized((5.0, -1.5, 1.0))).max() < F32


def test_equi_lut_inverts_the_theta_polynomial():
    """Each LUT entry is the angle whose distorted angle is its table radius
    (the Newton inversion of make_equi62_lut1d_params), the table covers the
    image with its 10-pixel margin, and it increases."""
    cam = equi()
    lut = np.array(cam.lut[:], np.float64)
    rd = cam.lut_step * np.arange(1024)
    t2 = lut * lut
    thetad = lut * (1 + sum(k * t2 ** (i + 1) for i, k in enumerate(EQUI_RADIAL)))
    assert np.abs(thetad - rd).max() < 1e-11
    assert np.all(np.diff(lut) > 0)
    corner = math.hypot((640 + 10 - 320) / 280.0, (480 + 10 - 240) / 282.0)
    assert abs(cam.lut_step * 1024 - corner) < 1e-12


@pytest.mark.parametrize("direction", [(0.21, 0.08, 1.0), (-0.6, 0.45, 1.0), (0.9, -0.7, 1.0),
                                       (0.02, -0.01, 1.0)])
def test_equi_round_trip_through_forward_model(direction):
    """test_camera_models.cpp:168-169 (2e-6) on the reference's camera and
    three more directions: the forward model's pixel unprojects back to the
    direction within the LUT's interpolation error and the one-step
    tangential removal."""
    cam = equi()
    px = equi_forward(direction, 280.0, 282.0, 320.0, 240.0, EQUI_RADIAL, EQUI_TANGENTIAL)
    d = O.unproject(cam, *px)
    assert np.abs(d - normalized(direction)).max() < 2e-6 + F32
    # projection (the device's LUT inversion) agrees with the forward model
    assert np.abs(O.project(cam, normalized(direction)) - px).max() < 2e-3


def test_equi_projection_of_identity_lut_is_exact():
    cam = equi((0.0,) * 6, (0.0, 0.0), 280.0, 280.0)
    for theta in (0.1, 0.4, 0.9):
        px = O.project(cam, (math.sin(theta), 0.0, math.cos(theta)))
        assert abs(px[0] - (320.0 + 280.0 * theta)) < 2e-4 and px[1] == 240.0


def test_default_equi_focal_length_of_the_generated_rigs():
    """The generated equi cameras use derive_default_camera_intrinsics with
    hfov 120 (camera_models.cpp:198-226; test_camera_models.cpp:104-110:
    f = 0.5 w / (0.5 hfov))."""
    for config, width in ((2, 1024), (3, 1280)):
        cams = [c for c in O.generate_rig(config, 0, 0, 60) if c.model == 1]
        assert cams
        for c in cams:
            f = 0.5 * width / (120.0 * math.pi / 360.0)
            assert abs(c.fx - f) < 1e-9 and abs(c.fy - f) < 1e-9
            assert c.cx == 0.5 * c.width and c.cy == 0.5 * c.height


# ------------------------------------------------------- direction debug image
def contract_rig():
    """The reference's camera contract rig (tests/camera_contract_fixtures.h):
    a distorted pinhole turned 0.35 rad about y at (1, -2, 3.5) and an equi
    camera turned -0.2 rad about x at (-1.5, 0.75, 2.25), both 64 x 48."""
    def rot(axis, a):
        c, s = math.cos(a), math.sin(a)
        if axis == "y":
            return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    pin = O.pinhole_camera(52.0, 50.0, 31.5, 23.5, 0.02, -0.01, 0.003, 0.001, -0.0015, width=64, height=48)
    eq = O.equi_camera(64, 48, 28.0, 29.0, 31.5, 23.5,
                       radial=(0.01, -0.003, 0.0008, -0.0002, 0.00005, -0.00001),
                       tangential=(0.0007, -0.0005))
    rots = (rot("y", 0.35), rot("x", -0.2))
    for cam, R, t in ((pin, rots[0], (1.0, -2.0, 3.5)), (eq, rots[1], (-1.5, 0.75, 2.25))):
        cam.origin[:] = t
        cam.basis_x[:], cam.basis_y[:], cam.basis_z[:] = tuple(R[:, 0]), tuple(R[:, 1]), tuple(R[:, 2])
    return (pin, eq), rots


def solve_forward(forward, pixel):
    """Normalized coordinates (x, y) that the forward model maps to `pixel`."""
    from scipy.optimize import least_squares
    r = least_squares(lambda v: forward(v) - np.asarray(pixel), x0=np.zeros(2), xtol=1e-15,
                      ftol=1e-15, gtol=1e-15)
    assert np.abs(forward(r.x) - np.asarray(pixel)).max() < 1e-9
    return r.x


def encode(d):
    return np.floor(255.0 * np.clip(0.5 * (np.asarray(d) + 1.0), 0.0, 1.0)).astype(int)


def test_direction_debug_matches_the_camera_contract():
    """The reference's direction-debug contract (tests/test_optix_direction.cpp:
    61-125): the image's RGB at sample pixels is the encoded world direction
    of the forward model's inverse within 2 codes, alpha is 255, and the equi
    sample differs from the pinhole sample by at least 12 codes."""
    (pin, eq), (Rp, Re) = contract_rig()
    imgs = []
    for cam in (pin, eq):
        img = np.zeros((cam.height, cam.width, 4), np.uint8)
        O.lib().rto_direction_debug(C.byref(cam), img.ctypes.data)
        assert (img[..., 3] == 255).all()
        imgs.append(img)

    def fwd_pin(v):
        xd = brown_conrady(np.asarray(v), 0.02, -0.01, 0.003, 0.001, -0.0015)
        return np.array([52.0 * xd[0] + 31.5, 50.0 * xd[1] + 23.5])

    def fwd_eq(v):
        return equi_forward(np.array([v[0], v[1], 1.0]), 28.0, 29.0, 31.5, 23.5,
                            (0.01, -0.003, 0.0008, -0.0002, 0.00005, -0.00001), (0.0007, -0.0005))

    for img, fwd, R, samples in ((imgs[0], fwd_pin, Rp, ((31.5, 23.5), (17.5, 11.5), (55.5, 38.5))),
                                 (imgs[1], fwd_eq, Re, ((31.5, 23.5), (45.5, 14.5), (58.5, 40.5)))):
        for px, py in samples:
            x, y = solve_forward(fwd, (px, py))
            world = R @ normalized((x, y, 1.0))
            got = img[int(py), int(px), :3].astype(int)
            assert np.abs(got - encode(world)).max() <= 2, (px, py, got, encode(world))
    a = imgs[1][14, 45, :3].astype(int)
    b = imgs[0][11, 17, :3].astype(int)
    assert np.abs(a - b).max() >= 12
