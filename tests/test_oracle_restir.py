"""Pins the oracle's ReSTIR DI reservoir math and analytic-light CDF.

Known answers are those of the reference's tests/test_restir_di.cpp (W = 4/3,
merged W = 0.6, M = 5, temporal M = 4, age 4, spatial W = 0.625, BASIC
W = 1.25, six surface-validity cases); random update/merge sequences are also
compared with the reference header compiled in place (oracle/_ref).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_ctypes as O


def res(**kw):
    r = O.Reservoir()
    r.light_index = -1
    r.material_type = r.primitive_type = r.primitive_index = -1
    for k, v in kw.items():
        setattr(r, k, v)
    return r


def update(r, light, target, weight, mult, u, u0=0.25, u1=0.75):
    return O.lib().rto_restir_update(C.byref(r), light, u0, u1, target, weight, mult, u)


def merge(cur, prev, target, max_count, u, temporal):
    sel = C.c_int()
    m = O.lib().rto_restir_merge(C.byref(cur), C.byref(prev), target, max_count, u,
                                 1 if temporal else 0, C.byref(sel))
    return m, sel.value


def test_reference_known_answers():
    L = O.lib()
    r = res()
    update(r, 3, 1.0, 1.0, 1, 0.0)
    update(r, 9, 3.0, 3.0, 1, 0.99)
    update(r, 11, 0.0, 0.0, 1, 0.5)
    L.rto_restir_finalize(C.byref(r))
    assert r.valid and r.light_index == 3 and r.candidate_count == 3
    assert abs(r.weight_sum - 4.0) < 1e-7
    assert abs(r.estimator_weight - 4.0 / 3.0) < 1e-7

    prev = res(light_index=21, sample_u0=0.1, sample_u1=0.2, weight_sum=16.0, selected_target=4.0,
               estimator_weight=0.5, candidate_count=8, age=3, valid=1, material_type=5,
               primitive_type=2, primitive_index=7)
    prev.position = O.f3(0.0, 0.0, 2.0)
    prev.normal = O.f3(0.0, 0.0, 1.0)
    merged = res()
    update(merged, 4, 1.0, 2.0, 1, 0.0)
    m, sel = merge(merged, prev, 2.0, 4, 0.0, True)
    L.rto_restir_finalize(C.byref(merged))
    assert merged.light_index == 21 and merged.candidate_count == 5
    assert merged.temporal_candidate_count == 4 and m == 4 and sel == 1 and merged.age == 4
    assert abs(merged.weight_sum - 6.0) < 1e-7 and abs(merged.estimator_weight - 0.6) < 1e-7

    spatial = res()
    update(spatial, 4, 1.0, 2.0, 1, 0.0)
    m, sel = merge(spatial, prev, 2.0, 3, 0.0, False)
    L.rto_restir_finalize(C.byref(spatial))
    assert spatial.candidate_count == 4 and spatial.spatial_candidate_count == 3
    assert spatial.temporal_candidate_count == 0 and m == 3 and sel == 1
    assert abs(spatial.weight_sum - 5.0) < 1e-7 and abs(spatial.estimator_weight - 0.625) < 1e-7

    corrected = res(light_index=8, weight_sum=10.0, selected_target=2.0, candidate_count=5, valid=1)
    L.rto_restir_finalize_basic(C.byref(corrected), 3.0, 12.0, 2)
    assert corrected.valid and corrected.bias_correction_source_count == 2
    assert abs(corrected.estimator_weight - 1.25) < 1e-7


def test_update_sequences_match_reference():
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(3)
    for _ in range(300):
        r = res()
        rs = np.zeros(9, np.float32)
        rs[8] = -1
        for _ in range(rng.integers(1, 40)):
            light = int(rng.integers(-1, 50))
            target = float(rng.choice([0.0, rng.exponential(2.0)]))
            weight = float(rng.choice([0.0, rng.exponential(3.0)]))
            mult = int(rng.integers(0, 4))
            u = float(rng.uniform(-0.1, 1.1))
            a = update(r, light, target, weight, mult, u)
            sel = C.c_int()
            R.ref_restir_update(rs.ctypes.data, light, target, weight, mult, u, C.byref(sel))
            assert a == sel.value
            assert r.weight_sum == rs[0] and r.selected_target == rs[1]
            assert r.candidate_count == int(rs[3]) and r.valid == int(rs[7])
            assert r.light_index == int(rs[8])
        O.lib().rto_restir_finalize(C.byref(r))
        R.ref_restir_finalize(rs.ctypes.data)
        assert r.estimator_weight == rs[2] and r.valid == int(rs[7])


def test_analytic_cdf_closes_and_is_monotone():
    """The generated 256 sphere lights, compiled by the oracle's light
    compiler: selection pdf = world area x luminance / total (grey radiance,
    so luminance = radiance), cdf non-decreasing and closed at 1."""
    from helpers import ALIGHT_DT
    g = O.GeneratedScene(2, 0)
    n = g.scene.analytic_light_count
    assert n == 256
    lights = np.ctypeslib.as_array(
        C.cast(g.scene.analytic_lights, C.POINTER(C.c_uint8)), (n * ALIGHT_DT.itemsize,)).view(ALIGHT_DT)
    sel, cdf = lights["selection_pdf"], lights["cdf"]
    assert np.all(np.diff(cdf) >= 0) and cdf[-1] == 1.0
    radius, radiance = lights["radius"].astype(np.float64), lights["radiance"][:, 0].astype(np.float64)
    w = 4 * np.pi * radius ** 2 * radiance
    assert np.allclose(sel, w / w.sum(), rtol=1e-5)
    assert np.allclose(lights["world_area"], 4 * np.pi * radius ** 2, rtol=1e-6)