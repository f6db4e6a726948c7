"""Plain vs learned path at equal relMSE on the four-camera mixed-model scene.

The learned path renders 1 spp with ReSTIR DI and the learned denoiser; the
plain path renders without either at increasing spp.  Both are scored by
relMSE (mean over pixels and RGB of (x - r)^2 / (r^2 + 1e-2)) against a
high-spp plain render of the same camera poses.  The plain frame time at the
learned path's error is interpolated log-log between the two bracketing spp
counts; the result is the ratio of that time to the learned path's frame
time (both device time of all four cameras, median over frames).
"""
from __future__ import annotations

import math
import os
import time

import numpy as np

from . import denoiser_train as T
import ray_tracing_amd as rt

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "denoiser.npy")


def frame_rel_mse(frames, ref) -> float:
    """relMSE over all cameras' pixels (pixel-weighted)."""
    num, den = 0.0, 0
    for x, r in zip(frames, ref):
        e = T.rel_mse(x, r)
        n = x[..., 0].size
        num += e * n
        den += n
    return num / den


def timed_sequence(dev, config, seed, spp, restir, denoise, nn, frame, frames, history=True):
    """Renders orbit frames frame-frames+1 .. frame; returns the last frame's
    beauty per camera and the median device ms of the timed frames.  Without
    `history` every frame starts from a reset accumulation (no temporal
    blending, the plain tracer's estimate of that frame alone)."""
    g = rt.Generated(config, seed)
    sc = rt.Scene(dev, g.desc)
    sc.build_bvh()
    ren = rt.Renderer(dev, sc)
    if nn is not None:
        ren.set_denoiser(nn)
    prof = g.profile
    prof.samples_per_pixel = spp
    prof.enable_restir_di = 1 if restir else 0
    prof.enable_denoise = 1 if denoise else 0
    ms = []
    for f in range(frame - frames + 1, frame + 1):
        if not history:
            ren.reset_accumulation()
        tm = ren.render(T.frame_rig(g, f), prof, 31 * spp + f)
        ms.append(tm.render_ms)
    rig = T.frame_rig(g, frame)
    beauty = [ren.download(c)["beauty"] for c in range(rig.active_count)]
    return beauty, float(np.median(ms[1:] if len(ms) > 1 else ms))


def compare(dev, config="instanced_mesh_restir", seed=1000, ref_spp=4096, frame=6, frames=5,
            max_plain_spp=512, log=print) -> dict:
    t0 = time.time()
    ref, _ = T.render(dev, config, seed, ref_spp, 3, frame=frame, restir=False)
    ref = [r["beauty"] for r in ref]
    log(f"[equal-error] reference {ref_spp} spp in {time.time() - t0:.1f} s")
    nn = rt.Denoiser(0, np.load(WEIGHTS))
    learned, learned_ms = timed_sequence(dev, config, seed, 1, True, True, nn, frame, frames)
    learned_err = frame_rel_mse(learned, ref)
    atrous, atrous_ms = timed_sequence(dev, config, seed, 1, True, True, None, frame, frames)
    atrous_err = frame_rel_mse(atrous, ref)
    log(f"[equal-error] learned: relMSE {learned_err:.4f} at {learned_ms:.3f} ms/frame; "
        f"a-trous: {atrous_err:.4f} at {atrous_ms:.3f} ms")
    plain = []
    spp = 1
    while spp <= max_plain_spp:
        b, ms = timed_sequence(dev, config, seed, spp, False, False, None, frame, 2, history=False)
        err = frame_rel_mse(b, ref)
        plain.append({"spp": spp, "relmse": err, "ms_per_frame": ms})
        log(f"[equal-error] plain {spp} spp: relMSE {err:.4f} at {ms:.3f} ms/frame")
        if err <= learned_err:
            break
        spp *= 2
    equal_ms = None
    if plain[-1]["relmse"] <= learned_err:
        if len(plain) == 1:
            equal_ms = plain[0]["ms_per_frame"]
        else:
            a, b = plain[-2], plain[-1]
            # log-log interpolation of frame time against error
            t = (math.log(a["relmse"]) - math.log(learned_err)) / (
                math.log(a["relmse"]) - math.log(b["relmse"]))
            equal_ms = math.exp(math.log(a["ms_per_frame"])
                                + t * (math.log(b["ms_per_frame"]) - math.log(a["ms_per_frame"])))
    res = {"scene": config, "seed": seed, "orbit_frame": frame, "reference_spp": ref_spp,
           "learned": {"relmse": learned_err, "ms_per_frame": learned_ms,
                       "method": "1 spp + ReSTIR DI + learned denoiser"},
           "atrous": {"relmse": atrous_err, "ms_per_frame": atrous_ms,
                      "method": "1 spp + ReSTIR DI + a-trous filter"},
           "plain": plain,
           "plain_ms_at_equal_relmse": equal_ms,
           "speedup_at_equal_relmse": (equal_ms / learned_ms) if equal_ms else None}
    log(f"[equal-error] plain at equal relMSE: {equal_ms} ms -> speedup "
        f"{res['speedup_at_equal_relmse']}")
    return res


if __name__ == "__main__":
    import json
    d = rt.Device(0)
    print(json.dumps(compare(d)), flush=True)
