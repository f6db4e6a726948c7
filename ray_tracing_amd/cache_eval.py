"""Radiance cache + learned denoiser against the plain tracer on the
four-camera 1280x720 workload (config radiance_cache_realtime).

Every method renders the same orbit sequence, ending at orbit frame `frame`,
and is scored by relMSE (mean over pixels and RGB of (x - r)^2 / (r^2 +
1e-2)) of that frame against a high-spp plain render of the same poses:
  nn        1 spp + ReSTIR DI + learned denoiser
  cache     1 spp + ReSTIR DI + radiance cache (paths end at vertex 2)
  cache_nn  1 spp + ReSTIR DI + radiance cache + learned denoiser
  plain     the plain tracer (no ReSTIR, no learned part, no temporal
            blending) at increasing spp, until it reaches the best learned
            method's error; its frame time there is interpolated log-log.
Frame time is the device time of the whole frame (all cameras, the cache's
training and inference and the denoiser included), median over the scored
sequence's last frames.
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
    num, den = 0.0, 0
    for x, r in zip(frames, ref):
        e = T.rel_mse(x, r)
        n = x[..., 0].size
        num += e * n
        den += n
    return num / den


def run_sequence(dev, config, seed, frame, frames, spp=1, restir=True, nn=None, cache=None,
                 history=True, timed=4, log=None):
    """Renders orbit frames frame - frames + 1 .. frame; returns the last
    frame's beauty per camera, the median device ms of the last `timed`
    frames and the cache's statistics per frame."""
    g = rt.Generated(config, seed)
    sc = rt.Scene(dev, g.desc)
    sc.build_bvh()
    ren = rt.Renderer(dev, sc)
    if nn is not None:
        ren.set_denoiser(nn)
    if cache is not None:
        ren.set_radiance_cache(cache)
    prof = g.profile
    prof.samples_per_pixel = spp
    prof.enable_restir_di = 1 if restir else 0
    prof.enable_denoise = 1 if nn is not None else 0
    ms, stats = [], []
    for f in range(frame - frames + 1, frame + 1):
        if not history:
            ren.reset_accumulation()
        tm = ren.render(T.frame_rig(g, f), prof, 31 * spp + f)
        ms.append(tm.render_ms)
        if cache is not None:
            stats.append(cache.stats())
            if log:
                s = stats[-1]
                log(f"    frame {f}: {tm.render_ms:.2f} ms, cache loss {s['last_loss']:.3f} "
                    f"ema {s['loss_ema']:.3f} usable {s['usable']} records {s['records']} "
                    f"queries {s['queries']}")
    rig = T.frame_rig(g, frame)
    beauty = [ren.download(c)["beauty"] for c in range(rig.active_count)]
    ren.set_radiance_cache(None)
    return beauty, float(np.median(ms[-timed:])), stats


def plain_at_error(dev, config, seed, frame, target_err, max_spp, log):
    plain = []
    spp = 1
    while spp <= max_spp:
        b, ms, _ = run_sequence(dev, config, seed, frame, 2, spp=spp, restir=False, history=False,
                                timed=1)
        err = frame_rel_mse(b, REF_CACHE["ref"])
        plain.append({"spp": spp, "relmse": err, "ms_per_frame": ms})
        log(f"[cache-eval] plain {spp} spp: relMSE {err:.4f} at {ms:.2f} ms/frame")
        if err <= target_err:
            break
        spp *= 2
    equal_ms = None
    if plain and plain[-1]["relmse"] <= target_err:
        if len(plain) == 1:
            equal_ms = plain[0]["ms_per_frame"]
        else:
            a, b = plain[-2], plain[-1]
            t = (math.log(a["relmse"]) - math.log(target_err)) / (
                math.log(a["relmse"]) - math.log(b["relmse"]))
            equal_ms = math.exp(math.log(a["ms_per_frame"])
                                + t * (math.log(b["ms_per_frame"]) - math.log(a["ms_per_frame"])))
    return plain, equal_ms


REF_CACHE: dict = {}


def compare(dev, config="radiance_cache_realtime", seed=1000, ref_spp=4096, frame=16, frames=16,
            max_plain_spp=512, log=print, cache_overrides=None) -> dict:
    t0 = time.time()
    ref, _ = T.render(dev, config, seed, ref_spp, 3, frame=frame, restir=False)
    REF_CACHE["ref"] = [r["beauty"] for r in ref]
    log(f"[cache-eval] reference {ref_spp} spp in {time.time() - t0:.1f} s")
    weights = np.load(WEIGHTS)
    results = {}
    cfg = dict(cache_overrides or {})
    for name, use_nn, use_cache in [("nn", True, False), ("cache", False, True),
                                    ("cache_nn", True, True)]:
        nn = rt.Denoiser(0, weights) if use_nn else None
        cache = rt.RadianceCache(0, **cfg) if use_cache else None
        log(f"[cache-eval] {name}")
        b, ms, stats = run_sequence(dev, config, seed, frame, frames, nn=nn, cache=cache,
                                    log=log if use_cache else None)
        err = frame_rel_mse(b, REF_CACHE["ref"])
        results[name] = {"relmse": err, "ms_per_frame": ms}
        if stats:
            results[name]["cache"] = stats[-1]
        log(f"[cache-eval] {name}: relMSE {err:.4f} at {ms:.2f} ms/frame")
    best = min(results, key=lambda k: results[k]["relmse"])
    plain, equal_ms = plain_at_error(dev, config, seed, frame, results[best]["relmse"],
                                     max_plain_spp, log)
    res = {"scene": config, "seed": seed, "orbit_frame": frame, "sequence_frames": frames,
           "reference_spp": ref_spp, "methods": results, "best": best, "plain": plain,
           "plain_ms_at_best_relmse": equal_ms,
           "speedup_at_equal_relmse": (equal_ms / results[best]["ms_per_frame"]) if equal_ms else None}
    log(f"[cache-eval] best {best}; plain at its relMSE: {equal_ms} ms -> speedup "
        f"{res['speedup_at_equal_relmse']}")
    return res


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser()
    ap.add_argument("--ref-spp", type=int, default=4096)
    ap.add_argument("--frame", type=int, default=16)
    ap.add_argument("--frames", type=int, default=16)
    ap.add_argument("--max-rel-loss", type=float, default=None)
    ap.add_argument("--out", default=None)
    ap.add_argument("--profile-frames", type=int, default=0,
                    help="only render this many frames of cache + denoiser (for a kernel trace)")
    a = ap.parse_args()
    ov = {}
    if a.max_rel_loss is not None:
        ov["max_rel_loss"] = a.max_rel_loss
    d = rt.Device(0)
    if a.profile_frames > 0:
        _, ms, _ = run_sequence(d, "radiance_cache_realtime", 1000, a.profile_frames,
                                a.profile_frames, nn=rt.Denoiser(0, np.load(WEIGHTS)),
                                cache=rt.RadianceCache(0, **ov), timed=a.profile_frames)
        print(f"[cache-eval] cache + denoiser: {a.profile_frames} frames, median {ms:.2f} ms/frame",
              flush=True)
        raise SystemExit(0)
    r = compare(d, ref_spp=a.ref_spp, frame=a.frame, frames=a.frames, cache_overrides=ov)
    txt = json.dumps(r)
    print(txt, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt + "\n")
