"""Offline training and evaluation of the learned denoiser (the 4-level
encoder-decoder of rt_nn.h, 565,443 parameters).

Training pairs come from this path tracer on seeded procedural scenes:
- the Cornell rooms with 1-8 random spheres, materials and light of the
  scene generator's `cornell_random` scenes (seeds 0..N-1), plus a few
  seeds of the instanced ReSTIR scene and the random-sphere scene the same
  network denoises in the realtime loop;
- inputs are 1-spp frames with their albedo / normal / depth buffers, taken
  the way the realtime loop produces them (ReSTIR DI on the instanced scene
  after a few frames of its camera orbit);
- targets are 1024-spp renders of the same camera poses by the plain tracer.

Training runs on the GPU through the library's own BF16 MFMA forward /
backward / Adam kernels (rt_nn_train_step): batches of 8 random 256x256
crops, Adam with lr 1e-3 on a cosine schedule, minimising the log-space
squared error weighted by (1 + t)^2 / (t^2 + 1e-2) (t the target), which is
the relMSE below to first order.  Evaluation scores relMSE =
mean over pixels and RGB of (x - r)^2 / (r^2 + 1e-2) against 4096-spp plain
renders of held-out seeds (1000, 1001, ...) for: the raw 1-spp frame, the
a-trous filter, and the network.

    python -m ray_tracing_amd.denoiser_train --out ray_tracing_amd/weights/denoiser.npy
    python -m ray_tracing_amd.denoiser_train --weights ray_tracing_amd/weights/denoiser.npy \
        --eval-seeds 50 --report eval.json        # evaluation only
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ray_tracing_amd as rt  # noqa: E402

EPS_REL = 1e-2
CONFIGS = {"random_spheres": 0, "instanced_mesh_restir": 2, "cornell_random": 4}


def rel_mse(x: np.ndarray, ref: np.ndarray) -> float:
    x = np.asarray(x, np.float64)[..., :3]
    ref = np.asarray(ref, np.float64)[..., :3]
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + EPS_REL)))


def frame_rig(g: rt.Generated, frame: int):
    try:
        return g.rig_at(frame, 60)
    except rt.RtError:
        return g.rig()


def render(dev, config: str, seed: int, spp: int, stream: int, frame: int = 0, restir=None,
           warm_frames: int = 0, denoise: bool = False, timing: bool = False):
    """Frames of every camera at orbit frame `frame`; `warm_frames` earlier
    orbit frames are rendered first so ReSTIR has its temporal history."""
    g = rt.Generated(config, seed)
    sc = rt.Scene(dev, g.desc)
    sc.build_bvh()
    ren = rt.Renderer(dev, sc)
    prof = g.profile
    prof.samples_per_pixel = spp
    prof.enable_denoise = 1 if denoise else 0
    if restir is not None:
        prof.enable_restir_di = 1 if restir else 0
    for f in range(frame - warm_frames, frame):
        ren.render(frame_rig(g, f), prof, 7919 * stream + f)
    rig = frame_rig(g, frame)
    tm = ren.render(rig, prof, 7919 * stream + frame)
    out = []
    for c in range(rig.active_count):
        d = ren.download(c)
        gd = ren.guides(c)
        d["resolved"] = gd["resolved"]
        out.append(d)
    return out, tm


def make_pairs(dev, config: str, seeds, target_spp: int, log=print):
    pairs = []
    restir_scene = config == "instanced_mesh_restir"
    for s in seeds:
        t0 = time.time()
        inp, _ = render(dev, config, s, 1, 1, frame=4 if restir_scene else 0,
                        warm_frames=4 if restir_scene else 0)
        tgt, _ = render(dev, config, s, target_spp, 2, frame=4 if restir_scene else 0,
                        restir=False)
        for a, b in zip(inp, tgt):
            pairs.append((a, b["beauty"]))
        log(f"[train] {config} seed {s}: {len(inp)} images, {time.time() - t0:.1f} s")
    return pairs


def crop_batch(rng, pairs, batch: int, size: int):
    frames, targets, sizes = [], [], []
    weights = np.array([p[0]["depth"].size for p in pairs], np.float64)
    weights /= weights.sum()
    for _ in range(batch):
        a, t = pairs[rng.choice(len(pairs), p=weights)]
        h, w = a["depth"].shape
        cy = rng.integers(0, max(h - size, 0) + 1)
        cx = rng.integers(0, max(w - size, 0) + 1)
        hh, ww = min(size, h), min(size, w)
        sl = (slice(cy, cy + hh), slice(cx, cx + ww))
        frames.append({k: a[k][sl] for k in ("beauty", "normal", "albedo", "depth")})
        targets.append(t[sl])
        sizes.append((ww, hh))
    cat = {k: np.concatenate([f[k].reshape(-1, *f[k].shape[2:]) for f in frames])
           for k in ("beauty", "normal", "albedo", "depth")}
    tgt = np.concatenate([t.reshape(-1, 4) for t in targets])
    return sizes, cat, tgt


# (input channels from this level, from the coarser level, outputs) of the
# encoder-decoder's 15 convolutions, in parameter order
CONVS = [(16, 0, 32), (32, 0, 32), (32, 0, 48), (48, 0, 48), (48, 0, 64), (64, 0, 64),
         (64, 0, 96), (96, 0, 96), (96, 0, 96), (64, 96, 64), (64, 0, 64), (48, 64, 48),
         (48, 0, 48), (32, 48, 32), (32, 0, 32)]


def init_params(rng) -> np.ndarray:
    """He-style init for the convolutions, a zero head (the untrained network
    returns its input)."""
    n = rt.lib().rt_nn_param_count()
    p = np.zeros(n, np.float32)
    o = 0
    for c0, c1, cout in CONVS:
        fan = 9 * (c0 + c1)
        nw = cout * fan
        p[o:o + nw] = rng.normal(0, np.sqrt(2.0 / fan), nw)
        o += nw + cout
    return p


def train(pairs, steps: int, batch: int, crop: int, lr: float, seed: int = 0, log=print):
    rng = np.random.default_rng(seed)
    nn = rt.Denoiser(0, init_params(rng))
    t0 = time.time()
    hist = []
    for i in range(steps):
        sizes, frame, tgt = crop_batch(rng, pairs, batch, crop)
        cur = lr * 0.5 * (1.0 + np.cos(np.pi * i / steps))
        loss, _ = nn.train_step(sizes, frame, tgt, float(cur))
        hist.append(loss)
        if i % 100 == 0 or i == steps - 1:
            log(f"[train] step {i} loss {np.mean(hist[-100:]):.5f} lr {cur:.2e} "
                f"({time.time() - t0:.1f} s)")
    return nn, hist


def evaluate(dev, nn, config: str, seed: int, ref_spp: int, log=print) -> dict:
    restir_scene = config == "instanced_mesh_restir"
    frame = 4 if restir_scene else 0
    warm = 4 if restir_scene else 0
    t0 = time.time()
    ref, _ = render(dev, config, seed, ref_spp, 3, frame=frame, restir=False)
    raw, tm_raw = render(dev, config, seed, 1, 4, frame=frame, warm_frames=warm, denoise=True)
    res = {"config": config, "seed": seed, "ref_spp": ref_spp, "cameras": []}
    for c, (r, x) in enumerate(zip(ref, raw)):
        h, w = x["depth"].shape
        learned = nn.infer([(w, h)], x).reshape(h, w, 4)
        res["cameras"].append({"camera": c, "size": [w, h],
                               "relmse_raw": rel_mse(x["resolved"], r["beauty"]),
                               "relmse_atrous": rel_mse(x["beauty"], r["beauty"]),
                               "relmse_learned": rel_mse(learned, r["beauty"])})
    for k in ("relmse_raw", "relmse_atrous", "relmse_learned"):
        res[k] = float(np.mean([c[k] for c in res["cameras"]]))
    log(f"[eval] {config} seed {seed}: raw {res['relmse_raw']:.4f} a-trous "
        f"{res['relmse_atrous']:.4f} learned {res['relmse_learned']:.4f} ({time.time() - t0:.1f} s)")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "denoiser.npy"))
    ap.add_argument("--report", default="")
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--target-spp", type=int, default=1024)
    ap.add_argument("--ref-spp", type=int, default=4096)
    ap.add_argument("--cornell-seeds", type=int, default=48)
    ap.add_argument("--instanced-seeds", type=int, default=4)
    ap.add_argument("--spheres-seeds", type=int, default=4)
    ap.add_argument("--eval-seeds", type=int, default=50, help="held-out cornell_random seeds")
    ap.add_argument("--eval-start", type=int, default=1000, help="first held-out seed")
    ap.add_argument("--weights", default="", help="evaluate these weights instead of training")
    a = ap.parse_args()
    dev = rt.Device(0)
    report = {}
    if a.weights:
        nn = rt.Denoiser(0, np.load(a.weights))
        report["weights"] = os.path.basename(a.weights)
    else:
        pairs = make_pairs(dev, "cornell_random", range(a.cornell_seeds), a.target_spp)
        pairs += make_pairs(dev, "instanced_mesh_restir", range(a.instanced_seeds), a.target_spp)
        pairs += make_pairs(dev, "random_spheres", range(a.spheres_seeds), a.target_spp)
        nn, hist = train(pairs, a.steps, a.batch, a.crop, a.lr)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        np.save(a.out, nn.params())
        print(f"[train] wrote {a.out}", flush=True)
        report["train"] = {"steps": a.steps, "batch": a.batch, "crop": a.crop, "lr": a.lr,
                           "target_spp": a.target_spp, "pairs": len(pairs),
                           "cornell_seeds": a.cornell_seeds, "instanced_seeds": a.instanced_seeds,
                           "spheres_seeds": a.spheres_seeds,
                           "loss_first": float(np.mean(hist[:50])),
                           "loss_last": float(np.mean(hist[-50:]))}
    seeds = [a.eval_start + i for i in range(a.eval_seeds)]
    evals = [evaluate(dev, nn, "cornell_random", s, a.ref_spp) for s in seeds]
    summary = {k: float(np.mean([e[k] for e in evals]))
               for k in ("relmse_raw", "relmse_atrous", "relmse_learned")}
    print(f"[eval] cornell_random seeds {seeds[0]}..{seeds[-1]}: " +
          " ".join(f"{k} {v:.4f}" for k, v in summary.items()), flush=True)
    report["eval_cornell_random"] = {"seeds": seeds, "ref_spp": a.ref_spp, **summary}
    report["eval"] = evals + [evaluate(dev, nn, "random_spheres", 1000, a.ref_spp),
                              evaluate(dev, nn, "instanced_mesh_restir", 1000, a.ref_spp)]
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    print(json.dumps(report["eval_cornell_random"]), flush=True)


if __name__ == "__main__":
    main()
