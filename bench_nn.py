"""Device-resident timing of the learned denoiser's kernels (no host
transfers): one inference over a batch of camera images and one training
step over 8 crops of 256x256, through the C ABI's device-pointer entry points
(rt_nn_infer, rt_nn_train_step) on buffers already in HBM.

Reports ms per call and the arithmetic rate: the dense convolution FLOPs of
the encoder-decoder (2 * 9 * cin * cout per pixel at each level, a level
having a quarter of the pixels of the one above) over the measured time, and
the fraction of MI355X's dense BF16 MFMA peak (2.5 PFLOP/s) that is.

    python bench_nn.py            # 4x640x480, 4x1280x720, 8x256^2 training
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import ray_tracing_amd as rt  # noqa: E402
from ray_tracing_amd import abi  # noqa: E402

BF16_PEAK_TFLOPS = 2500.0
# (channels from this level, from the coarser level, outputs, level)
CONVS = [(16, 0, 32, 0), (32, 0, 32, 0), (32, 0, 48, 1), (48, 0, 48, 1), (48, 0, 64, 2),
         (64, 0, 64, 2), (64, 0, 96, 3), (96, 0, 96, 3), (96, 0, 96, 3), (64, 96, 64, 2),
         (64, 0, 64, 2), (48, 64, 48, 1), (48, 0, 48, 1), (32, 48, 32, 0), (32, 0, 32, 0)]


def level_pixels(sizes, level: int) -> int:
    n = 0
    for w, h in sizes:
        for _ in range(level):
            w, h = (w + 1) // 2, (h + 1) // 2
        n += w * h
    return n


def forward_flops(sizes) -> float:
    """Dense multiply-add FLOPs of the 15 convolutions and the 1x1 head."""
    f = 0.0
    for c0, c1, co, lvl in CONVS:
        f += 2.0 * 9 * (c0 + c1) * co * level_pixels(sizes, lvl)
    return f + 2.0 * 32 * 3 * level_pixels(sizes, 0)


def train_flops(sizes) -> float:
    """Forward, input gradient (every convolution but the first) and weight
    gradient: three times the forward less the first layer's input gradient."""
    first = 2.0 * 9 * 16 * 32 * level_pixels(sizes, 0)
    return 3.0 * forward_flops(sizes) - first


def _frame(torch, P: int, seed: int):
    g = torch.Generator(device="cuda").manual_seed(seed)
    beauty = torch.rand((P, 4), device="cuda", generator=g) * 2
    normal = torch.rand((P, 4), device="cuda", generator=g)
    albedo = torch.rand((P, 4), device="cuda", generator=g)
    depth = torch.rand((P,), device="cuda", generator=g) * 10 + 1
    target = torch.rand((P, 4), device="cuda", generator=g)
    return beauty, normal, albedo, depth, target


def _p(t):
    return C.c_void_p(t.data_ptr())


def time_infer(nn, torch, sizes, reps: int, warmup: int) -> float:
    im = abi.nn_images(sizes)
    P = im.offset[im.count]
    b, n, a, d, _ = _frame(torch, P, 1)
    out = torch.empty((P, 4), device="cuda")
    call = lambda: rt.check(rt.lib().rt_nn_infer(nn.h, C.byref(im), _p(b), _p(n), _p(a), _p(d),
                                                 _p(out)))
    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()   # synchronises the library's stream before it returns
    return (time.perf_counter() - t0) * 1e3 / reps


def time_train(nn, torch, sizes, reps: int, warmup: int) -> float:
    im = abi.nn_images(sizes)
    P = im.offset[im.count]
    b, n, a, d, t = _frame(torch, P, 2)
    loss = C.c_float()
    call = lambda: rt.check(rt.lib().rt_nn_train_step(nn.h, C.byref(im), _p(b), _p(n), _p(a), _p(d),
                                                      _p(t), C.c_float(1e-4), C.byref(loss), None))
    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    return (time.perf_counter() - t0) * 1e3 / reps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--json", default="")
    a = ap.parse_args()
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("bench_nn needs a GPU")
    weights = os.path.join(HERE, "ray_tracing_amd", "weights", "denoiser.npy")
    nn = rt.Denoiser(0, np.load(weights))
    res = []
    for name, sizes in (("infer 4x640x480", [(640, 480)] * 4),
                        ("infer 4x1280x720", [(1280, 720)] * 4)):
        # the convolution tile shapes autotuned for this batch (rt_nn_autotune)
        tuned = nn.autotune(sizes, reps=5)
        ms = time_infer(nn, torch, sizes, a.reps, a.warmup)
        fl = forward_flops(sizes)
        res.append({"case": name, "ms": ms, "gflop": fl / 1e9, "tflops": fl / ms / 1e9,
                    "frac_bf16_peak": fl / ms / 1e9 / BF16_PEAK_TFLOPS, "conv_rows": tuned["conv_rows"]})
    sizes = [(256, 256)] * 8
    tuned = nn.autotune(sizes, reps=5)
    ms = time_train(nn, torch, sizes, a.reps, a.warmup)
    fl = train_flops(sizes)
    res.append({"case": "train step 8x256^2 (forward, backward, Adam)", "ms": ms,
                "gflop": fl / 1e9, "tflops": fl / ms / 1e9,
                "frac_bf16_peak": fl / ms / 1e9 / BF16_PEAK_TFLOPS, "conv_rows": tuned["conv_rows"]})
    for r in res:
        print(f"[nn] {r['case']}: {r['ms']:.3f} ms, {r['gflop']:.1f} GFLOP, "
              f"{r['tflops']:.1f} TFLOP/s ({100 * r['frac_bf16_peak']:.2f} % of BF16 peak), "
              f"tile rows per level {r['conv_rows']}", flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
