"""Device-resident timing of the viewer tone map (rt_tone_map_rgba8): float4
frames already in HBM, RGBA8 out, `--iters` back-to-back launches on the
device's stream between two synchronisations.  Algorithmic traffic is 20
bytes per pixel (16 read, 4 written).

    python bench_present.py [--pixels N ...] [--iters K]
"""
import argparse
import json
import time

import ray_tracing_amd as rt


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--pixels", type=int, nargs="+",
                    default=[4 * 640 * 480, 4 * 1280 * 720, 4 * 3840 * 2160])
    ap.add_argument("--iters", type=int, default=200)
    a = ap.parse_args(argv)
    import torch
    torch.cuda.init()   # torch's bundled HIP runtime opens the device before librt_hip.so's
    dev = rt.Device(0)
    rows = []
    for n in a.pixels:
        src = torch.rand((n, 4), device="cuda") * 2
        dst = torch.empty((n, 4), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        for _ in range(5):
            rt.tone_map_rgba8(dev, src.data_ptr(), dst.data_ptr(), n)
        dev.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.iters):
            rt.tone_map_rgba8(dev, src.data_ptr(), dst.data_ptr(), n)
        dev.synchronize()
        us = (time.perf_counter() - t0) / a.iters * 1e6
        row = {"pixels": n, "us_per_launch": round(us, 2),
               "GB_per_s": round(20.0 * n / us / 1e3, 1)}
        print(json.dumps(row), flush=True)
        rows.append(row)
        del src, dst
    dev.close()
    return rows


if __name__ == "__main__":
    main()
