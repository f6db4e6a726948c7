"""Benchmark of the path-tracing core on one MI355X.

Workload (default): cornell_mesh -- Cornell room + 81,920-triangle icosphere
with one OpenPBR material, 4 pinhole cameras 640x480, 16 spp, depth 8; one
step = one frame of all four cameras through the wavefront pipeline.  The
scene, BVH and buffers are resident in HBM before the timed region.

Prints one JSON line: Mpaths/s (whole job), ms per frame, the roofline of
the dominant kernel (closest-hit traversal, algorithmic bytes from measured
node visits / primitive tests) and the CPU oracle timed on this host on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# canonical unit sizes for algorithmic traffic (bytes)
RAY_B, HIT_B, NODE_B, PRIM_IDX_B, TRI_B, SPHERE_B = 32, 16, 32, 4, 36, 16
ANY_RESULT_B = 4


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(config: str, seconds: float, threads: int) -> dict:
    """Times the CPU oracle on whole cameras of the workload, camera after
    camera (then again with the next sample stream), until `seconds` of work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes as C

    import oracle_ctypes as O
    import ray_tracing_amd as rt

    g = rt.Generated(config, 0)
    desc = g.desc
    s = O.Scene()
    for name in ("sphere_count", "quad_count", "triangle_count", "material_count",
                 "openpbr_count", "analytic_light_count", "background_mode"):
        setattr(s, name, getattr(desc, name))
    s.background = O.f3(*desc.background)
    s.background_top = O.f3(*desc.background_top)
    s.spheres = C.cast(desc.spheres, C.POINTER(O.Sphere))
    s.quads = C.cast(desc.quads, C.POINTER(O.Quad))
    s.triangles = C.cast(desc.triangles, C.POINTER(O.Triangle))
    s.materials = C.cast(desc.materials, C.POINTER(O.Material))
    s.openpbr = C.cast(desc.openpbr, C.POINTER(O.OpenPbr))
    s.analytic_lights = C.cast(desc.analytic_lights, C.POINTER(O.AnalyticLight))
    bvh = O.OracleBvh(s)
    prof = O.Profile()
    C.memmove(C.byref(prof), C.byref(g.g.profile), C.sizeof(prof))
    ncam = g.g.camera_count
    paths = 0
    done = []
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < seconds and k < 64:
        ci, stream = k % ncam, k // ncam
        cam = O.Camera()
        C.memmove(C.byref(cam), C.byref(g.g.cameras[ci]), C.sizeof(cam))
        out = O.render_camera(s, bvh, cam, ci, prof, stream, threads=threads)
        paths += out["paths"]
        done.append(f"cam{ci}/stream{stream}")
        k += 1
    dt = time.perf_counter() - t0
    return {"value": paths / dt / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"{config}: {len(done)} whole camera images ({', '.join(done)}; "
                      f"{paths} paths, {prof.samples_per_pixel} spp, depth {prof.max_bounces}) "
                      f"in {dt:.2f} s on {threads} threads"}


TRAFFIC_PROFILE = {"cornell_mesh": "k_trace_fused_cornell_mesh_pmc.md",
                   "instanced_mesh_restir": "k_trace_fused_instanced_mesh_restir_pmc.md"}


def measured_traffic(config: str, tuning: dict, rays_per_launch: float):
    """HBM bytes per k_extend launch from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE scaled by the copy
    calibration).  When this run's launch shape differs from the profiled one
    the figure is scaled by traced rays per launch and says so."""
    name = TRAFFIC_PROFILE.get(config)
    if name is None:
        return None, None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        text = open(path).read()
    except OSError:
        return None, None
    vals = {}
    for line in text.splitlines():
        if ":" in line and line.split(":")[0] in ("traffic_bytes_per_launch",
                                                   "rays_per_launch", "tuning"):
            k, v = line.split(":", 1)
            vals[k] = v.strip()
    try:
        traffic = float(vals["traffic_bytes_per_launch"])
        rays = float(vals["rays_per_launch"])
        shape = dict(kv.split("=") for kv in vals["tuning"].split())
    except (KeyError, ValueError):
        return None, None
    same = all(int(shape[k]) == int(tuning.get(k, -1)) for k in shape)
    if same:
        return traffic, f"profiles/{name}"
    return traffic * rays_per_launch / rays, f"profiles/{name}, scaled by traced rays per launch"


def cam_desc(cam) -> str:
    return f"{cam.width}x{cam.height} {'equi' if cam.model == 1 else 'pinhole'}"


def measure(rt, dev, info, config, steps, warmup, tune, autotune_first, per_camera):
    """Throughput, frame time, stage times and the traversal roofline of one
    workload: scene, BVH and buffers resident, one step = one frame of every
    camera."""
    g = rt.Generated(config, 0)
    scene = rt.Scene(dev, g.desc)
    bst = scene.build_bvh()
    ren = rt.Renderer(dev, scene)
    if tune:
        ren.set_tuning(**tune)
    rig = g.rig()
    prof = g.profile
    # scenes with an orbiting rig advance one orbit step per frame (ReSTIR
    # history then reprojects real camera motion); the others keep a fixed rig
    try:
        orbit = [g.rig_at(i, 60) for i in range(60)]
    except rt.RtError:
        orbit = None

    def rig_for(i):
        return orbit[i % 60] if orbit else rig
    paths_per_frame = sum(rig.cameras[i].width * rig.cameras[i].height
                          for i in range(rig.active_count)) * prof.samples_per_pixel
    log(f"[bench] {info['name']} CUs={info['cu_count']} config={config} prims={scene.prim_count} "
        f"lbvh depth={bst.max_depth} 4-wide nodes={bst.wide_node_count} levels={bst.wide_levels} "
        f"build={bst.build_ms:.2f} ms paths/frame={paths_per_frame}")

    autotune = None
    if autotune_first:
        autotune = ren.autotune(rig_for(0), prof, 2)
        log(f"[bench] autotune: {autotune['trials']} trials, {autotune['baseline_ms']:.3f} -> "
            f"{autotune['best_ms']:.3f} ms, tuning {autotune['tuning']}")
    # the counting frame needs per-stage events: it runs without the graph
    run_tuning = ren.tuning()

    # traversal counts for the algorithmic-bytes roofline (separate, untimed frame)
    ren.set_tuning(count_traversal=1, use_graph=0)
    cnt = ren.render(rig_for(0), prof, 0)
    ren.set_tuning(count_traversal=0)
    # stage times (and the roofline's per-launch duration) come from one
    # eager frame with per-stage events, the same rig and stream as the
    # counting frame; the timed steps use the chosen tuning
    stage = ren.render(rig_for(0), prof, 0)
    ren.set_tuning(use_graph=run_tuning["use_graph"])
    nonsphere = cnt.prim_tests - cnt.sphere_tests
    extend_bytes = (cnt.extend_rays * (RAY_B + HIT_B) + cnt.node_visits * NODE_B
                    + nonsphere * (PRIM_IDX_B + TRI_B) + cnt.sphere_tests * (PRIM_IDX_B + SPHERE_B))
    # any-hit rays: the same formula with their early-exit counts and a 4-byte result
    any_nonsphere = cnt.shadow_prim_tests - cnt.shadow_sphere_tests
    shadow_bytes = (cnt.shadow_rays * (RAY_B + ANY_RESULT_B) + cnt.shadow_node_visits * NODE_B
                    + any_nonsphere * (PRIM_IDX_B + TRI_B)
                    + cnt.shadow_sphere_tests * (PRIM_IDX_B + SPHERE_B))
    trace_bytes = extend_bytes + shadow_bytes
    log(f"[bench] counting frame: rays={cnt.extend_rays} visits/ray="
        f"{cnt.node_visits / max(cnt.extend_rays, 1):.2f} (child boxes) tests/ray="
        f"{cnt.prim_tests / max(cnt.extend_rays, 1):.2f}; shadow rays={cnt.shadow_rays} "
        f"visits/ray={cnt.shadow_node_visits / max(cnt.shadow_rays, 1):.2f} "
        f"tests/ray={cnt.shadow_prim_tests / max(cnt.shadow_rays, 1):.2f}")

    for i in range(warmup):
        ren.render(rig_for(i), prof, 1000 + i)
    # the orbit poses' rigs (equirect lookup tables included) are inputs,
    # built before the timed region like the scene
    rigs = [rig_for(warmup + i) for i in range(steps)]
    # device time per frame from timed frames (each waits for its frame) of
    # the same poses as the timed steps
    frames = [ren.render(rigs[i % steps], prof, 2000 + i) for i in range(max(3, min(steps, 10)))]
    log(f"[bench] {config} device ms per frame: {[round(f.render_ms, 3) for f in frames]}")
    # the timed region: untimed asynchronous frames, each returning once
    # enqueued, so the host stages frame k + 1 while frame k runs
    ren.set_async(True)
    dev.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        ren.render(rigs[i], prof, i, timed=False)
        log(f"[bench] {config} step {i} enqueued at {(time.perf_counter() - t0) * 1e3:.1f} ms")
    ren.synchronize()
    dev.synchronize()
    wall = time.perf_counter() - t0
    ren.set_async(False)
    ms_per_step = wall * 1e3 / steps
    value = paths_per_frame * steps / wall / 1e6
    # every traversal launch of the frame: one per bounce (extension rays with
    # the previous bounce's shadow rays) plus one for the last shadow rays
    ext_ms = stage.extend_ms
    trace_ms = stage.extend_ms + stage.connect_ms
    launches = stage.extend_launches + stage.extend_launches // max(prof.max_bounces, 1)
    avg_launch_ms = trace_ms / max(launches, 1)
    achieved = trace_bytes / (trace_ms * 1e-3) / 1e9 if trace_ms > 0 else 0.0
    gpu_ms = sorted(f.render_ms for f in frames)[len(frames) // 2]

    traffic, traffic_source = measured_traffic(config, ren.tuning(),
                                               (cnt.extend_rays + cnt.shadow_rays) / max(launches, 1))
    cfg = {"workload": f"{config}: wavefront path tracer, "
                       f"{'+'.join(cam_desc(rig.cameras[i]) for i in range(rig.active_count))}, "
                       f"{prof.samples_per_pixel} spp, depth {prof.max_bounces}"
                       f"{', ReSTIR DI' if prof.enable_restir_di else ''}",
           "cameras": rig.active_count, "spp": prof.samples_per_pixel,
           "max_bounces": prof.max_bounces, "paths_per_frame": paths_per_frame,
           "gpu_ms_per_frame": round(gpu_ms, 3),
           "trace_ms_per_frame": round(trace_ms, 3),
           "shade_ms_per_frame": round(stage.shade_ms, 3),
           "last_shadow_launch_ms": round(stage.connect_ms, 3),
           "stage_times_from": "one eager frame with per-stage events",
           "extend_rays_per_frame": cnt.extend_rays,
           "shadow_rays_per_frame": cnt.shadow_rays,
           "child_boxes_per_ray": round(cnt.node_visits / max(cnt.extend_rays, 1), 3),
           "prim_tests_per_ray": round(cnt.prim_tests / max(cnt.extend_rays, 1), 3),
           "bvh_build_ms": round(bst.build_ms, 3),
           "tuning": ren.tuning(), "autotune": autotune}
    if per_camera and rig.active_count > 1:
        cfg["per_camera"] = camera_breakdown(rt, ren, rig, prof, steps)
    roof = {"bound": "hbm", "kernel": "k_trace_fused (closest-hit + any-hit traversal)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": int(traffic) if traffic is not None else None,
            "traffic_source": traffic_source,
            "bytes_per_launch": int(trace_bytes / max(launches, 1)),
            "closest_hit_bytes_per_frame": int(extend_bytes),
            "any_hit_bytes_per_frame": int(shadow_bytes),
            "avg_launch_ms": round(avg_launch_ms, 4), "launches_per_frame": launches}
    return {"value": round(value, 3), "ms_per_step": round(ms_per_step, 3), "config": cfg,
            "roofline": roof}


def camera_breakdown(rt, ren, rig, prof, steps):
    """Each camera of the rig rendered alone: frame time, throughput and
    rays per path (a camera that sees open sky or wall exteriors ends most
    paths after one or two rays, one inside the room traces many more)."""
    out = []
    for c in range(rig.active_count):
        one = rt.CameraRig()
        one.active_count = 1
        one.cameras[0] = rig.cameras[c]
        paths = one.cameras[0].width * one.cameras[0].height * prof.samples_per_pixel
        ren.set_tuning(count_traversal=1, use_graph=0)
        cnt = ren.render(one, prof, 0)
        ren.set_tuning(count_traversal=0)
        ms = sorted(ren.render(one, prof, 1 + i).render_ms for i in range(max(3, steps // 2)))
        med = ms[len(ms) // 2]
        out.append({"camera": c, "ms_per_frame": round(med, 3),
                    "mpaths_per_s": round(paths / (med * 1e-3) / 1e6, 1),
                    "extend_rays_per_path": round(cnt.extend_rays / paths, 3),
                    "child_boxes_per_ray": round(cnt.node_visits / max(cnt.extend_rays, 1), 2)})
        log(f"[bench] camera {c} alone: {out[-1]}")
    return out


def host_cpu():
    """The host's CPU model (lscpu) and logical core count."""
    model = None
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    return model, os.cpu_count() or 1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="cornell_mesh")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--wavefront-paths", type=int, default=0)
    ap.add_argument("--trace-block", type=int, default=0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (3 times), the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the plain vs learned comparison at equal relMSE")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tuning instead of running the autotuner first")
    ap.add_argument("--graph", type=int, default=-1, help="force hipGraph capture off (0) or on (1)")
    ap.add_argument("--second-config", default="instanced_mesh_restir",
                    help="a second workload measured the same way ('' skips it)")
    ap.add_argument("--no-per-camera", action="store_true",
                    help="skip the per-camera breakdown of the headline workload")
    ap.add_argument("--refill", type=int, default=0,
                    help="idle lanes before a traversal wave takes new rays (1..64)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this path tracer shards by camera with no "
              "collectives, and multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    if a.calibrate_copy:
        import torch  # device memory for the calibration copy only

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: 3 copies of {src.numel() * 4} bytes")
        return

    import ray_tracing_amd as rt

    dev = rt.Device(0)
    info = dev.info()
    tune = {}
    if a.wavefront_paths:
        tune["wavefront_paths"] = a.wavefront_paths
    if a.trace_block:
        tune["trace_block"] = a.trace_block
    if a.blocks_per_cu:
        tune["trace_blocks_per_cu"] = a.blocks_per_cu
    if a.refill:
        tune["trace_refill"] = a.refill
    if a.graph >= 0:
        tune["use_graph"] = a.graph
    m = measure(rt, dev, info, a.config, a.steps, a.warmup, tune, not a.no_autotune,
                per_camera=not a.no_per_camera)
    result = {
        "metric": "Mpaths/s & ms/frame 4-cam; plain vs learned path at equal relMSE",
        "value": m["value"],
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": m["config"],
        "roofline": m["roofline"],
    }
    if a.second_config and a.second_config != a.config:
        # the north star's four-camera mixed-model scene: the same measurement
        # (throughput, frame time, traversal roofline) beside the headline one
        try:
            m2 = measure(rt, dev, info, a.second_config, a.steps, a.warmup, tune,
                         not a.no_autotune, per_camera=False)
            result["mixed_model_4cam"] = {"value": m2["value"], "unit": "Mpaths/s",
                                          "ms_per_step": m2["ms_per_step"],
                                          "config": m2["config"], "roofline": m2["roofline"]}
        except Exception as e:  # the headline line is still printed
            result["mixed_model_4cam"] = {"error": f"{type(e).__name__}: {e}"}
    if not a.no_cpu_baseline:
        # every core this process may use, up to the 16 host CPUs a GPU box
        # allots to each GPU whatever os.cpu_count() reports
        model, logical = host_cpu()
        try:
            usable = len(os.sched_getaffinity(0))
        except AttributeError:
            usable = logical
        threads = max(1, min(usable, 16))
        result["cpu_baseline"] = cpu_baseline(a.config, a.cpu_seconds, threads)
        result["cpu_baseline"].update({"host_cpu_model": model, "host_logical_cpus": logical,
                                       "usable_cpus": usable})
    if not a.no_learned:
        # plain vs learned path at equal relMSE on the four-camera mixed-model
        # scene (untimed references; reported beside the throughput metric)
        try:
            from ray_tracing_amd import equal_error

            result["learned_vs_plain"] = equal_error.compare(dev, log=log)
        except Exception as e:  # the throughput line is still printed
            result["learned_vs_plain"] = {"error": f"{type(e).__name__}: {e}"}
        # radiance cache + learned denoiser on the four-camera 1280x720 depth-8
        # workload, against the plain tracer at equal relMSE
        try:
            from ray_tracing_amd import cache_eval

            result["cache_vs_plain"] = cache_eval.compare(dev, log=log)
        except Exception as e:
            result["cache_vs_plain"] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
