"""ray_tracing_amd -- MI355X (gfx950) path-tracing core.

Python view of the C ABI in include/rt/rt_hip.h, used by the tests and by
bench.py.  The compute path is the in-tree HIP library librt_hip.so; if it is
missing this module raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi
from .abi import (Camera, CameraRig, FrameTiming, GeneratedScene, Profile, SceneDesc, Tuning,
                  BvhStats, RAY_DTYPE, HIT_DTYPE, NODE_DTYPE)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librt_hip.so")

RT_OK = 0
STATUS_NAMES = {0: "RT_OK", 1: "RT_ERROR_INVALID_ARGUMENT", 2: "RT_ERROR_HIP",
                3: "RT_ERROR_NOT_READY", 4: "RT_ERROR_OUT_OF_MEMORY", 5: "RT_ERROR_UNSUPPORTED"}

# every entry point declared in include/rt/rt_hip.h
EXPORTED = [
    "rt_abi_version", "rt_last_error", "rt_device_create", "rt_device_destroy",
    "rt_device_synchronize", "rt_device_info", "rt_device_stream", "rt_scene_upload", "rt_scene_destroy",
    "rt_bvh_build", "rt_bvh_download", "rt_scene_update", "rt_scene_last_update", "rt_trace_closest", "rt_trace_any",
    "rt_trace_closest_device", "rt_trace_closest_device_noalloc", "rt_renderer_create", "rt_renderer_destroy", "rt_render_frame",
    "rt_render_frame_streams", "rt_renderer_set_async", "rt_renderer_synchronize", "rt_renderer_launch_diagnostics", "rt_analytic_lights_compile",
    "rt_render_direction_debug",
    "rt_frame_download", "rt_frame_device_view", "rt_reset_accumulation", "rt_reset_sequence",
    "rt_restir_diagnostics", "rt_renderer_set_tuning", "rt_renderer_get_tuning",
    "rt_scene_generate", "rt_scene_generated_free", "rt_restir_download",
    "rt_scene_generate_rig", "rt_frame_download_guides", "rt_bvh_download_wide",
    "rt_nn_param_count", "rt_nn_create", "rt_nn_destroy", "rt_nn_set_params", "rt_nn_get_params",
    "rt_nn_infer", "rt_nn_train_step", "rt_nn_infer_host", "rt_nn_train_step_host",
    "rt_nn_set_tiling", "rt_nn_get_tiling", "rt_nn_autotune",
    "rt_renderer_set_denoiser", "rt_renderer_set_denoiser_budget", "rt_renderer_denoiser_guard",
    "rt_renderer_autotune",
    "rt_cache_default_config", "rt_cache_mlp_param_count", "rt_cache_param_count",
    "rt_cache_create", "rt_cache_destroy", "rt_cache_set_params", "rt_cache_get_params",
    "rt_cache_set_bounds", "rt_cache_grid_info", "rt_cache_infer_host",
    "rt_cache_train_step_host", "rt_cache_get_stats", "rt_renderer_set_radiance_cache",
    "rt_camera_build_lut", "rt_presenter_create", "rt_presenter_destroy", "rt_presenter_present",
    "rt_presenter_pixels", "rt_presenter_download", "rt_tone_map_rgba8",
]

# rt_present_status (CudaGlPresentStatus, src/realtime/viewer/cuda_gl_presenter.h:14-18)
PRESENT_OK, PRESENT_INVALID_DEVICE_VIEW, PRESENT_RESOLUTION_MISMATCH = 0, 1, 2


class RtError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


_lib = None


def lib():
    """Loads librt_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        path = os.environ.get("RT_HIP_LIBRARY", LIB_PATH)  # a variant build for experiments
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: run __graft_entry__.build() or "
                              "python ray_tracing_amd/build.py")
        L = C.CDLL(path)
        abi.declare(L)
        _lib = L
    return _lib


def check(status: int):
    if status != RT_OK:
        raise RtError(status, lib().rt_last_error().decode())


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Device:
    def __init__(self, ordinal: int = 0):
        self.h = C.c_void_p()
        check(lib().rt_device_create(ordinal, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().rt_device_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(lib().rt_device_synchronize(self.h))

    def info(self):
        name = C.create_string_buffer(128)
        cu = C.c_int()
        tot = C.c_uint64()
        free = C.c_uint64()
        check(lib().rt_device_info(self.h, name, 128, C.byref(cu), C.byref(tot), C.byref(free)))
        return {"name": name.value.decode(), "cu_count": cu.value, "total_bytes": tot.value,
                "free_bytes": free.value}


class Scene:
    def __init__(self, device: Device, desc: SceneDesc, keepalive=None):
        self.device = device
        self.desc = desc
        self._keep = keepalive
        self.h = C.c_void_p()
        check(lib().rt_scene_upload(device.h, C.byref(desc), C.byref(self.h)))
        self.stats = None

    def close(self):
        if self.h:
            lib().rt_scene_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def prim_count(self) -> int:
        d = self.desc
        return d.sphere_count + d.quad_count + d.triangle_count

    def build_bvh(self) -> BvhStats:
        st = BvhStats()
        check(lib().rt_bvh_build(self.h, C.byref(st)))
        self.stats = st
        return st

    def update(self, desc: SceneDesc, keepalive=None) -> dict:
        """Replaces the scene's contents (rt_scene_update): returns the
        classification (reuse / update / refit / rebuild) and its timings."""
        u = abi.AccelerationUpdate()
        check(lib().rt_scene_update(self.h, C.byref(desc), C.byref(u)))
        self.desc = desc
        self._keep = keepalive
        if self.stats is not None:
            self.stats.prim_count = u.primitive_reference_count
            self.stats.node_count = u.node_count
            self.stats.wide_node_count = u.wide_node_count
        return self._update_dict(u)

    def last_update(self) -> dict:
        u = abi.AccelerationUpdate()
        check(lib().rt_scene_last_update(self.h, C.byref(u)))
        return self._update_dict(u)

    @staticmethod
    def _update_dict(u) -> dict:
        d = {k: getattr(u, k) for k, _ in abi.AccelerationUpdate._fields_}
        d["kind"] = abi.ACCEL_KINDS.get(u.kind, str(u.kind))
        return d

    def download_bvh(self):
        n = self.prim_count
        nodes_n = self.stats.node_count
        nodes = np.zeros(nodes_n, NODE_DTYPE)
        sp = np.zeros(n, np.uint32)
        mc = np.zeros(n, np.uint32)
        check(lib().rt_bvh_download(self.h, ptr(nodes), ptr(sp), ptr(mc)))
        return nodes, sp, mc

    def download_wide(self) -> np.ndarray:
        out = np.zeros(self.stats.wide_node_count, abi.WIDE_NODE_DTYPE)
        if len(out):
            check(lib().rt_bvh_download_wide(self.h, ptr(out)))
        return out

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        check(lib().rt_trace_closest(self.h, ptr(rays), len(rays), ptr(hits)))
        return hits

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        occ = np.zeros(len(rays), np.int32)
        check(lib().rt_trace_any(self.h, ptr(rays), len(rays), ptr(occ)))
        return occ


class Renderer:
    def __init__(self, device: Device, scene: Scene):
        self.device = device
        self.scene = scene
        self.h = C.c_void_p()
        check(lib().rt_renderer_create(device.h, scene.h, C.byref(self.h)))
        self.rig = None

    def close(self):
        if self.h:
            lib().rt_renderer_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, **kw):
        t = Tuning()
        t.use_graph = -1
        t.count_traversal = -1
        for k, v in kw.items():
            setattr(t, k, v)
        check(lib().rt_renderer_set_tuning(self.h, C.byref(t)))

    def tuning(self) -> dict:
        t = Tuning()
        check(lib().rt_renderer_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def render(self, rig: CameraRig, profile: Profile, sample_stream: int = 0,
               timed: bool = True, streams=None, keys=None) -> FrameTiming | None:
        """One frame of the rig. `streams` (four sample streams, one per
        camera slot) and `keys` (four camera keys) give each slot its own
        random-number identity; by default every slot draws `sample_stream`
        under its own index."""
        self.rig = rig
        tm = FrameTiming()
        out = C.byref(tm) if timed else None
        if streams is None and keys is None:
            check(lib().rt_render_frame(self.h, C.byref(rig), C.byref(profile), sample_stream, out))
        else:
            st = (C.c_uint32 * 4)(*(streams if streams is not None else [sample_stream] * 4))
            ky = (C.c_uint32 * 4)(*keys) if keys is not None else None
            check(lib().rt_render_frame_streams(self.h, C.byref(rig), C.byref(profile), st, ky, out))
        return tm if timed else None

    def set_async(self, enable: bool):
        """Untimed frames return once enqueued (host staging double-buffered);
        downloads still see complete frames."""
        check(lib().rt_renderer_set_async(self.h, 1 if enable else 0))

    def synchronize(self):
        check(lib().rt_renderer_synchronize(self.h))

    def launch_diagnostics(self) -> dict:
        a, u = C.c_uint64(), C.c_uint64()
        check(lib().rt_renderer_launch_diagnostics(self.h, C.byref(a), C.byref(u)))
        return {"allocation_count": a.value, "upload_count": u.value}

    def guides(self, camera: int) -> dict:
        """Flow (pixels to the previous frame), flow trust and the resolved
        beauty before denoising."""
        cam = self.rig.cameras[camera]
        w, h = cam.width, cam.height
        out = {"flow": np.zeros((h, w, 2), np.float32), "trust": np.zeros((h, w), np.float32),
               "resolved": np.zeros((h, w, 4), np.float32)}
        check(lib().rt_frame_download_guides(self.h, camera, ptr(out["flow"]), ptr(out["trust"]),
                                             ptr(out["resolved"])))
        return out

    def autotune(self, rig: CameraRig, profile: Profile, frames_per_trial: int = 2) -> dict:
        res = abi.AutotuneResult()
        check(lib().rt_renderer_autotune(self.h, C.byref(rig), C.byref(profile), frames_per_trial,
                                         C.byref(res)))
        out = {"baseline_ms": res.baseline_ms, "best_ms": res.best_ms, "trials": res.trials,
               "tuning": {k: getattr(res.best, k) for k, _ in Tuning._fields_}}
        if res.denoiser_tuned:
            d = res.denoiser
            out["denoiser"] = {"conv_rows": list(d.best.conv_rows), "best_ms": d.best_ms,
                               "baseline_ms": d.baseline_ms, "trials": d.trials}
        return out

    def set_denoiser(self, nn: "Denoiser | None"):
        self._nn = nn  # keeps the network alive while attached
        check(lib().rt_renderer_set_denoiser(self.h, nn.h if nn is not None else None))

    def set_denoiser_budget(self, max_mean_shift: float, max_residual_noise: float):
        check(lib().rt_renderer_set_denoiser_budget(self.h, max_mean_shift, max_residual_noise))

    def denoiser_guard(self) -> dict:
        g = abi.DenoiserGuard()
        check(lib().rt_renderer_denoiser_guard(self.h, C.byref(g)))
        out = {}
        for k, _ in abi.DenoiserGuard._fields_:
            v = getattr(g, k)
            out[k] = list(v) if hasattr(v, "__len__") else v
        return out

    def set_radiance_cache(self, cache: "RadianceCache | None"):
        self._cache = cache
        check(lib().rt_renderer_set_radiance_cache(self.h, cache.h if cache is not None else None))

    def reservoirs(self, camera: int) -> np.ndarray:
        cam = self.rig.cameras[camera]
        out = np.zeros(cam.width * cam.height, abi.RESERVOIR_DTYPE)
        check(lib().rt_restir_download(self.h, camera, ptr(out)))
        return out

    def restir_stats(self, camera: int) -> dict:
        st = abi.RestirStats()
        check(lib().rt_restir_diagnostics(self.h, camera, C.byref(st)))
        return {k: getattr(st, k) for k, _ in abi.RestirStats._fields_}

    def reset_accumulation(self):
        check(lib().rt_reset_accumulation(self.h))

    def download(self, camera: int) -> dict:
        cam = self.rig.cameras[camera]
        w, h = cam.width, cam.height
        out = {"beauty": np.zeros((h, w, 4), np.float32), "normal": np.zeros((h, w, 4), np.float32),
               "albedo": np.zeros((h, w, 4), np.float32), "depth": np.zeros((h, w), np.float32)}
        check(lib().rt_frame_download(self.h, camera, ptr(out["beauty"]), ptr(out["normal"]),
                                      ptr(out["albedo"]), ptr(out["depth"])))
        return out

    def device_view(self, camera: int) -> "abi.DeviceFrameView":
        v = abi.DeviceFrameView()
        check(lib().rt_frame_device_view(self.h, camera, C.byref(v)))
        return v


class Presenter:
    """CudaGlTexturePresenter (src/realtime/viewer/cuda_gl_presenter.h:20-40)
    over a device RGBA8 buffer of a fixed resolution."""

    def __init__(self, device: Device, width: int, height: int):
        self.device = device
        self.width, self.height = width, height
        self.h = C.c_void_p()
        check(lib().rt_presenter_create(device.h, width, height, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().rt_presenter_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def present(self, view) -> int:
        """Tone maps a DeviceFrameView (or None) and returns the present status."""
        st = C.c_int32(-1)
        check(lib().rt_presenter_present(self.h, C.byref(view) if view is not None else None,
                                         C.byref(st)))
        return st.value

    def device_pixels(self) -> int:
        p = C.c_void_p()
        check(lib().rt_presenter_pixels(self.h, C.byref(p)))
        return p.value

    def download(self) -> np.ndarray:
        out = np.empty((self.height, self.width, 4), np.uint8)
        check(lib().rt_presenter_download(self.h, ptr(out)))
        return out


def direction_debug(device: "Device", rig: CameraRig, camera: int = 0) -> np.ndarray:
    """rt_render_direction_debug: (height, width, 4) uint8 encoded world
    directions of the camera's pixel-centre rays."""
    cam = rig.cameras[camera]
    out = np.zeros((max(cam.height, 0), max(cam.width, 0), 4), np.uint8)
    check(lib().rt_render_direction_debug(device.h, C.byref(rig), camera,
                                          ptr(out) if out.size else None))
    return out


def compile_analytic_lights(lights) -> None:
    """rt_analytic_lights_compile in place over an array of analytic lights
    (a numpy structured array with AnalyticLight's layout, or a ctypes
    array): world areas, delta distant lights, selection pdf and cdf."""
    n = len(lights)
    addr = lights.ctypes.data if isinstance(lights, np.ndarray) else C.addressof(lights)
    check(lib().rt_analytic_lights_compile(C.c_void_p(addr), n))


def tone_map_rgba8(device: Device, d_beauty: int, d_rgba8: int, pixel_count: int):
    """tone_map_to_rgba8 between device buffers (addresses as ints), on the
    device's stream; does not wait."""
    check(lib().rt_tone_map_rgba8(device.h, C.c_void_p(d_beauty), C.c_void_p(d_rgba8),
                                  pixel_count))


class Denoiser:
    """The learned denoiser (BF16 MFMA CNN); host-array or device-pointer use."""

    def __init__(self, ordinal: int = 0, params: np.ndarray | None = None):
        self.h = C.c_void_p()
        check(lib().rt_nn_create(ordinal, C.byref(self.h)))
        self.n = lib().rt_nn_param_count()
        if params is not None:
            self.set_params(params)

    def close(self):
        if self.h:
            lib().rt_nn_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tiling(self, conv_rows):
        t = abi.NnTiling()
        t.conv_rows[:] = list(conv_rows)
        check(lib().rt_nn_set_tiling(self.h, C.byref(t)))

    def tiling(self) -> list:
        t = abi.NnTiling()
        check(lib().rt_nn_get_tiling(self.h, C.byref(t)))
        return list(t.conv_rows)

    def autotune(self, sizes, reps: int = 5) -> dict:
        """Searches the convolution tile rows per level for images of these
        [(width, height), ...] and applies the fastest."""
        im = abi.nn_images(sizes)
        res = abi.NnAutotuneResult()
        check(lib().rt_nn_autotune(self.h, C.byref(im), reps, C.byref(res)))
        return {"conv_rows": list(res.best.conv_rows), "best_ms": res.best_ms,
                "baseline_ms": res.baseline_ms, "trials": res.trials}

    def set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == self.n
        check(lib().rt_nn_set_params(self.h, ptr(p), self.n))

    def params(self) -> np.ndarray:
        p = np.zeros(self.n, np.float32)
        check(lib().rt_nn_get_params(self.h, ptr(p), self.n))
        return p

    @staticmethod
    def _frame(frame):
        f32 = lambda a: np.ascontiguousarray(a, np.float32)
        return (f32(frame["beauty"]), f32(frame["normal"]), f32(frame["albedo"]),
                f32(frame["depth"]))

    def infer(self, sizes, frame) -> np.ndarray:
        im = abi.nn_images(sizes)
        b, n, a, d = self._frame(frame)
        out = np.zeros((im.offset[im.count], 4), np.float32)
        check(lib().rt_nn_infer_host(self.h, C.byref(im), ptr(b), ptr(n), ptr(a), ptr(d), ptr(out)))
        return out

    def train_step(self, sizes, frame, target, lr: float, want_grad: bool = False):
        im = abi.nn_images(sizes)
        b, n, a, d = self._frame(frame)
        t = np.ascontiguousarray(target, np.float32)
        loss = C.c_float()
        grad = np.zeros(self.n, np.float32) if want_grad else None
        check(lib().rt_nn_train_step_host(self.h, C.byref(im), ptr(b), ptr(n), ptr(a), ptr(d),
                                          ptr(t), lr, C.byref(loss),
                                          ptr(grad) if grad is not None else None))
        return loss.value, grad


def cache_config(**overrides) -> "abi.CacheConfig":
    """The library's default radiance-cache configuration with overrides."""
    c = abi.CacheConfig()
    lib().rt_cache_default_config(C.byref(c))
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


class RadianceCache:
    """The online radiance cache (hash grid + BF16 MFMA MLP)."""

    def __init__(self, ordinal: int = 0, config=None, **overrides):
        self.config = config if config is not None else cache_config(**overrides)
        self.h = C.c_void_p()
        check(lib().rt_cache_create(ordinal, C.byref(self.config), C.byref(self.h)))
        self.n = lib().rt_cache_param_count(C.byref(self.config))
        self.n_mlp = lib().rt_cache_mlp_param_count()

    def close(self):
        if self.h:
            lib().rt_cache_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == self.n
        check(lib().rt_cache_set_params(self.h, ptr(p), self.n))

    def params(self) -> np.ndarray:
        p = np.zeros(self.n, np.float32)
        check(lib().rt_cache_get_params(self.h, ptr(p), self.n))
        return p

    def set_bounds(self, lo, hi):
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        check(lib().rt_cache_set_bounds(self.h, ptr(lo), ptr(hi)))

    def grid_info(self) -> dict:
        res = np.zeros(16, np.float32)
        dense = np.zeros(16, np.int32)
        lo = np.zeros(3, np.float32)
        inv = np.zeros(3, np.float32)
        check(lib().rt_cache_grid_info(self.h, ptr(res), ptr(dense), ptr(lo), ptr(inv)))
        return {"levels": self.config.levels, "log2_table": self.config.log2_table, "res": res,
                "dense": dense, "lo": lo, "inv_extent": inv}

    def infer(self, inputs: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 16)
        out = np.zeros((x.shape[0], 4), np.float32)
        check(lib().rt_cache_infer_host(self.h, ptr(x), x.shape[0], ptr(out)))
        return out

    def train_step(self, inputs: np.ndarray, targets: np.ndarray, want_grad: bool = False):
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 16)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 4)
        loss = C.c_float()
        grad = np.zeros(self.n, np.float32) if want_grad else None
        check(lib().rt_cache_train_step_host(self.h, ptr(x), ptr(t), x.shape[0], C.byref(loss),
                                             ptr(grad) if grad is not None else None))
        return loss.value, grad

    def stats(self) -> dict:
        st = abi.CacheStats()
        check(lib().rt_cache_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in abi.CacheStats._fields_}


class Generated:
    """A seeded benchmark scene from the library's generator (rt_scene_generate)."""

    CONFIGS = {"random_spheres": 0, "cornell_mesh": 1, "instanced_mesh_restir": 2,
               "radiance_cache_realtime": 3, "cornell_random": 4, "smoke": 5, "final_room": 6}

    def __init__(self, config, seed: int = 0):
        if isinstance(config, str):
            config = self.CONFIGS[config]
        self.config = config
        self.seed = seed
        self.p = C.POINTER(GeneratedScene)()
        check(lib().rt_scene_generate(config, seed, C.byref(self.p)))
        self.g = self.p.contents

    def __del__(self):
        try:
            lib().rt_scene_generated_free(self.p)
        except Exception:
            pass

    @property
    def desc(self) -> SceneDesc:
        return self.g.scene

    @property
    def profile(self) -> Profile:
        p = Profile()
        C.memmove(C.byref(p), C.byref(self.g.profile), C.sizeof(Profile))
        return p

    def rig_at(self, frame: int, frames: int = 60) -> CameraRig:
        r = CameraRig()
        check(lib().rt_scene_generate_rig(self.config, self.seed, frame, frames, C.byref(r)))
        return r

    def rig(self, count: int | None = None) -> CameraRig:
        r = CameraRig()
        n = self.g.camera_count if count is None else count
        r.active_count = n
        for i in range(n):
            C.memmove(C.byref(r.cameras[i]), C.byref(self.g.cameras[i]), C.sizeof(Camera))
        return r

    def arrays(self) -> dict:
        """Copies of the scene arrays as raw numpy byte records."""
        d = self.g.scene

        def grab(p, n, ctype):
            if n == 0:
                return np.zeros(0, dtype=np.dtype((np.void, C.sizeof(ctype))))
            raw = C.string_at(p, n * C.sizeof(ctype))
            return np.frombuffer(raw, dtype=np.dtype((np.void, C.sizeof(ctype)))).copy()

        return {
            "spheres": grab(d.spheres, d.sphere_count, abi.Sphere),
            "quads": grab(d.quads, d.quad_count, abi.Quad),
            "triangles": grab(d.triangles, d.triangle_count, abi.Triangle),
            "materials": grab(d.materials, d.material_count, abi.Material),
            "openpbr": grab(d.openpbr, d.openpbr_count, abi.OpenPbr),
            "analytic": grab(d.analytic_lights, d.analytic_light_count, abi.AnalyticLight),
        }


__all__ = ["Device", "Scene", "Renderer", "Denoiser", "Generated", "RtError", "lib", "abi", "EXPORTED",
           "RAY_DTYPE", "HIT_DTYPE", "NODE_DTYPE", "Camera", "CameraRig", "Profile",
           "FrameTiming", "SceneDesc", "Tuning"]
