"""orion_amd: MI355X-native (gfx950) path-tracing + SPPM hot path for orion.

Python access to the C ABI in include/orion_amd.h (liborion_amd.so, built
in-tree by ``python -m orion_amd.build``).  The library is the product; this
module only marshals numpy arrays across the boundary.  There is no CPU
fallback: every render / intersection call runs the HIP kernels and raises
if the library or a GPU is unavailable.
"""
import atexit
import ctypes as C
import os
import weakref
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
# ORION_AMD_LIB names another build of the library (the sanitizer build in
# orion_amd/_asan/, for instance); the default is the in-tree one
LIB_PATH = Path(os.environ["ORION_AMD_LIB"]).resolve() if os.environ.get("ORION_AMD_LIB") else _PKG / "liborion_amd.so"
MAT_FLOATS = 24  # ORION_MAT_FLOATS: floats per material record

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class OrionError(RuntimeError):
    pass


class SceneInfo(C.Structure):
    _fields_ = [("n_tris", C.c_int), ("n_mats", C.c_int), ("n_lights", C.c_int),
                ("width", C.c_int), ("height", C.c_int), ("integrator", C.c_int),
                ("max_depth", C.c_int), ("spp", C.c_int), ("sppm_iterations", C.c_int),
                ("sppm_photons", C.c_int), ("sppm_radius", C.c_float)]


class SceneStats(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_tris", C.c_int64), ("node_bytes", C.c_int64),
                ("tri_bytes", C.c_int64), ("shade_bytes", C.c_int64), ("bvh_depth", C.c_int),
                ("node_format", C.c_int), ("build_ms", C.c_double), ("upload_ms", C.c_double),
                ("stack_bound", C.c_int), ("trace_waves_per_simd", C.c_int)]


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_int64), ("closest_rays", C.c_int64), ("shadow_rays", C.c_int64),
                ("ref_equiv_rays", C.c_int64), ("photons", C.c_int64), ("deposits", C.c_int64),
                ("ms", C.c_double), ("node_visits_closest", C.c_int64),
                ("tri_tests_closest", C.c_int64), ("node_visits_shadow", C.c_int64),
                ("tri_tests_shadow", C.c_int64), ("trace_closest_ms", C.c_double),
                ("trace_shadow_ms", C.c_double), ("shade_ms", C.c_double),
                ("trace_closest_launches", C.c_int64), ("trace_shadow_launches", C.c_int64),
                ("shade_launches", C.c_int64), ("stack_overflows", C.c_int64),
                ("trace_pair_ms", C.c_double), ("trace_pair_launches", C.c_int64),
                ("shaded_segments", C.c_int64), ("photon_rays", C.c_int64),
                ("node_visits_photon", C.c_int64), ("tri_tests_photon", C.c_int64),
                ("photon_lookups", C.c_int64), ("photon_vp_scanned", C.c_int64),
                ("photon_trace_ms", C.c_double), ("photon_shade_ms", C.c_double),
                ("photon_trace_launches", C.c_int64), ("photon_shade_launches", C.c_int64),
                ("photon_hits", C.c_int64), ("photon_batch", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("spp_per_wave", C.c_int), ("wave_paths", C.c_int), ("rays_per_wave", C.c_int),
                ("use_graph", C.c_int), ("profile", C.c_int), ("lds_ring", C.c_int),
                ("queue_segment", C.c_int), ("split_trace", C.c_int), ("trace_plan", C.c_int)]


class SppmDebug(C.Structure):
    _fields_ = [("debug_iter", C.c_int), ("vp_cell_count", C.c_void_p),
                ("photon_cell_count", C.c_void_p), ("M", C.c_void_p), ("grid_info", C.c_void_p),
                ("grid_bounds", C.c_void_p), ("pairs", C.c_void_p), ("max_pairs", C.c_int64),
                ("cell_range", C.c_void_p), ("profile", C.c_int)]


class JobDesc(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("width", C.c_int), ("height", C.c_int), ("spp", C.c_int),
                ("max_depth", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("jobs_done", C.c_int64), ("waves", C.c_int64), ("paths", C.c_int64),
                ("closest_rays", C.c_int64), ("shadow_rays", C.c_int64), ("active_jobs", C.c_int),
                ("queued_jobs", C.c_int), ("hbm_in_use", C.c_int64), ("busy_ms", C.c_double),
                ("mixed_waves", C.c_int64), ("max_jobs_in_wave", C.c_int), ("hbm_peak", C.c_int64)]


# name -> (restype, argtypes); the header's full export list
_SIGS = {
    "orion_last_error": (C.c_char_p, []),
    "orion_version": (C.c_char_p, []),
    "orion_hscene_load_ori": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "orion_hscene_generate": (C.c_int, [C.c_char_p, C.c_uint64, C.c_void_p, C.c_int,
                                        C.POINTER(C.c_void_p)]),
    "orion_hscene_from_records": (C.c_int, [C.c_int, f32p, i32p, i32p, C.c_int, f32p, C.c_int,
                                            f32p, f32p, C.POINTER(C.c_void_p)]),
    "orion_hscene_info": (C.c_int, [C.c_void_p, C.POINTER(SceneInfo)]),
    "orion_hscene_records": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_void_p)] * 6),
    "orion_hscene_write_ori": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "orion_hscene_destroy": (None, [C.c_void_p]),
    "orion_hscene_bvh_check": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]),
    "orion_camera_record": (C.c_int, [f32p, C.c_float, C.c_int, C.c_int, f32p]),
    "orion_create_transform": (C.c_int, [C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p),
                                         f32p, f32p]),
    "orion_light_record": (C.c_int, [C.c_int, f32p, f32p, f32p, C.c_float, C.c_float, f32p]),
    "orion_hscene_add_envmap": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, f32p, f32p, f32p]),
    "orion_hscene_envmap_image": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                            C.POINTER(C.c_void_p)]),
    "orion_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "orion_ctx_destroy": (C.c_int, [C.c_void_p]),
    "orion_ctx_synchronize": (C.c_int, [C.c_void_p]),
    "orion_selfcheck_rcp": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "orion_scene_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "orion_scene_destroy": (C.c_int, [C.c_void_p]),
    "orion_scene_get_stats": (C.c_int, [C.c_void_p, C.POINTER(SceneStats)]),
    "orion_intersect_closest": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "orion_intersect_any": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "orion_intersect_closest_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                                  C.c_void_p, C.c_void_p, C.c_void_p]),
    "orion_render_path": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(Tuning),
                                    C.POINTER(RenderStats)]),
    "orion_film_clear": (C.c_int, [C.c_void_p]),
    "orion_film_read": (C.c_int, [C.c_void_p, f32p, i32p]),
    "orion_film_read_averaged": (C.c_int, [C.c_void_p, f32p]),
    "orion_film_average": (C.c_int, [C.c_int, C.c_int, f32p, i32p, f32p]),
    "orion_film_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, f32p]),
    "orion_render_sppm": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                    C.POINTER(SppmDebug), C.POINTER(RenderStats)]),
    "orion_render_batch": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(JobDesc),
                                     C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                     C.POINTER(Tuning), C.POINTER(RenderStats)]),
    "orion_server_create": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Tuning),
                                      C.POINTER(C.c_void_p)]),
    "orion_server_set_policy": (C.c_int, [C.c_void_p, C.c_int]),
    "orion_server_submit": (C.c_int, [C.c_void_p, C.POINTER(JobDesc), C.POINTER(C.c_uint64)]),
    "orion_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int), C.c_void_p,
                                    C.c_void_p, C.POINTER(C.c_double)]),
    "orion_server_peek": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]),
    "orion_server_drain": (C.c_int, [C.c_void_p]),
    "orion_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "orion_server_destroy": (C.c_int, [C.c_void_p]),
    "orion_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(JobDesc), C.POINTER(Tuning)]),
}

_lib = None

# Live servers, device scenes and devices.  At interpreter exit they are
# closed here, servers first, then scenes, then their devices, while the
# library and the HIP runtime are intact (module teardown would otherwise
# leave the order and the library's availability to the interpreter).
_LIVE = {"server": weakref.WeakSet(), "scene": weakref.WeakSet(), "device": weakref.WeakSet()}


def _close_all():
    for kind in ("server", "scene", "device"):
        for obj in list(_LIVE[kind]):
            try:
                obj.close()
            except Exception:
                pass


atexit.register(_close_all)


def lib():
    """Loads liborion_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise OrionError(f"{LIB_PATH} is missing: run `python -m orion_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def _check(rc):
    if rc != 0:
        msg = lib().orion_last_error()
        raise OrionError(f"orion_amd error {rc}: {msg.decode() if msg else ''}")


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class HostScene:
    """Resolved scene records on the host (orion_hscene)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        self._refresh()

    @classmethod
    def load_ori(cls, path):
        h = C.c_void_p()
        _check(lib().orion_hscene_load_ori(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def generate(cls, name, seed=0, params=None):
        h = C.c_void_p()
        p = None if params is None else np.ascontiguousarray(params, np.float64)
        _check(lib().orion_hscene_generate(name.encode(), seed, _ptr(p),
                                           0 if p is None else len(p), C.byref(h)))
        return cls(h)

    @classmethod
    def from_records(cls, tri_pn, tri_mat, tri_light, mats, lights, camera):
        tri_pn = np.ascontiguousarray(tri_pn, np.float32).reshape(-1)
        tri_mat = np.ascontiguousarray(tri_mat, np.int32)
        tri_light = np.ascontiguousarray(tri_light, np.int32)
        mats = np.ascontiguousarray(mats, np.float32).reshape(-1)
        lights = np.ascontiguousarray(lights, np.float32).reshape(-1)
        camera = np.ascontiguousarray(camera, np.float32)
        h = C.c_void_p()
        _check(lib().orion_hscene_from_records(len(tri_mat), tri_pn, tri_mat, tri_light,
                                               len(mats) // MAT_FLOATS, mats, len(lights) // 32, lights,
                                               camera, C.byref(h)))
        return cls(h)

    def _refresh(self):
        info = SceneInfo()
        _check(lib().orion_hscene_info(self.h, C.byref(info)))
        self.info = info
        ptrs = [C.c_void_p() for _ in range(6)]
        _check(lib().orion_hscene_records(self.h, *[C.byref(p) for p in ptrs]))
        n, m, l = info.n_tris, info.n_mats, info.n_lights

        def view(p, count, dt):
            if count == 0 or not p.value:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dt))),
                                         shape=(count,)).copy()
        self.tri_pn = view(ptrs[0], 18 * n, np.float32)
        self.tri_mat = view(ptrs[1], n, np.int32)
        self.tri_light = view(ptrs[2], n, np.int32)
        self.mats = view(ptrs[3], MAT_FLOATS * m, np.float32)
        self.lights = view(ptrs[4], 32 * l, np.float32)
        self.camera = view(ptrs[5], 16, np.float32)
        # envmap images (light record [1] indexes them), texels before L
        self.env_images = []
        k = 0
        while True:
            w, h, px = C.c_int(), C.c_int(), C.c_void_p()
            if lib().orion_hscene_envmap_image(self.h, k, C.byref(w), C.byref(h), C.byref(px)) != 0:
                break
            self.env_images.append(view(px, 3 * w.value * h.value, np.float32).reshape(h.value, w.value, 3))
            k += 1

    def add_envmap(self, L, rgb=None, light_to_world=None, world_to_light=None):
        """Appends an environment light: rgb (H, W, 3) texels (byte / 255 values)
        or None for the constant map L; orientation matrices default to identity."""
        eye = np.eye(4, dtype=np.float32)
        l2w = np.ascontiguousarray(eye if light_to_world is None else light_to_world, np.float32)
        w2l = np.ascontiguousarray(eye if world_to_light is None else world_to_light, np.float32)
        if rgb is None:
            hgt, wid, px = 1, 1, None
        else:
            rgb = np.ascontiguousarray(rgb, np.float32)
            hgt, wid = rgb.shape[:2]
            px = rgb
        _check(lib().orion_hscene_add_envmap(self.h, wid, hgt, _ptr(px), np.asarray(L, np.float32).reshape(3),
                                             l2w, w2l))
        self._refresh()

    # oracle_lib.OracleScene reads these names
    @property
    def n_tris(self):
        return self.info.n_tris

    @property
    def n_mats(self):
        return self.info.n_mats

    @property
    def n_lights(self):
        return self.info.n_lights

    @property
    def width(self):
        return self.info.width

    @property
    def height(self):
        return self.info.height

    def bvh_check(self):
        """Builds and structurally checks the device BVH on the host; returns
        (n_nodes, depth, stack_bound) or raises OrionError."""
        nn, d, sb = C.c_int64(), C.c_int(), C.c_int()
        _check(lib().orion_hscene_bvh_check(self.h, C.byref(nn), C.byref(d), C.byref(sb)))
        return nn.value, d.value, sb.value

    def write_ori(self, path, inline_limit=1000):
        _check(lib().orion_hscene_write_ori(self.h, str(path).encode(), inline_limit))

    def __del__(self):
        try:
            if self.h:
                lib().orion_hscene_destroy(self.h)
        except Exception:
            pass


IMAGE_FORMATS = {"bmp": 0, "pfm": 1, "exr": 2}


def film_average(rgb_sum, count):
    """RenderTarget::getAveraged of a host (sum, count) film."""
    rgb_sum = np.ascontiguousarray(rgb_sum, np.float32)
    count = np.ascontiguousarray(count, np.int32)
    h, w = count.shape
    out = np.zeros((h, w, 3), np.float32)
    _check(lib().orion_film_average(w, h, rgb_sum, count, out))
    return out


def write_image(path, rgb, fmt=None):
    """Writes an averaged film (H, W, 3) as BMP (saveAsBmp), PFM or EXR; the
    format defaults to the file extension."""
    fmt = fmt or str(path).rsplit(".", 1)[-1].lower()
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w = rgb.shape[:2]
    _check(lib().orion_film_write(str(path).encode(), IMAGE_FORMATS[fmt], w, h, rgb))


def camera_record(lookat, fov, width, height):
    out = np.zeros(16, np.float32)
    _check(lib().orion_camera_record(np.asarray(lookat, np.float32), fov, width, height, out))
    return out


def create_transform(pairs):
    n = len(pairs)
    keys = (C.c_char_p * max(n, 1))(*[k.encode() for k, _ in pairs])
    vals = (C.c_char_p * max(n, 1))(*[v.encode() for _, v in pairs])
    m = np.zeros(16, np.float32)
    mi = np.zeros(16, np.float32)
    _check(lib().orion_create_transform(n, keys, vals, m, mi))
    return m.reshape(4, 4), mi.reshape(4, 4)


def light_record(typ, I, frm, to=(0, 0, 1), angle=0.0, delta=0.0):
    out = np.zeros(32, np.float32)
    _check(lib().orion_light_record(typ, np.asarray(I, np.float32), np.asarray(frm, np.float32),
                                    np.asarray(to, np.float32), angle, delta, out))
    return out


class Device:
    """One GPU + one HIP stream (orion_ctx)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib().orion_ctx_create(device, C.byref(self.h)))
        _LIVE["device"].add(self)

    def synchronize(self):
        _check(lib().orion_ctx_synchronize(self.h))

    def selfcheck_rcp(self):
        """Floats in [2^-125, 2^125) whose device reciprocal differs from
        IEEE division (0 expected)."""
        n = C.c_uint64()
        _check(lib().orion_selfcheck_rcp(self.h, C.byref(n)))
        return n.value

    def close(self):
        # the library refuses (ORION_ERR_BUSY) while scenes or servers of this
        # context are alive; the handle is then kept for a later close
        if self.h and lib().orion_ctx_destroy(self.h) == 0:
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceScene:
    """A scene resident in HBM (BVH, triangles, shading data, film)."""

    def __init__(self, dev, hscene):
        self.dev = dev
        self.hs = hscene
        self.W, self.H = hscene.width, hscene.height
        self.h = C.c_void_p()
        _check(lib().orion_scene_upload(dev.h, hscene.h, C.byref(self.h)))
        _LIVE["scene"].add(self)

    def stats(self):
        st = SceneStats()
        _check(lib().orion_scene_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def intersect_closest(self, rays_soa):
        rays = np.ascontiguousarray(rays_soa, np.float32)
        n = rays.shape[1]
        t = np.zeros(n, np.float32); prim = np.zeros(n, np.int32)
        u = np.zeros(n, np.float32); v = np.zeros(n, np.float32); tie = np.zeros(n, np.int32)
        _check(lib().orion_intersect_closest(self.h, n, _ptr(rays), _ptr(t), _ptr(prim), _ptr(u),
                                             _ptr(v), _ptr(tie), 0))
        return t, prim, u, v, tie

    def intersect_any(self, rays_soa):
        rays = np.ascontiguousarray(rays_soa, np.float32)
        n = rays.shape[1]
        occ = np.zeros(n, np.int32)
        _check(lib().orion_intersect_any(self.h, n, _ptr(rays), _ptr(occ), 0))
        return occ

    def render_path(self, k0, k1, max_depth, spp_per_wave=0, rays_per_wave=0, profile=False,
                    use_graph=False, wave_paths=0, lds_ring=0, queue_segment=0, split_trace=False,
                    tuning=None, trace_plan=0):
        if tuning is not None:
            spp_per_wave, wave_paths = tuning.spp_per_wave, tuning.wave_paths
            rays_per_wave, use_graph = tuning.rays_per_wave, bool(tuning.use_graph)
            lds_ring, queue_segment = tuning.lds_ring, tuning.queue_segment
            split_trace, trace_plan = bool(tuning.split_trace), tuning.trace_plan
        st = RenderStats()
        tune = None  # NULL: the defaults
        if (spp_per_wave or wave_paths or rays_per_wave or use_graph or profile or lds_ring
                or queue_segment or split_trace or trace_plan):
            tune = C.byref(Tuning(spp_per_wave, wave_paths, rays_per_wave, int(use_graph),
                                  int(profile), lds_ring, queue_segment, int(split_trace), trace_plan))
        _check(lib().orion_render_path(self.h, k0, k1, max_depth, tune, C.byref(st)))
        return st

    def film_clear(self):
        _check(lib().orion_film_clear(self.h))

    def film(self):
        rgb = np.zeros(self.W * self.H * 3, np.float32)
        cnt = np.zeros(self.W * self.H, np.int32)
        _check(lib().orion_film_read(self.h, rgb, cnt))
        return rgb.reshape(self.H, self.W, 3), cnt.reshape(self.H, self.W)

    def film_averaged(self):
        """RenderTarget::getAveraged, computed on the device."""
        rgb = np.zeros(self.W * self.H * 3, np.float32)
        _check(lib().orion_film_read_averaged(self.h, rgb))
        return rgb.reshape(self.H, self.W, 3)

    def render_sppm(self, iters, photons, radius, max_depth, hash_size=0, debug_iter=-1, pairs=False,
                    profile=False):
        """pairs: also return the debug iteration's radix-sorted (key, pixel)
        pairs as an (n_pairs, 2) array and each cell's [start, end) range.
        profile: time the photon pass's launches into the statistics."""
        W, H = self.W, self.H
        hs = hash_size if hash_size > 0 else W * H
        dbg = {}
        dptr = None
        if profile and debug_iter < 0:
            d = SppmDebug(-1)
            d.profile = 1
            dptr = C.byref(d)
        if debug_iter >= 0:
            dbg = dict(vp_cell_count=np.zeros(hs, np.int32), photon_cell_count=np.zeros(hs, np.int32),
                       M=np.zeros(W * H, np.int32), grid_info=np.zeros(8, np.int32),
                       grid_bounds=np.zeros(6, np.float32))
            cap = 27 * W * H if pairs else 0  # a visible point's box spans at most 3 cells per axis
            if pairs:
                dbg["pairs"] = np.zeros(2 * cap, np.int32)
                dbg["cell_range"] = np.zeros(2 * hs, np.int32)
            d = SppmDebug(debug_iter, dbg["vp_cell_count"].ctypes.data,
                          dbg["photon_cell_count"].ctypes.data, dbg["M"].ctypes.data,
                          dbg["grid_info"].ctypes.data, dbg["grid_bounds"].ctypes.data,
                          dbg["pairs"].ctypes.data if pairs else None, cap,
                          dbg["cell_range"].ctypes.data if pairs else None, int(profile))
            dptr = C.byref(d)
        st = RenderStats()
        _check(lib().orion_render_sppm(self.h, iters, photons, radius, max_depth, hash_size, dptr,
                                       C.byref(st)))
        if pairs:
            n = int(dbg["grid_info"][3])
            dbg["pairs"] = dbg["pairs"][:2 * n].reshape(n, 2)
            dbg["cell_range"] = dbg["cell_range"].reshape(hs, 2)
        return st, dbg

    def close(self):
        if self.h:
            lib().orion_scene_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _descs(jobs):
    descs = (JobDesc * max(len(jobs), 1))()
    for i, (sc, W, H, spp, depth) in enumerate(jobs):
        descs[i] = JobDesc(sc.h.value, W, H, spp, depth)
    return descs


def autotune(dev, jobs):
    """Fastest Tuning for the job mix [(DeviceScene, width, height, spp, max_depth), ...]."""
    best = Tuning()
    _check(lib().orion_autotune(dev.h, len(jobs), _descs(jobs), C.byref(best)))
    return best


def render_batch(dev, jobs, wave_paths=0, rays_per_wave=0, use_graph=False, profile=False,
                 lds_ring=0, queue_segment=0, split_trace=False, tuning=None):
    """Renders jobs = [(DeviceScene, width, height, spp, max_depth), ...] in shared
    wavefronts; returns ([(rgb HxWx3, count HxW), ...], RenderStats)."""
    if tuning is not None:
        wave_paths, rays_per_wave = tuning.wave_paths, tuning.rays_per_wave
        use_graph, lds_ring = bool(tuning.use_graph), tuning.lds_ring
        queue_segment, split_trace = tuning.queue_segment, bool(tuning.split_trace)
    n = len(jobs)
    descs = _descs(jobs)
    outs = []
    rgb_p = (C.c_void_p * max(n, 1))()
    cnt_p = (C.c_void_p * max(n, 1))()
    for i, (sc, W, H, spp, depth) in enumerate(jobs):
        rgb = np.zeros((H, W, 3), np.float32)
        cnt = np.zeros((H, W), np.int32)
        outs.append((rgb, cnt))
        rgb_p[i] = rgb.ctypes.data
        cnt_p[i] = cnt.ctypes.data
    st = RenderStats()
    tune = Tuning(0, wave_paths, rays_per_wave, int(use_graph), int(profile), lds_ring,
                  queue_segment, int(split_trace))
    _check(lib().orion_render_batch(dev.h, n, descs, rgb_p, cnt_p, C.byref(tune), C.byref(st)))
    return outs, st


class Server:
    """Request loop over one device: submit jobs, poll or drain, read films and
    per-request times (submitted, admitted, finished; seconds since start)."""

    def __init__(self, dev, hbm_budget_bytes, wave_paths=0, use_graph=False, rays_per_wave=0,
                 lds_ring=0, queue_segment=0, split_trace=False, tuning=None):
        if tuning is not None:
            wave_paths, rays_per_wave = tuning.wave_paths, tuning.rays_per_wave
            use_graph, lds_ring = bool(tuning.use_graph), tuning.lds_ring
            queue_segment, split_trace = tuning.queue_segment, bool(tuning.split_trace)
        self.dev = dev
        self.h = C.c_void_p()
        self._shape = {}
        tune = Tuning(0, wave_paths, rays_per_wave, int(use_graph), 0, lds_ring, queue_segment,
                      int(split_trace))
        _check(lib().orion_server_create(dev.h, int(hbm_budget_bytes), C.byref(tune), C.byref(self.h)))
        _LIVE["server"].add(self)

    PACK, FAIR, ARRIVAL, AGED = 0, 1, 2, 3

    def set_policy(self, policy):
        _check(lib().orion_server_set_policy(self.h, policy))

    def submit(self, scene, width, height, spp, max_depth):
        jid = C.c_uint64()
        d = JobDesc(scene.h.value, width, height, spp, max_depth)
        _check(lib().orion_server_submit(self.h, C.byref(d), C.byref(jid)))
        self._shape[jid.value] = (height, width)
        return jid.value

    def poll(self, jid, want_film=True):
        """Returns None while running, else (rgb, count, times)."""
        H, W = self._shape[jid]
        done = C.c_int()
        times = (C.c_double * 3)()
        rgb = np.zeros((H, W, 3), np.float32) if want_film else None
        cnt = np.zeros((H, W), np.int32) if want_film else None
        _check(lib().orion_server_poll(self.h, jid, C.byref(done), _ptr(rgb), _ptr(cnt), times))
        if not done.value:
            return None
        return rgb, cnt, tuple(times)

    def peek(self, jid):
        """Progressive preview: (rgb sum, count, samples rendered so far)."""
        H, W = self._shape[jid]
        rgb = np.zeros((H, W, 3), np.float32)
        cnt = np.zeros((H, W), np.int32)
        k = C.c_int()
        _check(lib().orion_server_peek(self.h, jid, _ptr(rgb), _ptr(cnt), C.byref(k)))
        return rgb, cnt, k.value

    def drain(self):
        _check(lib().orion_server_drain(self.h))

    def stats(self):
        st = ServerStats()
        _check(lib().orion_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.h:
            _check(lib().orion_server_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
