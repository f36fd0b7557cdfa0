"""Python binding of libmcpt.so, the gfx950 path-tracing core (include/mcpt.h).

The library is loaded from this directory. There is no CPU fallback: if
libmcpt.so is missing the import fails, and every call that reaches the GPU
raises McptError when no gfx950 device is present.

Typical use::

    import importlib; mcpt = importlib.import_module("monocuda-pt_amd")
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    with mcpt.Context() as ctx:
        scene = ctx.scene(hot, cold, mats)
        scene.build()
        cam = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_CORNELL, 256, 256))
        film = ctx.film(256, 256)
        stats = scene.render(cam, film, mode=mcpt.MODE_PERSISTENT, spp=16, max_depth=8)
        rgb = film.read()
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmcpt.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `make -C monocuda-pt_amd` or __graft_entry__.build()")
_lib = C.CDLL(LIB_PATH)

# ------------------------------------------------------------------ constants
OK, DEFERRED, PENDING = 0, 1, 2
ERR_INVALID, ERR_HIP, ERR_BUDGET, ERR_NODEVICE = -1, -2, -3, -4
MAT_LAMBERTIAN, MAT_DIELECTRIC, MAT_CONDUCTOR, MAT_EMISSIVE, MAT_ISOTROPIC = 0, 1, 2, 3, 4
MODE_MEGAKERNEL, MODE_WAVEFRONT, MODE_PERSISTENT = 0, 1, 2
TONEMAP_NONE, TONEMAP_REINHARD, TONEMAP_ACES = 0, 1, 2
RENDER_COUNT_TRAVERSAL, RENDER_TIME_KERNELS, RENDER_REFERENCE_ORDER, RENDER_PERSISTENT_SHADOW = 1, 2, 4, 8
RENDER_SHADOW_PER_RAY = 16
NOALLOC_MAX_PATHS = 299584  # mcpt_render_noalloc: paths per pass in the fixed workspace
TEX_CHECKER, TEX_IMAGE = 1, 2
GEN_CORNELL, GEN_SOUP, GEN_ICOSPHERE, GEN_MEDIUM = 1, 2, 3, 5
MLP_FEATURES, MLP_INPUTS, MLP_WIDTH, MLP_HIDDEN, MLP_OUTPUTS = 11, 64, 64, 4, 3
MLP_PARAMS = 4 * 64 * 64 + 3 * 64
SCHED_SHORTEST_FIRST, SCHED_FIFO, SCHED_SHORTEST_FIRST_AGED = 0, 1, 2

# numpy layouts of the C records (byte-identical)
HOT_DTYPE = np.dtype([("v0", "<f4", 3), ("v1", "<f4", 3), ("v2", "<f4", 3)])
COLD_DTYPE = np.dtype([("uv0", "<f4", 2), ("uv1", "<f4", 2), ("uv2", "<f4", 2), ("n0", "<f4", 3),
                       ("n1", "<f4", 3), ("n2", "<f4", 3), ("material", "<i4")])
MAT_DTYPE = np.dtype([("type", "<i4"), ("albedo", "<f4", 3), ("emission", "<f4", 3), ("fuzz", "<f4"),
                      ("ior", "<f4"), ("density", "<f4"), ("phase_g", "<f4"), ("texture", "<i4")])
SPHERE_DTYPE = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("center1", "<f4", 3), ("moving", "<i4"),
                         ("material", "<i4"), ("pad", "<i4", 3)])
LIGHT_DTYPE = np.dtype([("tri", HOT_DTYPE), ("emission", "<f4", 3), ("area", "<f4"), ("normal", "<f4", 3)])
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("time", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim_type", "<i4"), ("prim_index", "<i4"), ("did_hit", "<i4")])
AABB_DTYPE = np.dtype([("min", "<f4", 3), ("pad0", "<f4"), ("max", "<f4", 3), ("pad1", "<f4")])
# four-wide node of mcpt_lbvh_dump_wide (64 bytes)
WIDE_DTYPE = np.dtype([("origin", "<f4", 3), ("meta", "<u4"), ("q", "u1", (6, 4)), ("pad", "u1", 8),
                       ("child", "<i4", 4)])
assert WIDE_DTYPE.itemsize == 64
assert HOT_DTYPE.itemsize == 36 and COLD_DTYPE.itemsize == 64 and MAT_DTYPE.itemsize == 48
assert LIGHT_DTYPE.itemsize == 64 and RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16
assert AABB_DTYPE.itemsize == 32


class CameraDesc(C.Structure):
    _fields_ = [("look_from", C.c_float * 3), ("look_at", C.c_float * 3), ("up", C.c_float * 3),
                ("vfov_degrees", C.c_float), ("aperture", C.c_float), ("focus_distance", C.c_float),
                ("width", C.c_int32), ("height", C.c_int32), ("background", C.c_float * 3), ("rr_cap", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("pixel00", C.c_float * 3), ("delta_u", C.c_float * 3),
                ("delta_v", C.c_float * 3), ("u", C.c_float * 3), ("v", C.c_float * 3), ("w", C.c_float * 3),
                ("lens_radius", C.c_float), ("width", C.c_int32), ("height", C.c_int32),
                ("background", C.c_float * 3), ("rr_cap", C.c_float)]


class EnvDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("intensity", C.c_float),
                ("rotation_radians", C.c_float), ("rgba", C.c_void_p)]


class LbvhDump(C.Structure):
    _fields_ = [("morton", C.c_void_p), ("sorted_keys", C.c_void_p), ("sorted_ids", C.c_void_p),
                ("left", C.c_void_p), ("right", C.c_void_p), ("left_is_leaf", C.c_void_p),
                ("right_is_leaf", C.c_void_p), ("parent", C.c_void_p), ("leaf_parent", C.c_void_p),
                ("internal_bounds", C.c_void_p), ("leaf_bounds", C.c_void_p), ("scene_bounds", C.c_void_p),
                ("max_depth", C.c_int32), ("build_ms", C.c_float)]


class RenderOpts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("spp", C.c_int32), ("sample_begin", C.c_int32), ("max_depth", C.c_int32),
                ("seed", C.c_uint64), ("refill_lanes", C.c_int32), ("flags", C.c_uint32), ("batch", C.c_int32),
                ("group_below", C.c_int32)]


class RenderStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("camera_ms", C.c_double), ("extend_ms", C.c_double),
                ("shade_ms", C.c_double), ("shadow_ms", C.c_double), ("finalize_ms", C.c_double),
                ("extend_launches", C.c_uint64), ("extend_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("paths", C.c_uint64), ("extend_steps", C.c_uint64), ("extend_leaves", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ObjOpts(C.Structure):
    _fields_ = [("translation", C.c_float * 3), ("uniform_scale", C.c_float), ("rotation_y", C.c_float),
                ("fallback", C.c_uint8 * MAT_DTYPE.itemsize), ("load_textures", C.c_int32)]


class ObjStats(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("n_tris", "n_materials", "skipped_faces", "texture_requests",
                                          "missing_libraries", "textures_loaded", "missing_textures",
                                          "reserved")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class PbrtInfo(C.Structure):
    _fields_ = [("n_tris", C.c_int32), ("n_spheres", C.c_int32), ("n_materials", C.c_int32),
                ("textures_loaded", C.c_int32), ("has_camera", C.c_int32), ("width", C.c_int32),
                ("height", C.c_int32), ("look_from", C.c_float * 3), ("look_at", C.c_float * 3), ("up", C.c_float * 3),
                ("vfov_degrees", C.c_float), ("background", C.c_float * 3), ("has_env_map", C.c_int32),
                ("env_intensity", C.c_float), ("env_rotation_degrees", C.c_float), ("env_path", C.c_char * 1024),
                ("shapes", C.c_int32), ("instances", C.c_int32), ("skipped_shapes", C.c_int32),
                ("approximated_lights", C.c_int32), ("unsupported_textures", C.c_int32),
                ("approximated_materials", C.c_int32)]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = v.decode(errors="replace") if k == "env_path" else (list(v) if hasattr(v, "__len__") else v)
        return out


class DensityGrid(C.Structure):
    """mcpt_density_grid (48 bytes): dense voxels, x fastest."""
    _fields_ = [("dims", C.c_int32 * 3), ("origin", C.c_float * 3), ("voxel_size", C.c_float * 3),
                ("reserved", C.c_int32), ("density", C.c_void_p)]


def _density_grid_desc(density, origin, voxel_size):
    keep = np.ascontiguousarray(density, np.float32)
    if keep.ndim != 3:
        raise ValueError("density must be a (nz, ny, nx) array")
    g = DensityGrid()
    g.dims[:] = [keep.shape[2], keep.shape[1], keep.shape[0]]
    g.origin[:] = [float(x) for x in origin]
    g.voxel_size[:] = [float(x) for x in (voxel_size if np.ndim(voxel_size) else (voxel_size,) * 3)]
    g.density = keep.ctypes.data
    return g, keep


class TextureDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("color_a", C.c_float * 3), ("color_b", C.c_float * 3), ("scale", C.c_float),
                ("width", C.c_int32), ("height", C.c_int32), ("rgba8", C.c_void_p)]


def _texture_desc(kind, color_a=(0, 0, 0), color_b=(1, 1, 1), scale=1.0, image=None):
    d = TextureDesc()
    d.type = kind
    d.color_a[:] = [float(x) for x in color_a]
    d.color_b[:] = [float(x) for x in color_b]
    d.scale = scale
    keep = None
    if image is not None:
        keep = np.ascontiguousarray(image, np.uint8)
        assert keep.ndim == 3 and keep.shape[2] == 4, "image must be (H, W, 4) RGBA8"
        d.height, d.width = keep.shape[:2]
        d.rgba8 = keep.ctypes.data
    return d, keep


class GenSizes(C.Structure):
    _fields_ = [("n_tris", C.c_int32), ("n_materials", C.c_int32)]


class BatchTune(C.Structure):
    _fields_ = [("samples_per_pass", C.c_int32), ("refill_lanes", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("reserved", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CacheOpts(C.Structure):
    _fields_ = [("train_paths", C.c_int32), ("train_batch", C.c_int32), ("train_steps", C.c_int32),
                ("terminate_depth", C.c_int32), ("lr", C.c_float), ("ema_decay", C.c_float),
                ("frame_spp", C.c_int32)]


class CacheStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("train_trace_ms", C.c_double), ("train_ms", C.c_double),
                ("render_ms", C.c_double), ("infer_ms", C.c_double), ("train_samples", C.c_uint64),
                ("queries", C.c_uint64), ("paths", C.c_uint64), ("loss", C.c_float), ("reserved", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerOpts(C.Structure):
    _fields_ = [("path_capacity", C.c_int32), ("max_jobs", C.c_int32), ("frames_per_job", C.c_int32),
                ("refill_lanes", C.c_int32), ("film_budget_bytes", C.c_uint64), ("policy", C.c_int32),
                ("group_below", C.c_int32)]


class Job(C.Structure):
    _fields_ = [("scene", C.c_int32), ("spp", C.c_int32), ("max_depth", C.c_int32), ("reserved", C.c_int32),
                ("seed", C.c_uint64), ("camera", Camera)]


class JobResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("submit_ms", C.c_double), ("start_ms", C.c_double), ("finish_ms", C.c_double),
                ("paths", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [("iterations", C.c_uint64), ("jobs_submitted", C.c_uint64), ("jobs_done", C.c_uint64),
                ("jobs_waiting", C.c_uint64), ("jobs_running", C.c_uint64), ("paths_started", C.c_uint64),
                ("extend_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("film_bytes_in_use", C.c_uint64),
                ("film_budget", C.c_uint64), ("busy_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


VP = C.c_void_p
_SIGS = {
    "mcpt_last_error": ([], C.c_char_p),
    "mcpt_version": ([], C.c_char_p),
    "mcpt_ctx_create": ([C.c_int, C.c_uint64, C.POINTER(VP)], C.c_int),
    "mcpt_ctx_destroy": ([VP], C.c_int),
    "mcpt_ctx_memory": ([VP, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], C.c_int),
    "mcpt_ctx_describe": ([VP, C.c_char_p, C.c_size_t], C.c_int),
    "mcpt_scene_create": ([VP, VP, VP, C.c_int32, VP, C.c_int32, C.POINTER(EnvDesc), C.POINTER(VP)], C.c_int),
    "mcpt_scene_create_ex": ([VP, VP, VP, C.c_int32, VP, C.c_int32, VP, C.c_int32, C.POINTER(EnvDesc),
                              C.POINTER(VP)], C.c_int),
    "mcpt_scene_destroy": ([VP], C.c_int),
    "mcpt_scene_light_count": ([VP, C.POINTER(C.c_int32)], C.c_int),
    "mcpt_scene_lights": ([VP, VP], C.c_int),
    "mcpt_scene_env_tables": ([VP, VP, VP, C.POINTER(C.c_float)], C.c_int),
    "mcpt_lbvh_build": ([VP, C.POINTER(LbvhDump)], C.c_int),
    "mcpt_lbvh_dump_wide": ([VP, VP, C.c_int32, C.POINTER(C.c_int32), VP], C.c_int),
    "mcpt_trace_closest": ([VP, VP, C.c_int32, VP, VP, VP], C.c_int),
    "mcpt_trace_any": ([VP, VP, C.c_int32, VP], C.c_int),
    "mcpt_camera_init": ([C.POINTER(CameraDesc), C.POINTER(Camera)], C.c_int),
    "mcpt_film_create": ([VP, C.c_int32, C.c_int32, C.POINTER(VP)], C.c_int),
    "mcpt_film_destroy": ([VP], C.c_int),
    "mcpt_film_clear": ([VP], C.c_int),
    "mcpt_film_read": ([VP, VP], C.c_int),
    "mcpt_film_write": ([VP, VP], C.c_int),
    "mcpt_render": ([VP, C.POINTER(Camera), C.POINTER(RenderOpts), VP, C.POINTER(RenderStats)], C.c_int),
    "mcpt_render_noalloc": ([VP, C.POINTER(Camera), C.POINTER(RenderOpts), VP], C.c_int),
    "mcpt_resolve": ([VP, C.c_int32, C.c_int32, C.c_float, VP], C.c_int),
    "mcpt_resolve_rgb8": ([VP, C.c_int32, C.c_int32, C.c_float, VP], C.c_int),
    "mcpt_auto_exposure": ([VP, C.c_int32, C.c_float, C.POINTER(C.c_float)], C.c_int),
    "mcpt_write_png": ([C.c_char_p, C.c_int32, C.c_int32, VP], C.c_int),
    "mcpt_server_create": ([VP, VP, C.c_int32, C.POINTER(ServerOpts), C.POINTER(VP)], C.c_int),
    "mcpt_server_submit": ([VP, C.POINTER(Job), C.POINTER(C.c_uint64)], C.c_int),
    "mcpt_server_poll": ([VP, C.c_uint64, C.POINTER(JobResult)], C.c_int),
    "mcpt_server_wait": ([VP, C.c_uint64, C.POINTER(JobResult), VP], C.c_int),
    "mcpt_server_release": ([VP, C.c_uint64], C.c_int),
    "mcpt_server_stats_get": ([VP, C.POINTER(ServerStats)], C.c_int),
    "mcpt_server_destroy": ([VP], C.c_int),
    "mcpt_lbvh_build_batch": ([VP, C.c_int32, C.POINTER(C.c_float)], C.c_int),
    "mcpt_render_batch": ([VP, VP, C.c_int32, C.POINTER(RenderOpts), VP, C.POINTER(BatchTune),
                           C.POINTER(RenderStats)], C.c_int),
    "mcpt_batch_autotune": ([VP, VP, C.c_int32, C.POINTER(RenderOpts), C.POINTER(BatchTune),
                             C.POINTER(C.c_double)], C.c_int),
    "mcpt_mlp_create": ([VP, C.c_uint64, C.POINTER(VP)], C.c_int),
    "mcpt_mlp_destroy": ([VP], C.c_int),
    "mcpt_mlp_get_params": ([VP, VP], C.c_int),
    "mcpt_mlp_set_params": ([VP, VP], C.c_int),
    "mcpt_mlp_forward": ([VP, VP, C.c_int32, VP], C.c_int),
    "mcpt_mlp_gradient": ([VP, VP, VP, C.c_int32, C.POINTER(C.c_float), VP], C.c_int),
    "mcpt_mlp_train": ([VP, VP, VP, C.c_int32, C.c_int32, C.c_float, C.POINTER(C.c_float)], C.c_int),
    "mcpt_render_cached": ([VP, C.POINTER(Camera), C.POINTER(RenderOpts), VP, VP, C.POINTER(CacheOpts),
                            C.POINTER(CacheStats)], C.c_int),
    "mcpt_gen_scene": ([C.c_int32, C.c_uint64, C.c_int64, VP, VP, VP, C.POINTER(GenSizes)], C.c_int),
    "mcpt_load_obj": ([C.c_char_p, C.POINTER(ObjOpts), VP, VP, VP, C.POINTER(ObjStats)], C.c_int),
    "mcpt_scene_add_texture": ([VP, C.POINTER(TextureDesc), C.POINTER(C.c_int32)], C.c_int),
    "mcpt_scene_set_density_grid": ([VP, C.POINTER(DensityGrid), C.POINTER(C.c_float)], C.c_int),
    "mcpt_load_ppm": ([C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), VP], C.c_int),
    "mcpt_load_png": ([C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), VP], C.c_int),
    "mcpt_load_tga": ([C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), VP], C.c_int),
    "mcpt_load_jpeg": ([C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), VP], C.c_int),
    "mcpt_load_pfm": ([C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), VP], C.c_int),
    "mcpt_load_hdr": ([C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), VP], C.c_int),
    "mcpt_load_exr": ([C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), VP], C.c_int),
    "mcpt_load_nvdb": ([C.c_char_p, C.POINTER(DensityGrid), VP, C.POINTER(C.c_float)], C.c_int),
    "mcpt_obj_texture": ([C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), VP], C.c_int),
    "mcpt_load_pbrt": ([C.c_char_p, C.c_int64, VP, VP, VP, VP, C.POINTER(PbrtInfo)], C.c_int),
    "mcpt_pbrt_texture": ([C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), VP], C.c_int),
    "mcpt_gen_sky": ([C.c_int32, C.c_int32, VP], C.c_int),
    "mcpt_gen_camera": ([C.c_int32, C.c_int32, C.c_int32, C.POINTER(CameraDesc)], C.c_int),
}
EXPORTS = sorted(_SIGS)
for _name, (_args, _res) in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.argtypes = _args
    _f.restype = _res


class McptError(RuntimeError):
    def __init__(self, code, where):
        msg = _lib.mcpt_last_error().decode(errors="replace")
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


def _check(code, where):
    if code < 0:
        raise McptError(code, where)
    return code


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def version():
    return _lib.mcpt_version().decode()


def _handles(objs):
    return (VP * len(objs))(*[o.h for o in objs])


def _cameras(cams):
    return (Camera * len(cams))(*cams)


def build_batch(scenes):
    """mcpt_lbvh_build_batch: every LBVH stage as one launch over all scenes;
    returns the device time in ms."""
    ms = C.c_float()
    _check(_lib.mcpt_lbvh_build_batch(_handles(scenes), len(scenes), C.byref(ms)), "mcpt_lbvh_build_batch")
    return ms.value


def render_batch(scenes, cameras, films, spp=1, sample_begin=0, max_depth=8, seed=0, tune=None):
    """mcpt_render_batch: scene i through cameras[i] into films[i], one wavefront for all."""
    o = RenderOpts(MODE_PERSISTENT, spp, sample_begin, max_depth, seed, 0, 0, 0, 0)
    st = RenderStats()
    t = C.byref(tune) if tune is not None else None
    _check(_lib.mcpt_render_batch(_handles(scenes), _cameras(cameras), len(scenes), C.byref(o), _handles(films), t,
                                  C.byref(st)), "mcpt_render_batch")
    return st.as_dict()


def batch_autotune(scenes, cameras, spp=1, sample_begin=0, max_depth=8, seed=0):
    """mcpt_batch_autotune: returns (BatchTune, search device ms)."""
    o = RenderOpts(MODE_PERSISTENT, spp, sample_begin, max_depth, seed, 0, 0, 0, 0)
    best = BatchTune()
    ms = C.c_double()
    _check(_lib.mcpt_batch_autotune(_handles(scenes), _cameras(cameras), len(scenes), C.byref(o), C.byref(best),
                                    C.byref(ms)), "mcpt_batch_autotune")
    return best, ms.value


# ------------------------------------------------------------ host helpers
def gen_scene(kind, seed=1, param=0):
    """Procedural scene -> (hot, cold, materials) numpy record arrays."""
    sz = GenSizes()
    _check(_lib.mcpt_gen_scene(kind, seed, param, None, None, None, C.byref(sz)), "mcpt_gen_scene")
    hot = np.zeros(sz.n_tris, HOT_DTYPE)
    cold = np.zeros(sz.n_tris, COLD_DTYPE)
    mats = np.zeros(sz.n_materials, MAT_DTYPE)
    _check(_lib.mcpt_gen_scene(kind, seed, param, _ptr(hot), _ptr(cold), _ptr(mats), None), "mcpt_gen_scene")
    return hot, cold, mats


def load_obj(path, translation=(0.0, 0.0, 0.0), scale=1.0, rotation_y=0.0, fallback=None, load_textures=False):
    """mcpt_load_obj: an OBJ (+ MTL) file -> (hot, cold, materials, stats),
    plus the list of map_Kd images (RGBA8 arrays; material texture k names
    the k-th, counting from 1) when load_textures is set. Material 0 is
    `fallback` (a MAT_DTYPE record; default Lambertian grey 0.73 linear),
    1 + i the file's i-th MTL material."""
    if fallback is None:
        fallback = np.zeros(1, MAT_DTYPE)[0]
        fallback["albedo"] = 0.73
    o = ObjOpts()
    o.translation[:] = [float(v) for v in translation]
    o.uniform_scale = scale
    o.rotation_y = rotation_y
    C.memmove(o.fallback, np.asarray(fallback, MAT_DTYPE).tobytes(), MAT_DTYPE.itemsize)
    o.load_textures = 1 if load_textures else 0
    st = ObjStats()
    p = os.fsencode(path)
    _check(_lib.mcpt_load_obj(p, C.byref(o), None, None, None, C.byref(st)), "mcpt_load_obj")
    hot = np.zeros(st.n_tris, HOT_DTYPE)
    cold = np.zeros(st.n_tris, COLD_DTYPE)
    mats = np.zeros(st.n_materials, MAT_DTYPE)
    _check(_lib.mcpt_load_obj(p, C.byref(o), _ptr(hot), _ptr(cold), _ptr(mats), C.byref(st)), "mcpt_load_obj")
    if not load_textures:
        return hot, cold, mats, st.as_dict()
    images = []
    for k in range(st.textures_loaded):
        w, h = C.c_int32(), C.c_int32()
        _check(_lib.mcpt_obj_texture(k, C.byref(w), C.byref(h), None), "mcpt_obj_texture")
        img = np.zeros((h.value, w.value, 4), np.uint8)
        _check(_lib.mcpt_obj_texture(k, C.byref(w), C.byref(h), _ptr(img)), "mcpt_obj_texture")
        images.append(img)
    return hot, cold, mats, st.as_dict(), images


def load_pbrt(path, max_prims=0):
    """mcpt_load_pbrt: a pbrt-v3 scene -> (hot, cold, spheres, materials,
    info, images); material texture k names images[k - 1]."""
    p = os.fsencode(path)
    info = PbrtInfo()
    _check(_lib.mcpt_load_pbrt(p, max_prims, None, None, None, None, C.byref(info)), "mcpt_load_pbrt")
    hot = np.zeros(info.n_tris, HOT_DTYPE)
    cold = np.zeros(info.n_tris, COLD_DTYPE)
    sph = np.zeros(info.n_spheres, SPHERE_DTYPE)
    mats = np.zeros(info.n_materials, MAT_DTYPE)
    _check(_lib.mcpt_load_pbrt(p, max_prims, _ptr(hot), _ptr(cold), _ptr(sph), _ptr(mats), C.byref(info)),
           "mcpt_load_pbrt")
    images = []
    for k in range(info.textures_loaded):
        w, h = C.c_int32(), C.c_int32()
        _check(_lib.mcpt_pbrt_texture(k, C.byref(w), C.byref(h), None), "mcpt_pbrt_texture")
        img = np.zeros((h.value, w.value, 4), np.uint8)
        _check(_lib.mcpt_pbrt_texture(k, C.byref(w), C.byref(h), _ptr(img)), "mcpt_pbrt_texture")
        images.append(img)
    return hot, cold, sph, mats, info.as_dict(), images


def load_ppm(path):
    """mcpt_load_ppm: a binary PPM as an (H, W, 4) RGBA8 array."""
    w, h = C.c_int32(), C.c_int32()
    p = os.fsencode(path)
    _check(_lib.mcpt_load_ppm(p, C.byref(w), C.byref(h), None), "mcpt_load_ppm")
    out = np.zeros((h.value, w.value, 4), np.uint8)
    _check(_lib.mcpt_load_ppm(p, C.byref(w), C.byref(h), _ptr(out)), "mcpt_load_ppm")
    return out


def load_png(path):
    """mcpt_load_png: a PNG file as an (H, W, 4) RGBA8 array."""
    w, h = C.c_int32(), C.c_int32()
    p = os.fsencode(path)
    _check(_lib.mcpt_load_png(p, C.byref(w), C.byref(h), None), "mcpt_load_png")
    out = np.zeros((h.value, w.value, 4), np.uint8)
    _check(_lib.mcpt_load_png(p, C.byref(w), C.byref(h), _ptr(out)), "mcpt_load_png")
    return out


def load_tga(path):
    """mcpt_load_tga: a TGA file as an (H, W, 4) RGBA8 array."""
    w, h = C.c_int32(), C.c_int32()
    p = os.fsencode(path)
    _check(_lib.mcpt_load_tga(p, C.byref(w), C.byref(h), None), "mcpt_load_tga")
    out = np.zeros((h.value, w.value, 4), np.uint8)
    _check(_lib.mcpt_load_tga(p, C.byref(w), C.byref(h), _ptr(out)), "mcpt_load_tga")
    return out


def load_jpeg(path):
    """mcpt_load_jpeg: a baseline JPEG file as an (H, W, 4) RGBA8 array."""
    w, h = C.c_int32(), C.c_int32()
    p = os.fsencode(path)
    _check(_lib.mcpt_load_jpeg(p, C.byref(w), C.byref(h), None), "mcpt_load_jpeg")
    out = np.zeros((h.value, w.value, 4), np.uint8)
    _check(_lib.mcpt_load_jpeg(p, C.byref(w), C.byref(h), _ptr(out)), "mcpt_load_jpeg")
    return out


def _load_float_image(fn, name, path):
    w, h = C.c_int32(), C.c_int32()
    p = os.fsencode(path)
    _check(fn(p, C.byref(w), C.byref(h), None), name)
    out = np.zeros((h.value, w.value, 4), np.float32)
    _check(fn(p, C.byref(w), C.byref(h), _ptr(out)), name)
    return out


def load_pfm(path):
    """mcpt_load_pfm: a little-endian PFM as an (H, W, 4) RGBA FP32 array,
    top row first."""
    return _load_float_image(_lib.mcpt_load_pfm, "mcpt_load_pfm", path)


def load_hdr(path):
    """mcpt_load_hdr: a Radiance RGBE file as an (H, W, 4) RGBA FP32 array."""
    return _load_float_image(_lib.mcpt_load_hdr, "mcpt_load_hdr", path)


def load_exr(path):
    """mcpt_load_exr: a scanline OpenEXR file as an (H, W, 4) RGBA FP32 array."""
    return _load_float_image(_lib.mcpt_load_exr, "mcpt_load_exr", path)


def load_nvdb(path):
    """mcpt_load_nvdb: a NanoVDB float grid as (density (nz, ny, nx) float32,
    origin, voxel_size, max_density), ready for Scene.set_density_grid."""
    import numpy as np
    g = DensityGrid()
    mx = C.c_float(0.0)
    p = os.fspath(path).encode()
    _check(_lib.mcpt_load_nvdb(p, C.byref(g), None, C.byref(mx)), "mcpt_load_nvdb")
    nx, ny, nz = g.dims
    out = np.zeros((nz, ny, nx), np.float32)
    _check(_lib.mcpt_load_nvdb(p, C.byref(g), _ptr(out), C.byref(mx)), "mcpt_load_nvdb")
    return out, tuple(g.origin), tuple(g.voxel_size), float(mx.value)


def load_env(path):
    """An environment map file by its extension, as the reference's --env
    option chooses (src/path_tracer.cu:4197-4199, src/image_io.cpp:41-60):
    .pfm, .exr or .hdr."""
    low = os.fspath(path).lower()
    if low.endswith(".pfm"):
        return load_pfm(path)
    if low.endswith(".exr"):
        return load_exr(path)
    if low.endswith(".hdr"):
        return load_hdr(path)
    raise ValueError(f"{path}: environment maps are read from .pfm, .exr or .hdr files")


def write_png(path, rgb8):
    """mcpt_write_png: an (H, W, 3) uint8 image to a PNG file."""
    rgb8 = np.ascontiguousarray(rgb8, np.uint8)
    h, w = rgb8.shape[:2]
    _check(_lib.mcpt_write_png(os.fsencode(path), w, h, _ptr(rgb8)), "mcpt_write_png")


def gen_sky(width=512, height=256):
    rgba = np.zeros((height, width, 4), np.float32)
    _check(_lib.mcpt_gen_sky(width, height, _ptr(rgba)), "mcpt_gen_sky")
    return rgba


def gen_camera(kind, width, height):
    d = CameraDesc()
    _check(_lib.mcpt_gen_camera(kind, width, height, C.byref(d)), "mcpt_gen_camera")
    return d


def camera_init(desc):
    cam = Camera()
    _check(_lib.mcpt_camera_init(C.byref(desc), C.byref(cam)), "mcpt_camera_init")
    return cam


def make_env(rgba, intensity=1.0, rotation_radians=0.0):
    rgba = np.ascontiguousarray(rgba, np.float32)
    e = EnvDesc(rgba.shape[1], rgba.shape[0], intensity, rotation_radians, rgba.ctypes.data)
    e._keep = rgba
    return e


# ------------------------------------------------------------ device objects
class Context:
    """A device context. Closing it first closes every scene and film made
    from it (the C library requires children to go before their context)."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        self.h = VP()
        self._children = weakref.WeakSet()
        _check(_lib.mcpt_ctx_create(device, hbm_budget_bytes, C.byref(self.h)), "mcpt_ctx_create")

    def close(self):
        if self.h:
            # servers first: they hold the scenes they render
            for child in sorted(self._children, key=lambda c: not isinstance(c, Server)):
                child.close()
            _check(_lib.mcpt_ctx_destroy(self.h), "mcpt_ctx_destroy")
            self.h = VP()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def memory(self):
        used, budget = C.c_uint64(), C.c_uint64()
        _check(_lib.mcpt_ctx_memory(self.h, C.byref(used), C.byref(budget)), "mcpt_ctx_memory")
        return used.value, budget.value

    def describe(self):
        buf = C.create_string_buffer(1024)
        _check(_lib.mcpt_ctx_describe(self.h, buf, 1024), "mcpt_ctx_describe")
        return buf.value.decode()

    def scene(self, hot, cold, mats, env=None, spheres=None):
        return Scene(self, hot, cold, mats, env, spheres)

    def film(self, width, height):
        return Film(self, width, height)

    def server(self, scenes, **opts):
        return Server(self, scenes, **opts)

    def mlp(self, seed=1):
        return Mlp(self, seed)


class Scene:
    def __init__(self, ctx, hot, cold, mats, env=None, spheres=None):
        self.ctx = ctx
        self.hot = np.ascontiguousarray(hot, HOT_DTYPE)
        self.cold = np.ascontiguousarray(cold, COLD_DTYPE)
        self.mats = np.ascontiguousarray(mats, MAT_DTYPE)
        self.spheres = np.ascontiguousarray(spheres if spheres is not None else np.zeros(0, SPHERE_DTYPE),
                                            SPHERE_DTYPE)
        self.n = len(self.hot) + len(self.spheres)  # BVH primitives: spheres, then triangles
        self.h = VP()
        ctx._children.add(self)
        envp = C.byref(env) if env is not None else None
        if len(self.spheres):
            _check(_lib.mcpt_scene_create_ex(ctx.h, _ptr(self.hot) if len(self.hot) else None,
                                             _ptr(self.cold) if len(self.cold) else None, len(self.hot),
                                             _ptr(self.spheres), len(self.spheres), _ptr(self.mats), len(self.mats),
                                             envp, C.byref(self.h)), "mcpt_scene_create_ex")
        else:
            _check(_lib.mcpt_scene_create(ctx.h, _ptr(self.hot), _ptr(self.cold), len(self.hot), _ptr(self.mats),
                                          len(self.mats), envp, C.byref(self.h)), "mcpt_scene_create")

    def close(self):
        if self.h:
            if self.ctx.h:
                _check(_lib.mcpt_scene_destroy(self.h), "mcpt_scene_destroy")
            self.h = VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build(self, dump=False):
        """Run the GPU LBVH build; with dump=True return every stage as arrays."""
        if not dump:
            d = LbvhDump()
            _check(_lib.mcpt_lbvh_build(self.h, C.byref(d)), "mcpt_lbvh_build")
            return {"max_depth": d.max_depth, "build_ms": d.build_ms}
        n, ni = self.n, max(self.n - 1, 0)
        out = {
            "morton": np.zeros(n, np.uint32), "sorted_keys": np.zeros(n, np.uint32),
            "sorted_ids": np.zeros(n, np.int32), "left": np.zeros(ni, np.int32), "right": np.zeros(ni, np.int32),
            "left_is_leaf": np.zeros(ni, np.uint8), "right_is_leaf": np.zeros(ni, np.uint8),
            "parent": np.zeros(ni, np.int32), "leaf_parent": np.zeros(n, np.int32),
            "internal_bounds": np.zeros(ni, AABB_DTYPE), "leaf_bounds": np.zeros(n, AABB_DTYPE),
            "scene_bounds": np.zeros(1, AABB_DTYPE),
        }
        d = LbvhDump(*[_ptr(out[k]) if out[k].size else None for k in
                       ("morton", "sorted_keys", "sorted_ids", "left", "right", "left_is_leaf", "right_is_leaf",
                        "parent", "leaf_parent", "internal_bounds", "leaf_bounds", "scene_bounds")])
        _check(_lib.mcpt_lbvh_build(self.h, C.byref(d)), "mcpt_lbvh_build")
        out["max_depth"] = d.max_depth
        out["build_ms"] = d.build_ms
        return out

    def dump_wide(self):
        """The four-wide tree (WIDE_DTYPE records, root first) and each leaf's
        rank in the reference's leaf-test order."""
        n = C.c_int32()
        _check(_lib.mcpt_lbvh_dump_wide(self.h, None, 0, C.byref(n), None), "mcpt_lbvh_dump_wide")
        nodes = np.zeros(n.value, WIDE_DTYPE)
        tie = np.zeros(self.n, np.int32)
        _check(_lib.mcpt_lbvh_dump_wide(self.h, _ptr(nodes), n.value, C.byref(n), _ptr(tie)), "mcpt_lbvh_dump_wide")
        return nodes, tie

    def add_texture(self, kind, color_a=(0, 0, 0), color_b=(1, 1, 1), scale=1.0, image=None):
        """mcpt_scene_add_texture; returns the index k (materials use texture = k + 1)."""
        d, _keep = _texture_desc(kind, color_a, color_b, scale, image)
        k = C.c_int32()
        _check(_lib.mcpt_scene_add_texture(self.h, C.byref(d), C.byref(k)), "mcpt_scene_add_texture")
        return k.value

    def set_density_grid(self, density, origin, voxel_size):
        """mcpt_scene_set_density_grid: density is a (nz, ny, nx) float32
        array whose voxel [k, j, i] sits at origin + (i, j, k) * voxel_size;
        every isotropic material becomes a heterogeneous medium. Returns
        the grid's largest value."""
        g, _keep = _density_grid_desc(density, origin, voxel_size)
        mx = C.c_float()
        _check(_lib.mcpt_scene_set_density_grid(self.h, C.byref(g), C.byref(mx)), "mcpt_scene_set_density_grid")
        return mx.value

    def lights(self):
        n = C.c_int32()
        _check(_lib.mcpt_scene_light_count(self.h, C.byref(n)), "mcpt_scene_light_count")
        out = np.zeros(n.value, LIGHT_DTYPE)
        if n.value:
            _check(_lib.mcpt_scene_lights(self.h, _ptr(out)), "mcpt_scene_lights")
        return out

    def env_tables(self, width, height):
        row = np.zeros(height + 1, np.float32)
        col = np.zeros(height * (width + 1), np.float32)
        tot = C.c_float()
        _check(_lib.mcpt_scene_env_tables(self.h, _ptr(row), _ptr(col), C.byref(tot)), "mcpt_scene_env_tables")
        return row, col, tot.value

    def trace_closest(self, rays, counts=False):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        steps = np.zeros(len(rays), np.uint32) if counts else None
        leaves = np.zeros(len(rays), np.uint32) if counts else None
        _check(_lib.mcpt_trace_closest(self.h, _ptr(rays), len(rays), _ptr(hits), _ptr(steps), _ptr(leaves)),
               "mcpt_trace_closest")
        return (hits, steps, leaves) if counts else hits

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint8)
        _check(_lib.mcpt_trace_any(self.h, _ptr(rays), len(rays), _ptr(occ)), "mcpt_trace_any")
        return occ

    def render(self, camera, film, mode=MODE_PERSISTENT, spp=1, sample_begin=0, max_depth=8, seed=0,
               refill_lanes=0, flags=0, batch=0, group_below=0):
        o = RenderOpts(mode, spp, sample_begin, max_depth, seed, refill_lanes, flags, batch, group_below)
        st = RenderStats()
        _check(_lib.mcpt_render(self.h, C.byref(camera), C.byref(o), film.h, C.byref(st)), "mcpt_render")
        return st.as_dict()

    def render_cached(self, camera, film, mlp, spp=1, sample_begin=0, max_depth=8, seed=0, refill_lanes=0,
                      train_paths=0, train_batch=0, train_steps=0, terminate_depth=0, lr=0.0, ema_decay=0.0,
                      frame_spp=0):
        o = RenderOpts(MODE_PERSISTENT, spp, sample_begin, max_depth, seed, refill_lanes, 0, 0, 0)
        co = CacheOpts(train_paths, train_batch, train_steps, terminate_depth, lr, ema_decay, frame_spp)
        st = CacheStats()
        _check(_lib.mcpt_render_cached(self.h, C.byref(camera), C.byref(o), film.h, mlp.h, C.byref(co),
                                       C.byref(st)), "mcpt_render_cached")
        return st.as_dict()


class Film:
    def __init__(self, ctx, width, height):
        self.ctx, self.w, self.hgt = ctx, width, height
        self.h = VP()
        ctx._children.add(self)
        _check(_lib.mcpt_film_create(ctx.h, width, height, C.byref(self.h)), "mcpt_film_create")

    def close(self):
        if self.h:
            if self.ctx.h:
                _lib.mcpt_film_destroy(self.h)
            self.h = VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self):
        _check(_lib.mcpt_film_clear(self.h), "mcpt_film_clear")

    def read(self):
        out = np.zeros((self.hgt, self.w, 3), np.float32)
        _check(_lib.mcpt_film_read(self.h, _ptr(out)), "mcpt_film_read")
        return out

    def write(self, rgb):
        rgb = np.ascontiguousarray(rgb, np.float32).reshape(self.hgt, self.w, 3)
        _check(_lib.mcpt_film_write(self.h, _ptr(rgb)), "mcpt_film_write")

    def resolve(self, spp, tonemap=TONEMAP_NONE, exposure=1.0):
        out = np.zeros((self.hgt, self.w, 3), np.float32)
        _check(_lib.mcpt_resolve(self.h, spp, tonemap, exposure, _ptr(out)), "mcpt_resolve")
        return out

    def resolve_rgb8(self, spp, tonemap=TONEMAP_NONE, exposure=1.0):
        """mcpt_resolve_rgb8: the resolved image as 8-bit RGB (H, W, 3)."""
        out = np.zeros((self.hgt, self.w, 3), np.uint8)
        _check(_lib.mcpt_resolve_rgb8(self.h, spp, tonemap, exposure, _ptr(out)), "mcpt_resolve_rgb8")
        return out

    def auto_exposure(self, spp, target_gray=0.18):
        """mcpt_auto_exposure: target_gray / log-average luminance."""
        e = C.c_float()
        _check(_lib.mcpt_auto_exposure(self.h, spp, target_gray, C.byref(e)), "mcpt_auto_exposure")
        return e.value


class Server:
    """Continuous-batching render server over resident scenes
    (mcpt_server_*). Jobs render into their own films; a finished job's
    film stays on the device until release()."""

    def __init__(self, ctx, scenes, path_capacity=0, max_jobs=0, frames_per_job=0, refill_lanes=0,
                 film_budget_bytes=0, policy=0, group_below=0):
        self.ctx = ctx
        self.scenes = list(scenes)
        self.h = VP()
        ctx._children.add(self)
        arr = (VP * len(self.scenes))(*[sc.h for sc in self.scenes])
        o = ServerOpts(path_capacity, max_jobs, frames_per_job, refill_lanes, film_budget_bytes, policy,
                       group_below)
        _check(_lib.mcpt_server_create(ctx.h, arr, len(self.scenes), C.byref(o), C.byref(self.h)),
               "mcpt_server_create")

    def close(self):
        if self.h:
            if self.ctx.h:
                _lib.mcpt_server_destroy(self.h)
            self.h = VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, scene, camera, spp, max_depth=8, seed=0):
        """Returns (job_id, admitted_now)."""
        j = Job(scene, spp, max_depth, 0, seed, camera)
        jid = C.c_uint64()
        code = _check(_lib.mcpt_server_submit(self.h, C.byref(j), C.byref(jid)), "mcpt_server_submit")
        return jid.value, code == OK

    def poll(self, job_id):
        r = JobResult()
        _check(_lib.mcpt_server_poll(self.h, job_id, C.byref(r)), "mcpt_server_poll")
        return r.as_dict()

    def wait(self, job_id, read=True):
        """Block until the job is done; returns (result, film rgb or None)."""
        r = JobResult()
        rgb = None
        if read:
            info = self.poll(job_id)
            rgb = np.zeros((info["height"], info["width"], 3), np.float32)
        _check(_lib.mcpt_server_wait(self.h, job_id, C.byref(r), _ptr(rgb)), "mcpt_server_wait")
        return r.as_dict(), rgb

    def release(self, job_id):
        _check(_lib.mcpt_server_release(self.h, job_id), "mcpt_server_release")

    def stats(self):
        st = ServerStats()
        _check(_lib.mcpt_server_stats_get(self.h, C.byref(st)), "mcpt_server_stats_get")
        return st.as_dict()


class Mlp:
    """The radiance-cache network (mcpt_mlp_*)."""

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        self.h = VP()
        ctx._children.add(self)
        _check(_lib.mcpt_mlp_create(ctx.h, seed, C.byref(self.h)), "mcpt_mlp_create")

    def close(self):
        if self.h:
            if self.ctx.h:
                _lib.mcpt_mlp_destroy(self.h)
            self.h = VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(MLP_PARAMS, np.float32)
        _check(_lib.mcpt_mlp_get_params(self.h, _ptr(out)), "mcpt_mlp_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == MLP_PARAMS
        _check(_lib.mcpt_mlp_set_params(self.h, _ptr(p)), "mcpt_mlp_set_params")

    def forward(self, features):
        f = np.ascontiguousarray(features, np.float32).reshape(-1, MLP_FEATURES)
        out = np.zeros((len(f), 3), np.float32)
        _check(_lib.mcpt_mlp_forward(self.h, _ptr(f), len(f), _ptr(out)), "mcpt_mlp_forward")
        return out

    def gradient(self, features, targets):
        f = np.ascontiguousarray(features, np.float32).reshape(-1, MLP_FEATURES)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        _check(_lib.mcpt_mlp_gradient(self.h, _ptr(f), _ptr(t), len(f), C.byref(loss), _ptr(g)), "mcpt_mlp_gradient")
        return loss.value, g

    def train(self, features, targets, batch=16384, lr=0.005):
        f = np.ascontiguousarray(features, np.float32).reshape(-1, MLP_FEATURES)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        loss = C.c_float()
        _check(_lib.mcpt_mlp_train(self.h, _ptr(f), _ptr(t), len(f), batch, lr, C.byref(loss)), "mcpt_mlp_train")
        return loss.value

