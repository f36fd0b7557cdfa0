"""ctypes binding of liboracle.so, the CPU oracle (test infrastructure only).

Only the test suite, the smoke check and the benchmark's CPU-baseline leg
import this module; the product library never does. Record layouts are the
ones of include/mcpt.h.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

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


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        VP = C.c_void_p
        sigs = {
            "oracle_camera_init": ([VP, VP], C.c_int),
            "oracle_env_tables": ([VP, C.c_int, C.c_int, VP, VP, C.POINTER(C.c_float)], None),
            "oracle_finalize": ([VP, C.c_int, VP, VP, VP, VP], None),
            "oracle_expand_bits": ([C.c_uint32], C.c_uint32),
            "oracle_quantize": ([C.c_float, C.c_float, C.c_float], C.c_uint32),
            "oracle_sort_stable": ([VP, C.c_int, VP, VP], None),
            "oracle_sort_reference": ([VP, C.c_int, VP, VP], None),
            "oracle_karras": ([VP, C.c_int, VP, VP, VP, VP, VP, VP], None),
            "oracle_scene_create": ([VP, VP, C.c_int, VP, C.c_int, VP, C.POINTER(VP)], C.c_int),
            "oracle_scene_create_ex": ([VP, VP, C.c_int, VP, C.c_int, VP, C.c_int, VP, C.POINTER(VP)], C.c_int),
            "oracle_scene_destroy": ([VP], None),
            "oracle_sphere_uv": ([VP, C.c_int, VP], None),
            "oracle_scene_add_texture": ([VP, VP], C.c_int),
            "oracle_scene_set_density_grid": ([VP, VP, C.POINTER(C.c_float)], C.c_int),
            "oracle_grid_density": ([VP, VP, C.c_int, VP], None),
            "oracle_grid_track": ([VP, C.c_int, C.c_float, VP, C.c_int, C.c_uint64, VP], None),
            "oracle_sample_albedo": ([VP, VP, VP, C.c_int, VP], None),
            "oracle_scene_dump": ([VP, VP], C.c_int),
            "oracle_scene_lights": ([VP, VP, C.c_int], C.c_int),
            "oracle_trace_closest": ([VP, VP, C.c_int, VP, VP, VP, C.c_int], None),
            "oracle_trace_any": ([VP, VP, C.c_int, VP, C.c_int], None),
            "oracle_brute_closest": ([VP, C.c_int, VP, C.c_int, VP], None),
            "oracle_render": ([VP, VP, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int, VP, VP, C.c_int,
                               VP], C.c_int),
            "oracle_resolve": ([VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, VP], None),
            "oracle_resolve_rgb8": ([VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, VP], None),
            "oracle_auto_exposure": ([VP, C.c_int, C.c_int, C.c_int, C.c_float], C.c_float),
            "oracle_fmath": ([C.c_int, VP, VP, VP, C.c_int], None),
            "oracle_pcg32": ([C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, VP], None),
            "oracle_sampling": ([C.c_int, VP, C.c_int, VP], C.c_int),
            "oracle_env_probe": ([VP, C.c_int, C.c_int, C.c_float, C.c_int, VP, C.c_int, VP], C.c_int),
            "oracle_conductor": ([VP, VP, C.c_float, C.c_int, C.c_uint64, VP], None),
            "oracle_camera_rays": ([VP, VP, C.c_int, C.c_uint32, C.c_uint64, VP], None),
            "oracle_tri_interp": ([VP, VP, VP, C.c_int, VP, VP], None),
        }
        for name, (args, res) in sigs.items():
            f = getattr(_lib, name)
            f.argtypes = args
            f.restype = res
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class DensityGrid(C.Structure):
    """mcpt_density_grid: dense voxels, x fastest"""
    _fields_ = [("dims", C.c_int32 * 3), ("origin", C.c_float * 3), ("voxel_size", C.c_float * 3),
                ("reserved", C.c_int32), ("density", C.c_void_p)]


def density_grid_desc(density, origin, voxel_size):
    """(desc, keep): density is a (nz, ny, nx) float32 array"""
    keep = np.ascontiguousarray(density, np.float32)
    if keep.ndim != 3:
        raise ValueError("density must be (nz, ny, nx)")
    g = DensityGrid()
    g.dims[:] = [keep.shape[2], keep.shape[1], keep.shape[0]]
    g.origin[:] = [float(x) for x in origin]
    g.voxel_size[:] = [float(x) for x in (voxel_size if np.ndim(voxel_size) else (voxel_size,) * 3)]
    g.density = keep.ctypes.data
    return g, keep


class EnvDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("intensity", C.c_float),
                ("rotation_radians", C.c_float), ("rgba", C.c_void_p)]


class Camera(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("pixel00", C.c_float * 3), ("delta_u", C.c_float * 3),
                ("delta_v", C.c_float * 3), ("u", C.c_float * 3), ("v", C.c_float * 3), ("w", C.c_float * 3),
                ("lens_radius", C.c_float), ("width", C.c_int32), ("height", C.c_int32),
                ("background", C.c_float * 3), ("rr_cap", C.c_float)]


class RenderStats(C.Structure):
    _fields_ = [("extend_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("paths", C.c_uint64),
                ("extend_steps", C.c_uint64), ("extend_leaves", C.c_uint64), ("seconds", C.c_double),
                ("threads", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def camera_init(desc):
    """desc: a ctypes structure laid out as mcpt_camera_desc."""
    cam = Camera()
    if lib().oracle_camera_init(C.addressof(desc), C.addressof(cam)) != 0:
        raise ValueError("camera size not reproducible")
    return cam


def env_tables(rgba):
    rgba = np.ascontiguousarray(rgba, np.float32)
    h, w = rgba.shape[:2]
    row = np.zeros(h + 1, np.float32)
    col = np.zeros(h * (w + 1), np.float32)
    tot = C.c_float()
    lib().oracle_env_tables(_p(rgba), w, h, _p(row), _p(col), C.byref(tot))
    return row, col, tot.value


def finalize(hot):
    hot = np.ascontiguousarray(hot, HOT_DTYPE)
    n = len(hot)
    boxes = np.zeros(n, AABB_DTYPE)
    sb = np.zeros(1, AABB_DTYPE)
    cen = np.zeros((n, 3), np.float32)
    mort = np.zeros(n, np.uint32)
    lib().oracle_finalize(_p(hot), n, _p(boxes), _p(sb), _p(cen), _p(mort))
    return boxes, sb[0], cen, mort


def sort_stable(morton):
    morton = np.ascontiguousarray(morton, np.uint32)
    k = np.zeros_like(morton)
    i = np.zeros(len(morton), np.int32)
    lib().oracle_sort_stable(_p(morton), len(morton), _p(k), _p(i))
    return k, i


def sort_reference(morton):
    morton = np.ascontiguousarray(morton, np.uint32)
    k = np.zeros_like(morton)
    i = np.zeros(len(morton), np.int32)
    lib().oracle_sort_reference(_p(morton), len(morton), _p(k), _p(i))
    return k, i


def karras(keys):
    keys = np.ascontiguousarray(keys, np.uint32)
    n = len(keys)
    ni = max(n - 1, 0)
    L, R, P = (np.zeros(ni, np.int32) for _ in range(3))
    LL, RL = np.zeros(ni, np.uint8), np.zeros(ni, np.uint8)
    LP = np.full(n, -1, np.int32)
    lib().oracle_karras(_p(keys), n, _p(L), _p(R), _p(LL), _p(RL), _p(P), _p(LP))
    return {"left": L, "right": R, "left_is_leaf": LL, "right_is_leaf": RL, "parent": P, "leaf_parent": LP}


def sphere_uv(normals):
    """computeSphereUvFromOutwardNormal for (n, 3) unit normals -> (n, 2)."""
    nrm = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    out = np.zeros((len(nrm), 2), np.float32)
    lib().oracle_sphere_uv(_p(nrm), len(nrm), _p(out))
    return out


class Scene:
    def __init__(self, hot, cold, mats, sky=None, intensity=1.0, rotation=0.0, spheres=None):
        self.hot = np.ascontiguousarray(hot, HOT_DTYPE)
        self.cold = np.ascontiguousarray(cold, COLD_DTYPE)
        self.mats = np.ascontiguousarray(mats, MAT_DTYPE)
        self.spheres = np.ascontiguousarray(spheres if spheres is not None else np.zeros(0, SPHERE_DTYPE),
                                            SPHERE_DTYPE)
        self.n = len(self.hot) + len(self.spheres)  # BVH primitives
        env = None
        if sky is not None:
            self._sky = np.ascontiguousarray(sky, np.float32)
            env = EnvDesc(self._sky.shape[1], self._sky.shape[0], intensity, rotation, self._sky.ctypes.data)
        self.h = C.c_void_p()
        rc = lib().oracle_scene_create_ex(_p(self.hot) if len(self.hot) else None,
                                          _p(self.cold) if len(self.cold) else None, len(self.hot),
                                          _p(self.spheres) if len(self.spheres) else None, len(self.spheres),
                                          _p(self.mats), len(self.mats),
                                          C.byref(env) if env is not None else None, C.byref(self.h))
        if rc != 0:
            raise ValueError("oracle_scene_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_scene_destroy(self.h)
            self.h = None

    def add_texture(self, kind, color_a=(0, 0, 0), color_b=(1, 1, 1), scale=1.0, image=None):
        """the oracle's copy of a scene texture; kind 1 checker, 2 image
        ((H, W, 4) RGBA8); returns the index (materials use index + 1)"""
        class Desc(C.Structure):
            _fields_ = [("type", C.c_int32), ("color_a", C.c_float * 3), ("color_b", C.c_float * 3),
                        ("scale", C.c_float), ("width", C.c_int32), ("height", C.c_int32), ("rgba8", C.c_void_p)]
        d = Desc()
        d.type = kind
        d.color_a[:] = [float(x) for x in color_a]
        d.color_b[:] = [float(x) for x in color_b]
        d.scale = scale
        keep = None
        if image is not None:
            keep = np.ascontiguousarray(image, np.uint8)
            d.height, d.width = keep.shape[:2]
            d.rgba8 = keep.ctypes.data
        k = lib().oracle_scene_add_texture(self.h, C.byref(d))
        if k < 0:
            raise ValueError("bad texture")
        return k

    def set_density_grid(self, density, origin, voxel_size):
        """the oracle's copy of a density grid ((nz, ny, nx) float32);
        returns the largest voxel value"""
        g, keep = density_grid_desc(density, origin, voxel_size)
        mx = C.c_float(0.0)
        if lib().oracle_scene_set_density_grid(self.h, C.byref(g), C.byref(mx)) < 0:
            raise ValueError("bad density grid")
        self._grid = keep
        return mx.value

    def grid_density(self, pts):
        pts = np.ascontiguousarray(np.asarray(pts, np.float32).reshape(-1, 3))
        out = np.zeros(len(pts), np.float32)
        lib().oracle_grid_density(self.h, _p(pts), len(pts), _p(out))
        return out

    def grid_track(self, fn, sigma, rays, seed=0):
        """fn "free_flight" (delta tracking) or "transmittance" (ratio
        tracking) along rays of 7 floats: origin, direction, length"""
        rays = np.ascontiguousarray(np.asarray(rays, np.float32).reshape(-1, 7))
        out = np.zeros(len(rays), np.float32)
        lib().oracle_grid_track(self.h, {"free_flight": 0, "transmittance": 1}[fn], sigma, _p(rays), len(rays),
                                seed, _p(out))
        return out

    def sample_albedo(self, mat, uv):
        """albedo of material record `mat` at (u, v) rows"""
        m = np.ascontiguousarray(np.asarray(mat, MAT_DTYPE).reshape(1))
        uv = np.ascontiguousarray(np.asarray(uv, np.float32).reshape(-1, 2))
        out = np.zeros((len(uv), 3), np.float32)
        lib().oracle_sample_albedo(self.h, _p(m), _p(uv), len(uv), _p(out))
        return out

    def dump(self):
        n, ni = self.n, max(self.n - 1, 0)
        out = {
            "morton": np.zeros(n, np.uint32), "sorted_keys": np.zeros(n, np.uint32),
            "sorted_ids": np.zeros(n, np.int32), "left": np.zeros(ni, np.int32), "right": np.zeros(ni, np.int32),
            "left_is_leaf": np.zeros(ni, np.uint8), "right_is_leaf": np.zeros(ni, np.uint8),
            "parent": np.zeros(ni, np.int32), "leaf_parent": np.zeros(n, np.int32),
            "internal_bounds": np.zeros(ni, AABB_DTYPE), "leaf_bounds": np.zeros(n, AABB_DTYPE),
            "scene_bounds": np.zeros(1, AABB_DTYPE),
        }

        class Dump(C.Structure):
            _fields_ = [(k, C.c_void_p) for k in ("morton", "sorted_keys", "sorted_ids", "left", "right",
                                                   "left_is_leaf", "right_is_leaf", "parent", "leaf_parent",
                                                   "internal_bounds", "leaf_bounds", "scene_bounds")] + \
                       [("max_depth", C.c_int32), ("build_ms", C.c_float)]

        d = Dump(*[_p(out[k]) if out[k].size else None for k, _ in Dump._fields_[:12]])
        lib().oracle_scene_dump(self.h, C.byref(d))
        out["max_depth"] = d.max_depth
        return out

    def lights(self):
        n = lib().oracle_scene_lights(self.h, None, 0)
        out = np.zeros(n, LIGHT_DTYPE)
        lib().oracle_scene_lights(self.h, _p(out), n)
        return out

    def trace_closest(self, rays, threads=0):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        steps = np.zeros(len(rays), np.uint32)
        leaves = np.zeros(len(rays), np.uint32)
        lib().oracle_trace_closest(self.h, _p(rays), len(rays), _p(hits), _p(steps), _p(leaves), threads)
        return hits, steps, leaves

    def trace_any(self, rays, threads=0):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint8)
        lib().oracle_trace_any(self.h, _p(rays), len(rays), _p(occ), threads)
        return occ

    def render(self, cam, spp, max_depth, seed=0, sample_begin=0, rows=None, accum=None, accum64=False,
               threads=0):
        W, H = cam.width, cam.height
        if accum is None:
            accum = np.zeros((H, W, 3), np.float32)
        a64 = np.zeros((H, W, 3), np.float64) if accum64 else None
        r0, r1 = rows if rows else (0, H)
        st = RenderStats()
        rc = lib().oracle_render(self.h, C.addressof(cam), spp, sample_begin, max_depth, seed, r0, r1, _p(accum),
                                 _p(a64), threads, C.byref(st))
        if rc != 0:
            raise ValueError("oracle_render failed")
        return accum, a64, st.as_dict()


FMATH = {"sin": 0, "cos": 1, "asin": 2, "acos": 3, "atan2": 4, "log": 5, "exp": 6, "pow": 7}


def fmath(name, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y if y is not None else np.zeros_like(x), np.float32)
    out = np.zeros_like(x)
    lib().oracle_fmath(FMATH[name], _p(x), _p(y), _p(out), len(x))
    return out


SAMPLING = {"refract": (0, 7, 3), "schlick": (1, 2, 1), "dielectric": (2, 8, 3), "hg": (3, 6, 3), "hg_pdf": (4, 2, 1)}


def sampling(name, args):
    """Evaluate one of the sampling functions on rows of arguments (float32):
    refract(d3, n3, eta), schlick(cos, eta), dielectric(d3, n3, eta, u),
    hg(in3, g, u1, u2), hg_pdf(cos, g)."""
    fn, nin, nout = SAMPLING[name]
    a = np.ascontiguousarray(np.asarray(args, np.float32).reshape(-1, nin))
    out = np.zeros((len(a), nout), np.float32)
    if lib().oracle_sampling(fn, _p(a), len(a), _p(out)) != 0:
        raise ValueError(name)
    return out if nout > 1 else out[:, 0]


ENV = {"sample": (0, 4, 4), "pdf": (1, 3, 1), "uv": (2, 3, 2), "direction": (3, 2, 3)}


def env_probe(rgba, rot, name, args):
    """Environment sampling on an (h, w, 4) map: sample(u4) -> (dir3, pdf),
    pdf(dir3), uv(dir3) -> (u, v), direction(u, v) -> dir3."""
    rgba = np.ascontiguousarray(rgba, np.float32)
    h, w = rgba.shape[:2]
    fn, nin, nout = ENV[name]
    a = np.ascontiguousarray(np.asarray(args, np.float32).reshape(-1, nin))
    out = np.zeros((len(a), nout), np.float32)
    if lib().oracle_env_probe(_p(rgba), w, h, rot, fn, _p(a), len(a), _p(out)) != 0:
        raise ValueError(name)
    return out if nout > 1 else out[:, 0]


def conductor(d, n, fuzz, count, seed=0):
    """count draws of the conductor scatter for incoming d and normal n"""
    d = np.ascontiguousarray(d, np.float32)
    n = np.ascontiguousarray(n, np.float32)
    out = np.zeros((count, 3), np.float32)
    lib().oracle_conductor(_p(d), _p(n), fuzz, count, seed, _p(out))
    return out


def camera_rays(cam, pixels, sample=0, seed=0):
    """(origins, directions, times) of the camera rays of (x, y) pixels"""
    px = np.ascontiguousarray(pixels, np.int32).reshape(-1, 2)
    out = np.zeros((len(px), 7), np.float32)
    lib().oracle_camera_rays(C.addressof(cam), _p(px), len(px), sample, seed, _p(out))
    return out[:, :3], out[:, 3:6], out[:, 6]


def tri_interp(hot, cold, points):
    """(uv, shading normal) of one triangle (records of HOT_DTYPE and
    COLD_DTYPE) at the given points"""
    hot = np.ascontiguousarray(hot, HOT_DTYPE).reshape(1)
    cold = np.ascontiguousarray(cold, COLD_DTYPE).reshape(1)
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    uv = np.zeros((len(pts), 2), np.float32)
    nrm = np.zeros((len(pts), 3), np.float32)
    lib().oracle_tri_interp(_p(hot), _p(cold), _p(pts), len(pts), _p(uv), _p(nrm))
    return uv, nrm


def pcg32(pixel, sample, seed, n):
    out = np.zeros(n, np.uint32)
    lib().oracle_pcg32(pixel, sample, seed, n, _p(out))
    return out


def brute_closest(hot, rays):
    hot = np.ascontiguousarray(hot, HOT_DTYPE)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(len(rays), HIT_DTYPE)
    lib().oracle_brute_closest(_p(hot), len(hot), _p(rays), len(rays), _p(hits))
    return hits


def resolve_rgb8(accum, spp, tonemap=0, exposure=1.0):
    accum = np.ascontiguousarray(accum, np.float32)
    h, w = accum.shape[:2]
    out = np.zeros((h, w, 3), np.uint8)
    lib().oracle_resolve_rgb8(_p(accum), w, h, spp, tonemap, exposure, _p(out))
    return out


def auto_exposure(accum, spp, target=0.18):
    accum = np.ascontiguousarray(accum, np.float32)
    h, w = accum.shape[:2]
    return lib().oracle_auto_exposure(_p(accum), w, h, spp, target)


def resolve(accum, spp, tonemap=0, exposure=1.0):
    accum = np.ascontiguousarray(accum, np.float32)
    H, W = accum.shape[:2]
    out = np.zeros_like(accum)
    lib().oracle_resolve(_p(accum), W, H, spp, tonemap, exposure, _p(out))
    return out
