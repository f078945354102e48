"""Python host side of the MI355X spectral path tracer.

A thin ctypes layer over libslr_hip.so (include/slr_hip.h), shaped like the
reference's host sequence (HostProgram/main.cpp:22-70): build a Scene from
meshes, spectra, materials and a PerspectiveCamera, build it (SBVH -> QBVH),
then PTRenderer(spp).render(scene, settings) into a film. Ray batches go
through Scene.intersect / Scene.occluded, the QBVH queries.

The library is required: importing works without it, but every call that
needs it raises RuntimeError when libslr_hip.so has not been built
(python -m slr_amd.build). There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libslr_hip.so")

SLRH_OK = 0
SPECTRUM_REFLECTANCE = 0
SPECTRUM_LIGHT_SOURCE = 1
MEM_HOST = 0
MEM_DEVICE = 1
FILM_STRATA16_F64 = 0
FILM_XYZ_F64 = 1
FILM_SRGB8 = 2

EXPORTS = [
    "slrh_create", "slrh_destroy", "slrh_last_error", "slrh_version",
    "slrh_spectrum_rgb", "slrh_spectrum_regular", "slrh_spectrum_evaluate", "slrh_cmf_strata",
    "slrh_material_lambert", "slrh_material_oren_nayar", "slrh_material_ggx_conductor",
    "slrh_material_ggx_dielectric", "slrh_material_diffuse_emitter",
    "slrh_mesh_add", "slrh_prototype_add", "slrh_instance_add", "slrh_instance_get", "slrh_scene_compat",
    "slrh_medium_density_grid", "slrh_medium_query", "slrh_camera_perspective",
    "slrh_scene_create", "slrh_scene_build", "slrh_scene_destroy",
    "slrh_scene_accel_stats", "slrh_scene_build_stats", "slrh_scene_sbvh_dump", "slrh_scene_qbvh_dump",
    "slrh_scene_accelerators", "slrh_scene_query_accel", "slrh_scene_standard_bvh_dump",
    "slrh_scene_load_text", "slrh_scene_geometry", "slrh_scene_camera_get", "slrh_material_info",
    "slrh_intersect", "slrh_occluded", "slrh_intersect_noalloc", "slrh_scene_query_counts",
    "slrh_film_create", "slrh_film_destroy", "slrh_film_clear", "slrh_film_read", "slrh_film_load", "slrh_film_bmp",
    "slrh_render_progressive",
    "slrh_environment", "slrh_render", "slrh_render_batch", "slrh_trace_variants",
    "slrh_server_create", "slrh_server_destroy", "slrh_submit", "slrh_poll", "slrh_wait", "slrh_cancel",
    "slrh_server_stats",
]

_lib = None


class SlrhError(RuntimeError):
    pass


class _Rays(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]


class _Hits(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("t", "prim", "b0", "b1", "inst")]


class _Frame(ctypes.Structure):
    _fields_ = [("l2w", ctypes.c_float * 16)] + [(n, ctypes.c_float) for n in
                ("sensitivity", "aspect", "fov_y", "lens_radius", "img_dist", "obj_dist")] + \
               [(n, ctypes.c_uint32) for n in ("width", "height", "spp", "seed")] + [("film", ctypes.c_void_p)]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("path_state_bytes", ctypes.c_uint64), ("tune_trace", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class _Request(ctypes.Structure):
    _fields_ = [("frame", _Frame), ("max_depth", ctypes.c_uint32), ("slots_per_pixel", ctypes.c_uint32)]


COMPAT_INSTANCE_SHADOW = 1
COMPAT_VOLPT_TANGENT_NEE = 2
FLAG_TIME_KERNELS = 1
FLAG_NO_GRAPH = 2
FLAG_PARALLEL_SHADE = 4
FLAG_SHADE_BLOCK_128 = 8
FLAG_TUNE_POOL = 16


class _SceneFileInfo(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("spp", ctypes.c_uint32),
                ("integrator", ctypes.c_uint32), ("rng_seed", ctypes.c_uint32), ("brightness", ctypes.c_float),
                ("sensitivity", ctypes.c_float), ("num_meshes", ctypes.c_uint32), ("num_triangles", ctypes.c_uint32),
                ("num_materials", ctypes.c_uint32), ("num_spectra", ctypes.c_uint32)]


class _Settings(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in
                ("width", "height", "spp", "max_depth", "seed", "integrator", "slots_per_pixel", "flags",
                 "tune_trace", "reserved")]


def load_library(path=None):
    """Loads libslr_hip.so; raises RuntimeError if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError("libslr_hip.so not built (%s); run python -m slr_amd.build" % p)
    lib = ctypes.CDLL(p)
    vp, i32, u32, f32, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_float, ctypes.c_double
    pp = ctypes.POINTER(ctypes.c_void_p)
    pi = ctypes.POINTER(ctypes.c_int)
    sig = {
        "slrh_create": (i32, [i32, pp]),
        "slrh_destroy": (i32, [vp]),
        "slrh_last_error": (ctypes.c_char_p, [vp]),
        "slrh_version": (i32, []),
        "slrh_spectrum_rgb": (i32, [vp, i32, f32, f32, f32, pi]),
        "slrh_spectrum_regular": (i32, [vp, f32, f32, vp, i32, pi]),
        "slrh_spectrum_evaluate": (i32, [vp, i32, i32, vp, vp]),
        "slrh_cmf_strata": (i32, [vp, vp, vp, vp]),
        "slrh_material_lambert": (i32, [vp, i32, pi]),
        "slrh_material_oren_nayar": (i32, [vp, i32, f32, pi]),
        "slrh_material_ggx_conductor": (i32, [vp, i32, i32, f32, f32, pi]),
        "slrh_material_ggx_dielectric": (i32, [vp, i32, i32, f32, f32, pi]),
        "slrh_material_diffuse_emitter": (i32, [vp, i32, i32, pi]),
        "slrh_mesh_add": (i32, [vp, u32, vp, vp, vp, u32, vp, vp, i32, pi]),
        "slrh_prototype_add": (i32, [vp, u32, vp, vp, vp, u32, vp, vp, i32, pi]),
        "slrh_instance_add": (i32, [vp, i32, vp, i32, pi]),
        "slrh_scene_compat": (i32, [vp, u32]),
        "slrh_medium_density_grid": (i32, [vp, vp, vp, u32, u32, u32, i32, i32, f32, pi]),
        "slrh_medium_query": (i32, [vp, vp, u32, i32, u32, vp, i32, vp]),
        "slrh_instance_get": (i32, [vp, i32, vp, vp]),
        "slrh_camera_perspective": (i32, [vp, vp, f32, f32, f32, f32, f32, f32]),
        "slrh_scene_create": (i32, [vp, pp]),
        "slrh_scene_build": (i32, [vp]),
        "slrh_scene_destroy": (i32, [vp]),
        "slrh_scene_accel_stats": (i32, [vp, vp]),
        "slrh_scene_build_stats": (i32, [vp, vp]),
        "slrh_scene_sbvh_dump": (i32, [vp, vp, vp, vp, vp, vp]),
        "slrh_scene_qbvh_dump": (i32, [vp, vp, vp, vp, vp, vp]),
        "slrh_scene_accelerators": (i32, [vp, u32]),
        "slrh_scene_query_accel": (i32, [vp, i32]),
        "slrh_scene_standard_bvh_dump": (i32, [vp, vp, vp, vp, vp, vp, vp]),
        "slrh_scene_load_text": (i32, [vp, ctypes.c_char_p, vp]),
        "slrh_scene_geometry": (i32, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "slrh_scene_camera_get": (i32, [vp, vp, vp]),
        "slrh_material_info": (i32, [vp, i32, pi, pi, pi, pi, vp, vp]),
        "slrh_intersect": (i32, [vp, vp, vp, u32, i32, vp]),
        "slrh_scene_query_counts": (i32, [vp, vp, vp]),
        "slrh_occluded": (i32, [vp, vp, vp, u32, i32, vp]),
        "slrh_intersect_noalloc": (i32, [vp, vp, vp, u32, vp]),
        "slrh_film_create": (i32, [vp, u32, u32, pp]),
        "slrh_film_destroy": (i32, [vp]),
        "slrh_film_clear": (i32, [vp]),
        "slrh_film_read": (i32, [vp, i32, f64, vp, ctypes.c_size_t]),
        "slrh_film_load": (i32, [vp, vp, ctypes.c_size_t]),
        "slrh_film_bmp": (i32, [vp, f32, f32, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
        "slrh_render_progressive": (i32, [vp, vp, vp, vp, vp, vp]),
        "slrh_render": (i32, [vp, vp, vp, vp]),
        "slrh_render_batch": (i32, [vp, vp, u32, vp, vp]),
        "slrh_environment": (i32, [vp, vp, u32, u32, f32]),
        "slrh_server_create": (i32, [vp, vp, pp]),
        "slrh_server_destroy": (i32, [vp]),
        "slrh_submit": (i32, [vp, vp, ctypes.POINTER(ctypes.c_uint64)]),
        "slrh_poll": (i32, [vp, ctypes.c_uint64, pi]),
        "slrh_wait": (i32, [vp, ctypes.c_uint64, i32, pi]),
        "slrh_cancel": (i32, [vp, ctypes.c_uint64]),
        "slrh_server_stats": (i32, [vp, vp]),
        "slrh_trace_variants": (i32, [vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


class Context:
    """slrh_create / slrh_destroy: one per HIP device."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        r = self.lib.slrh_create(device, ctypes.byref(h))
        if r != SLRH_OK:
            raise SlrhError("slrh_create failed: %d" % r)
        self.h = h

    def check(self, r, what):
        if r != SLRH_OK:
            msg = self.lib.slrh_last_error(self.h)
            raise SlrhError("%s failed (%d): %s" % (what, r, msg.decode() if msg else ""))

    def close(self):
        if self.h:
            self.lib.slrh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """Scene assembled through the C ABI (libSLR Scene + nodes + materials)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.lib = ctx.lib
        h = ctypes.c_void_p()
        ctx.check(self.lib.slrh_scene_create(ctx.h, ctypes.byref(h)), "slrh_scene_create")
        self.h = h
        self.num_prims = 0

    def _id(self, fn, name, *args):
        out = ctypes.c_int()
        self.ctx.check(fn(self.h, *args, ctypes.byref(out)), name)
        return out.value

    def spectrum_rgb(self, rgb, light=False):
        t = SPECTRUM_LIGHT_SOURCE if light else SPECTRUM_REFLECTANCE
        return self._id(self.lib.slrh_spectrum_rgb, "slrh_spectrum_rgb", t, *[float(x) for x in rgb])

    def spectrum_regular(self, lo, hi, values):
        v = _f32(np.atleast_1d(values))
        return self._id(self.lib.slrh_spectrum_regular, "slrh_spectrum_regular", float(lo), float(hi), _ptr(v), len(v))

    def spectrum_constant(self, value):
        return self.spectrum_regular(360.0, 830.0, [value])

    def spectrum_evaluate(self, sid, lambdas):
        lam = _f32(np.atleast_1d(lambdas))
        out = np.zeros(len(lam), np.float32)
        self.ctx.check(self.lib.slrh_spectrum_evaluate(self.h, sid, len(lam), _ptr(lam), _ptr(out)),
                       "slrh_spectrum_evaluate")
        return out

    def lambert(self, refl):
        return self._id(self.lib.slrh_material_lambert, "slrh_material_lambert", refl)

    def oren_nayar(self, refl, sigma):
        return self._id(self.lib.slrh_material_oren_nayar, "slrh_material_oren_nayar", refl, float(sigma))

    def ggx_conductor(self, eta, k, ax, ay=None):
        return self._id(self.lib.slrh_material_ggx_conductor, "slrh_material_ggx_conductor", eta, k, float(ax),
                        float(ax if ay is None else ay))

    def ggx_dielectric(self, eta_ext, eta_int, ax, ay=None):
        return self._id(self.lib.slrh_material_ggx_dielectric, "slrh_material_ggx_dielectric", eta_ext, eta_int,
                        float(ax), float(ax if ay is None else ay))

    def diffuse_emitter(self, base, emittance):
        return self._id(self.lib.slrh_material_diffuse_emitter, "slrh_material_diffuse_emitter", base, emittance)

    def add_mesh(self, pos, idx, mat, nrm=None, tan=None):
        pos = _f32(pos).reshape(-1, 3)
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 3)
        nrm = None if nrm is None else _f32(nrm).reshape(-1, 3)
        tan = None if tan is None else _f32(tan).reshape(-1, 3)
        if np.isscalar(mat):
            matp, matv = None, int(mat)
        else:
            matp, matv = np.ascontiguousarray(mat, dtype=np.uint32), 0
        mid = self._id(self.lib.slrh_mesh_add, "slrh_mesh_add", len(pos), _ptr(pos), _ptr(nrm), _ptr(tan), len(idx),
                       _ptr(idx), _ptr(matp), matv)
        self.num_prims += len(idx)
        return mid

    def add_prototype(self, pos, idx, mat, nrm=None, tan=None):
        """A mesh in object space that appears only through add_instance."""
        pos = _f32(pos).reshape(-1, 3)
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 3)
        nrm = None if nrm is None else _f32(nrm).reshape(-1, 3)
        tan = None if tan is None else _f32(tan).reshape(-1, 3)
        if np.isscalar(mat):
            matp, matv = None, int(mat)
        else:
            matp, matv = np.ascontiguousarray(mat, dtype=np.uint32), 0
        pid = self._id(self.lib.slrh_prototype_add, "slrh_prototype_add", len(pos), _ptr(pos), _ptr(nrm), _ptr(tan),
                       len(idx), _ptr(idx), _ptr(matp), matv)
        self.num_prims += len(idx)
        return pid

    def add_instance(self, proto, l2w, material=-1):
        """Places prototype `proto` with the object-to-world affine transform
        l2w (3x4 or 4x4, row-major); material >= 0 replaces its materials."""
        m = np.asarray(l2w, dtype=np.float64)
        m = _f32(m.reshape(-1, 4)[:3].copy())
        return self._id(self.lib.slrh_instance_add, "slrh_instance_add", int(proto), _ptr(m), int(material))

    def instance_transform(self, inst):
        """(object-to-world rows 3x4, world-to-object matrix 4x4), float32."""
        a = np.zeros(12, np.float32)
        b = np.zeros(16, np.float32)
        self.ctx.check(self.lib.slrh_instance_get(self.h, int(inst), _ptr(a), _ptr(b)), "slrh_instance_get")
        return a.reshape(3, 4), b.reshape(4, 4)

    def set_medium(self, bounds, density, sigma_s, sigma_e, g=0.0):
        """Heterogeneous medium: bounds (min xyz, max xyz), density (nz, ny, nx)
        lattice, sigma_s / sigma_e spectrum ids (extinction constant), HG g."""
        b = _f32(np.asarray(bounds, np.float32).ravel())
        d = _f32(np.asarray(density, np.float32))
        nz, ny, nx = d.shape
        mid = ctypes.c_int()
        self.ctx.check(self.lib.slrh_medium_density_grid(self.h, _ptr(b), _ptr(d), nx, ny, nz, int(sigma_s),
                                                         int(sigma_e), float(g), ctypes.byref(mid)),
                       "slrh_medium_density_grid")
        return mid.value

    def medium_query(self, rays, mode=0, seed=1):
        """Per ray (n, 8) {org, dir, tmin, tmax}: mode 0 a transmittance
        estimate, 1 a sampled free-path distance (tmax if none)."""
        rays = _f32(rays).reshape(-1, 8)
        n = len(rays)
        cols = [np.ascontiguousarray(rays[:, k]) for k in range(8)]
        r = _Rays(*[c.ctypes.data for c in cols])
        out = np.zeros(n, np.float32)
        self.ctx.check(self.lib.slrh_medium_query(self.h, ctypes.byref(r), n, int(mode), int(seed) & 0xFFFFFFFF,
                                                  _ptr(out), MEM_HOST, None), "slrh_medium_query")
        return out

    def set_compat(self, flags):
        self.ctx.check(self.lib.slrh_scene_compat(self.h, int(flags)), "slrh_scene_compat")

    def set_camera(self, l2w, aspect, fov_y, lens_radius=0.0, img_dist=1.0, obj_dist=1.0, sensitivity=1.0):
        m = _f32(np.asarray(l2w, dtype=np.float32).reshape(4, 4).T.copy())  # column-major
        self.ctx.check(self.lib.slrh_camera_perspective(self.h, _ptr(m), float(sensitivity), float(aspect),
                                                        float(fov_y), float(lens_radius), float(img_dist),
                                                        float(obj_dist)), "slrh_camera_perspective")

    def set_environment(self, rgb, scale=1.0):
        """rgb: (height, width, 3) linear sRGB equirectangular image."""
        img = _f32(np.asarray(rgb, dtype=np.float32))
        h, w = img.shape[0], img.shape[1]
        self.ctx.check(self.lib.slrh_environment(self.h, _ptr(img), w, h, float(scale)), "slrh_environment")

    def build(self):
        self.ctx.check(self.lib.slrh_scene_build(self.h), "slrh_scene_build")

    def accel_stats(self):
        st = np.zeros(9)
        self.ctx.check(self.lib.slrh_scene_accel_stats(self.h, _ptr(st)), "slrh_scene_accel_stats")
        keys = ["sbvh_nodes", "sbvh_refs", "sbvh_depth", "sbvh_cost", "qbvh_nodes", "qbvh_refs", "qbvh_depth",
                "qbvh_cost", "build_seconds"]
        return dict(zip(keys, st.tolist()))

    def build_stats(self):
        """{build_seconds, device_nodes}: all SBVH builds of the scene, and
        their nodes binned on the device (slrh_scene_build_stats)."""
        st = np.zeros(2)
        self.ctx.check(self.lib.slrh_scene_build_stats(self.h, _ptr(st)), "slrh_scene_build_stats")
        return {"build_seconds": float(st[0]), "device_nodes": int(st[1])}

    def _dump(self, fn, nbox, ninfo, name):
        nn, nr = ctypes.c_uint32(), ctypes.c_uint32()
        self.ctx.check(fn(self.h, None, None, None, ctypes.byref(nn), ctypes.byref(nr)), name)
        boxes = np.zeros((nn.value, nbox), np.float32)
        info = np.zeros((nn.value, ninfo), np.uint32)
        refs = np.zeros(nr.value, np.int32)
        self.ctx.check(fn(self.h, _ptr(boxes), _ptr(info), _ptr(refs), None, None), name)
        return boxes, info, refs

    def sbvh_dump(self):
        return self._dump(self.lib.slrh_scene_sbvh_dump, 6, 5, "slrh_scene_sbvh_dump")

    def qbvh_dump(self):
        return self._dump(self.lib.slrh_scene_qbvh_dump, 24, 7, "slrh_scene_qbvh_dump")

    def standard_bvh_dump(self):
        """(boxes (n, 6), info (n, 5), refs, depth, SAH cost) of the binned-SAH BVH."""
        nn, nr = ctypes.c_uint32(), ctypes.c_uint32()
        st = np.zeros(2)
        fn = self.lib.slrh_scene_standard_bvh_dump
        self.ctx.check(fn(self.h, None, None, None, ctypes.byref(nn), ctypes.byref(nr), _ptr(st)),
                       "slrh_scene_standard_bvh_dump")
        boxes = np.zeros((nn.value, 6), np.float32)
        info = np.zeros((nn.value, 5), np.uint32)
        refs = np.zeros(nr.value, np.int32)
        self.ctx.check(fn(self.h, _ptr(boxes), _ptr(info), _ptr(refs), None, None, None),
                       "slrh_scene_standard_bvh_dump")
        return boxes, info, refs, int(st[0]), float(st[1])

    def load_text(self, text):
        """slrh_scene_load_text: adds the scene a scene-language text describes;
        returns the render settings it sets (a dict)."""
        info = _SceneFileInfo()
        self.ctx.check(self.lib.slrh_scene_load_text(self.h, text.encode(), ctypes.byref(info)), "slrh_scene_load_text")
        self.num_prims += info.num_triangles
        return {k: getattr(info, k) for k, _ in _SceneFileInfo._fields_}

    def geometry(self):
        """World-level triangles: dict of pos, nrm, tan (nv, 3), idx (nt, 3), mat (nt,)."""
        nv, nt = ctypes.c_uint32(0), ctypes.c_uint32(0)
        fn = self.lib.slrh_scene_geometry
        self.ctx.check(fn(self.h, ctypes.byref(nv), ctypes.byref(nt), None, None, None, None, None), "slrh_scene_geometry")
        g = {k: np.zeros((nv.value, 3), np.float32) for k in ("pos", "nrm", "tan")}
        g["idx"] = np.zeros((nt.value, 3), np.uint32)
        g["mat"] = np.zeros(nt.value, np.uint32)
        self.ctx.check(fn(self.h, ctypes.byref(nv), ctypes.byref(nt), _ptr(g["pos"]), _ptr(g["nrm"]), _ptr(g["tan"]),
                          _ptr(g["idx"]), _ptr(g["mat"])), "slrh_scene_geometry")
        return g

    def camera(self):
        """(l2w 4x4 row-major, {sensitivity, aspect, fov_y, lens_radius, img_dist, obj_dist})."""
        l2w = np.zeros(16, np.float32)
        p = np.zeros(6, np.float32)
        self.ctx.check(self.lib.slrh_scene_camera_get(self.h, _ptr(l2w), _ptr(p)), "slrh_scene_camera_get")
        keys = ("sensitivity", "aspect", "fov_y", "lens_radius", "img_dist", "obj_dist")
        return l2w.reshape(4, 4).T.copy(), dict(zip(keys, p.tolist()))

    def material_info(self, mid):
        out = [ctypes.c_int() for _ in range(4)]
        p0, p1 = ctypes.c_float(), ctypes.c_float()
        self.ctx.check(self.lib.slrh_material_info(self.h, int(mid), *[ctypes.byref(x) for x in out],
                                                   ctypes.byref(p0), ctypes.byref(p1)), "slrh_material_info")
        return {"model": out[0].value, "spec0": out[1].value, "spec1": out[2].value, "emittance": out[3].value,
                "p0": p0.value, "p1": p1.value}

    def set_accelerators(self, standard_bvh=False):
        self.ctx.check(self.lib.slrh_scene_accelerators(self.h, 1 if standard_bvh else 0), "slrh_scene_accelerators")

    def query_accel(self, standard_bvh):
        """Ray batches through the binned-SAH BVH (True) or the QBVH (False)."""
        self.ctx.check(self.lib.slrh_scene_query_accel(self.h, 1 if standard_bvh else 0), "slrh_scene_query_accel")

    def intersect(self, rays, with_inst=False):
        """rays: (n, 8) float32 {org xyz, dir xyz, tmin, tmax} on the host.
        Returns (t, prim, b0, b1), and inst (-1: world space) with with_inst."""
        rays = _f32(rays).reshape(-1, 8)
        n = len(rays)
        cols = [np.ascontiguousarray(rays[:, k]) for k in range(8)]
        r = _Rays(*[c.ctypes.data for c in cols])
        t = np.zeros(n, np.float32)
        prim = np.zeros(n, np.int32)
        b0 = np.zeros(n, np.float32)
        b1 = np.zeros(n, np.float32)
        inst = np.zeros(n, np.int32) if with_inst else None
        h = _Hits(t.ctypes.data, prim.ctypes.data, b0.ctypes.data, b1.ctypes.data,
                  inst.ctypes.data if with_inst else None)
        self.ctx.check(self.lib.slrh_intersect(self.h, ctypes.byref(r), ctypes.byref(h), n, MEM_HOST, None),
                       "slrh_intersect")
        return (t, prim, b0, b1, inst) if with_inst else (t, prim, b0, b1)

    def intersect_device(self, cols, outs, n, stream=None):
        """Device-pointer form: cols = 8 device addresses, outs = 4 (or 5, with
        instance ids) device addresses."""
        r = _Rays(*cols)
        h = _Hits(*(list(outs) + [None] * (5 - len(outs))))
        self.ctx.check(self.lib.slrh_intersect(self.h, ctypes.byref(r), ctypes.byref(h), n, MEM_DEVICE, stream),
                       "slrh_intersect")

    def intersect_noalloc(self, cols, outs, n, stream=None):
        """slrh_intersect_noalloc: the device-pointer closest hit that
        allocates nothing (asynchronous on `stream`; synchronise before
        reading the outputs). cols = 8 device addresses, outs = 4 or 5."""
        r = _Rays(*cols)
        h = _Hits(*(list(outs) + [None] * (5 - len(outs))))
        self.ctx.check(self.lib.slrh_intersect_noalloc(self.h, ctypes.byref(r), ctypes.byref(h), n, stream),
                       "slrh_intersect_noalloc")

    def query_counts(self):
        """(QBVH nodes popped, triangles tested) of the last intersect /
        occluded batch (slrh_scene_query_counts)."""
        n, t = ctypes.c_double(), ctypes.c_double()
        self.ctx.check(self.lib.slrh_scene_query_counts(self.h, ctypes.byref(n), ctypes.byref(t)),
                       "slrh_scene_query_counts")
        return n.value, t.value

    def occluded(self, rays):
        rays = _f32(rays).reshape(-1, 8)
        n = len(rays)
        cols = [np.ascontiguousarray(rays[:, k]) for k in range(8)]
        r = _Rays(*[c.ctypes.data for c in cols])
        vis = np.zeros(n, np.float32)
        self.ctx.check(self.lib.slrh_occluded(self.h, ctypes.byref(r), _ptr(vis), n, MEM_HOST, None), "slrh_occluded")
        return vis

    def close(self):
        if self.h:
            self.lib.slrh_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Film:
    """FP64 16-stratum film (ImageSensor)."""

    def __init__(self, ctx, width, height):
        self.ctx = ctx
        self.lib = ctx.lib
        self.width, self.height = width, height
        h = ctypes.c_void_p()
        ctx.check(self.lib.slrh_film_create(ctx.h, width, height, ctypes.byref(h)), "slrh_film_create")
        self.h = h

    def clear(self):
        self.ctx.check(self.lib.slrh_film_clear(self.h), "slrh_film_clear")

    def strata(self, scale=1.0):
        out = np.zeros((self.height, self.width, 16), np.float64)
        self.ctx.check(self.lib.slrh_film_read(self.h, FILM_STRATA16_F64, float(scale), _ptr(out), out.nbytes),
                       "slrh_film_read")
        return out

    def xyz(self, scale=1.0):
        out = np.zeros((self.height, self.width, 3), np.float64)
        self.ctx.check(self.lib.slrh_film_read(self.h, FILM_XYZ_F64, float(scale), _ptr(out), out.nbytes),
                       "slrh_film_read")
        return out

    def srgb8(self, scale=1.0):
        out = np.zeros((self.height, self.width, 3), np.uint8)
        self.ctx.check(self.lib.slrh_film_read(self.h, FILM_SRGB8, float(scale), _ptr(out), out.nbytes),
                       "slrh_film_read")
        return out

    def load(self, strata):
        """Replaces the sums with (height, width, 16) float64 values."""
        a = np.ascontiguousarray(strata, np.float64).reshape(self.height, self.width, 16)
        self.ctx.check(self.lib.slrh_film_load(self.h, _ptr(a), a.nbytes), "slrh_film_load")

    def bmp(self, scale=1.0, sensitivity=1.0):
        """The BMP file bytes ImageSensor::saveImage would write for these sums."""
        n = ctypes.c_size_t()
        self.ctx.check(self.lib.slrh_film_bmp(self.h, float(scale), float(sensitivity), None, 0, ctypes.byref(n)),
                       "slrh_film_bmp")
        out = np.zeros(n.value, np.uint8)
        self.ctx.check(self.lib.slrh_film_bmp(self.h, float(scale), float(sensitivity), _ptr(out), out.nbytes,
                                              ctypes.byref(n)), "slrh_film_bmp")
        return out.tobytes()

    def close(self):
        if self.h:
            self.lib.slrh_film_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cmf_strata():
    lib = load_library()
    x, y, z = np.zeros(16, np.float32), np.zeros(16, np.float32), np.zeros(16, np.float32)
    integ = np.zeros(1, np.float32)
    lib.slrh_cmf_strata(_ptr(x), _ptr(y), _ptr(z), _ptr(integ))
    return x, y, z, float(integ[0])


def trace_variants():
    """[(lds_depth, refill_threshold, fused)] of the traversal kernel variants."""
    lib = load_library()
    n = ctypes.c_int(64)
    d = np.zeros(64, np.int32)
    r = np.zeros(64, np.int32)
    f = np.zeros(64, np.int32)
    lib.slrh_trace_variants(ctypes.byref(n), _ptr(d), _ptr(r), _ptr(f))
    return list(zip(d[:n.value].tolist(), r[:n.value].tolist(), f[:n.value].tolist()))


_SNAPSHOT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32)

STAT_KEYS = ["rays", "samples", "device_seconds", "nonfinite", "iterations", "trace_seconds", "trace_launches",
             "extend_rays", "shade_seconds", "other_seconds", "shadow_rays", "nodes_visited", "tris_tested",
             "nodes_shadow", "tris_shadow", "trace_variant", "medium_events", "medium_shadows", "medium_seconds",
             "medium_lookups", "medium_cells", "medium_rays", "medium_guard", "nodes_lds", "inst_entered",
             "shade_block", "slots_per_pixel"]


def _fill_frame(fr, f, default_spp):
    cam = f["camera"]
    m = np.asarray(cam["l2w"], dtype=np.float32).reshape(4, 4).T.ravel()  # column-major
    for k in range(16):
        fr.l2w[k] = float(m[k])
    fr.sensitivity = float(cam.get("sensitivity", 1.0))
    fr.aspect = float(cam["aspect"])
    fr.fov_y = float(cam["fov_y"])
    fr.lens_radius = float(cam.get("lens_radius", 0.0))
    fr.img_dist = float(cam.get("img_dist", 1.0))
    fr.obj_dist = float(cam.get("obj_dist", 1.0))
    fr.width, fr.height = f["film"].width, f["film"].height
    fr.spp = int(f.get("spp", default_spp))
    fr.seed = int(f.get("seed", 1509761209)) & 0xFFFFFFFF
    fr.film = f["film"].h


class PTRenderer:
    """PTRenderer(spp).render(scene, settings) (libSLR/Renderer/PTRenderer.h:16-44)."""

    def __init__(self, spp):
        self.spp = int(spp)

    integrator = 0   # SLRH_PT

    def render(self, scene, film, seed=1509761209, max_depth=0, slots_per_pixel=0, time_kernels=False,
               tune_trace=0, graph=True, parallel_shade=False, shade_block_128=False, tune_pool=False):
        flags = (FLAG_TIME_KERNELS if time_kernels else 0) | (0 if graph else FLAG_NO_GRAPH) | \
            (FLAG_PARALLEL_SHADE if parallel_shade else 0) | (FLAG_SHADE_BLOCK_128 if shade_block_128 else 0) | \
            (FLAG_TUNE_POOL if tune_pool else 0)
        s = _Settings(film.width, film.height, self.spp, max_depth, seed & 0xFFFFFFFF, self.integrator,
                      slots_per_pixel, flags, tune_trace, 0)
        st = np.zeros(len(STAT_KEYS))
        scene.ctx.check(scene.lib.slrh_render(scene.h, ctypes.byref(s), film.h, _ptr(st)), "slrh_render")
        return dict(zip(STAT_KEYS, st.tolist()))

    def render_progressive(self, scene, film, snapshot=None, seed=1509761209, max_depth=0, slots_per_pixel=0):
        """slrh_render_progressive: passes of 1, 1, 2, 4, ... samples; snapshot(film, samples_done) is
        called after each (a true return stops the render)."""
        def cb(user, f, done):
            return 1 if (snapshot is not None and snapshot(film, int(done))) else 0
        fn = _SNAPSHOT_FN(cb)
        s = _Settings(film.width, film.height, self.spp, max_depth, seed & 0xFFFFFFFF, self.integrator,
                      slots_per_pixel, 0, 0, 0)
        st = np.zeros(len(STAT_KEYS))
        scene.ctx.check(scene.lib.slrh_render_progressive(scene.h, ctypes.byref(s), film.h, fn, None, _ptr(st)),
                        "slrh_render_progressive")
        return dict(zip(STAT_KEYS, st.tolist()))

    def render_batch(self, scene, frames, max_depth=0, slots_per_pixel=0, time_kernels=False, tune_trace=0,
                     graph=True, parallel_shade=False, shade_block_128=False, tune_pool=False):
        """frames: list of dicts {film, camera: dict(l2w, aspect, fov_y, lens_radius, img_dist, obj_dist,
        sensitivity), seed, spp (default self.spp)}; all frames in one wavefront."""
        arr = (_Frame * max(1, len(frames)))()
        for i, f in enumerate(frames):
            _fill_frame(arr[i], f, self.spp)
        flags = (FLAG_TIME_KERNELS if time_kernels else 0) | (0 if graph else FLAG_NO_GRAPH) | \
            (FLAG_PARALLEL_SHADE if parallel_shade else 0) | (FLAG_SHADE_BLOCK_128 if shade_block_128 else 0) | \
            (FLAG_TUNE_POOL if tune_pool else 0)
        s = _Settings(0, 0, self.spp, max_depth, 0, self.integrator, slots_per_pixel, flags, tune_trace, 0)
        st = np.zeros(len(STAT_KEYS))
        scene.ctx.check(scene.lib.slrh_render_batch(scene.h, arr, len(frames), ctypes.byref(s), _ptr(st)),
                        "slrh_render_batch")
        return dict(zip(STAT_KEYS, st.tolist()))


class VolumetricPTRenderer(PTRenderer):
    """VolumetricPTRenderer(spp) (libSLR/Renderer/VolumetricPTRenderer.h:16-44):
    the scene's medium scatters and attenuates; no fixed path-length cap."""
    integrator = 1   # SLRH_VOL_PT


# the entries slrh_render fills with the bidirectional integrator's counters
# in place of the medium's (which it does not have)
BPT_STAT_KEYS = {"medium_events": "light_vertices", "medium_shadows": "eye_vertices", "medium_lookups": "connections",
                 "medium_cells": "lens_contributions", "medium_guard": "capped_subpaths"}


class BPTRenderer(PTRenderer):
    """BPTRenderer(spp) (libSLR/Renderer/BPTRenderer.h:18-124): light and eye
    subpaths per pixel sample, every connection between them weighted by the
    reference's MIS; light vertices seen by the lens add to the pixel they
    project to. max_depth caps the eye subpath (0: no cap, as the reference)."""
    integrator = 2   # SLRH_BPT

    @staticmethod
    def _rename(st):
        return {BPT_STAT_KEYS.get(k, k): v for k, v in st.items()}

    def render(self, scene, film, seed=1509761209, max_depth=0, slots_per_pixel=0, time_kernels=False, **kw):
        return self._rename(PTRenderer.render(self, scene, film, seed=seed, max_depth=max_depth,
                                              slots_per_pixel=slots_per_pixel, time_kernels=time_kernels, **kw))

    def render_batch(self, scene, frames, max_depth=0, slots_per_pixel=0, time_kernels=False, **kw):
        return self._rename(PTRenderer.render_batch(self, scene, frames, max_depth=max_depth,
                                                    slots_per_pixel=slots_per_pixel, time_kernels=time_kernels, **kw))


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_CANCELLED, REQ_FAILED = range(5)
SERVER_STAT_KEYS = ["completed", "queued", "running", "p50_ms", "p99_ms", "mean_ms", "samples", "busy_seconds"]


class Server:
    """Continuous-batching render server over a built scene (slrh_server_*)."""

    def __init__(self, scene, path_state_bytes=2 << 30, tune_trace=0):
        self.scene = scene
        self.lib = scene.lib
        h = ctypes.c_void_p()
        cfg = _ServerConfig(int(path_state_bytes), int(tune_trace), 0)
        scene.ctx.check(self.lib.slrh_server_create(scene.h, ctypes.byref(cfg), ctypes.byref(h)), "slrh_server_create")
        self.h = h
        self._films = {}

    def submit(self, film, camera, spp, seed=1509761209, max_depth=0, slots_per_pixel=0):
        req = _Request()
        _fill_frame(req.frame, dict(film=film, camera=camera, spp=spp, seed=seed), spp)
        req.max_depth = int(max_depth)
        req.slots_per_pixel = int(slots_per_pixel)
        t = ctypes.c_uint64()
        self.scene.ctx.check(self.lib.slrh_submit(self.h, ctypes.byref(req), ctypes.byref(t)), "slrh_submit")
        self._films[t.value] = film   # kept alive until the request settles
        return t.value

    def poll(self, ticket):
        st = ctypes.c_int()
        self.scene.ctx.check(self.lib.slrh_poll(self.h, ctypes.c_uint64(ticket), ctypes.byref(st)), "slrh_poll")
        return st.value

    def wait(self, ticket, timeout_ms=-1):
        st = ctypes.c_int()
        self.scene.ctx.check(self.lib.slrh_wait(self.h, ctypes.c_uint64(ticket), int(timeout_ms), ctypes.byref(st)),
                             "slrh_wait")
        return st.value

    def cancel(self, ticket):
        self.scene.ctx.check(self.lib.slrh_cancel(self.h, ctypes.c_uint64(ticket)), "slrh_cancel")

    def stats(self):
        out = np.zeros(len(SERVER_STAT_KEYS))
        self.scene.ctx.check(self.lib.slrh_server_stats(self.h, _ptr(out)), "slrh_server_stats")
        return dict(zip(SERVER_STAT_KEYS, out.tolist()))

    def close(self):
        if self.h:
            self.lib.slrh_server_destroy(self.h)
            self.h = None
            self._films.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_scene(ctx, desc):
    """Builds a Scene from a description dict produced by slr_amd.scenes."""
    sc = Scene(ctx)
    mat_ids = []
    for m in desc["materials"]:
        kind = m["kind"]
        if kind == "lambert" and "const" in m:
            mid = sc.lambert(sc.spectrum_constant(m["const"]))
        elif kind == "lambert":
            mid = sc.lambert(sc.spectrum_rgb(m["rgb"]))
        elif kind == "oren_nayar":
            mid = sc.oren_nayar(sc.spectrum_rgb(m["rgb"]), m["sigma"])
        elif kind == "ggx_conductor":
            mid = sc.ggx_conductor(sc.spectrum_constant(m["eta"]), sc.spectrum_constant(m["k"]), m["alpha"],
                                   m.get("alpha_y"))
        elif kind == "ggx_dielectric":
            mid = sc.ggx_dielectric(sc.spectrum_constant(m["eta_ext"]), sc.spectrum_constant(m["eta_int"]),
                                    m["alpha"], m.get("alpha_y"))
        elif kind == "emitter":
            base = mat_ids[m["base"]] if m.get("base", -1) >= 0 else -1
            le = sc.spectrum_constant(m["const"]) if "const" in m else sc.spectrum_rgb(m["rgb"], light=True)
            mid = sc.diffuse_emitter(base, le)
        else:
            raise ValueError(kind)
        mat_ids.append(mid)
    for mesh in desc["meshes"]:
        mat = mesh["mat"]
        mat = mat_ids[mat] if np.isscalar(mat) else np.asarray(mat_ids, np.uint32)[np.asarray(mat)]
        sc.add_mesh(mesh["pos"], mesh["idx"], mat, mesh.get("nrm"), mesh.get("tan"))
    for mesh in desc.get("prototypes", []):
        mat = mesh["mat"]
        mat = mat_ids[mat] if np.isscalar(mat) else np.asarray(mat_ids, np.uint32)[np.asarray(mat)]
        sc.add_prototype(mesh["pos"], mesh["idx"], mat, mesh.get("nrm"), mesh.get("tan"))
    for inst in desc.get("instances", []):
        m = inst.get("mat", -1)
        sc.add_instance(inst["proto"], inst["l2w"], mat_ids[m] if m >= 0 else -1)
    flags = (COMPAT_INSTANCE_SHADOW if desc.get("compat_instance_shadow") else 0) | \
        (COMPAT_VOLPT_TANGENT_NEE if desc.get("compat_volpt_tangent_nee") else 0)
    if flags:
        sc.set_compat(flags)
    med = desc.get("medium")
    if med is not None:
        sc.set_medium(med["bounds"], med["density"], sc.spectrum_constant(med["sigma_s"]),
                      sc.spectrum_constant(med["sigma_e"]), med.get("g", 0.0))
    if desc.get("standard_bvh"):
        sc.set_accelerators(standard_bvh=True)
    cam = desc["camera"]
    sc.set_camera(cam["l2w"], cam["aspect"], cam["fov_y"], cam.get("lens_radius", 0.0), cam.get("img_dist", 1.0),
                  cam.get("obj_dist", 1.0), cam.get("sensitivity", 1.0))
    if desc.get("environment") is not None:
        env = desc["environment"]
        sc.set_environment(env["rgb"], env.get("scale", 1.0))
    sc.build()
    return sc
