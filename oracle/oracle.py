"""ctypes access to the CPU restatement (oracle/slr_oracle.c).

Test infrastructure only (tests, the smoke check, bench.py's cpu_baseline
leg). build() compiles it with gcc into oracle/_build/liboracle.so.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "slr_oracle.c")
LIB = os.path.join(HERE, "_build", "liboracle.so")
_lib = None

KIND = {"lambert": 0, "oren_nayar": 1, "ggx_conductor": 2, "ggx_dielectric": 3}


def build():
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    # no -march: SSE2 only and no fused multiply-add, like the reference build
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-o", tmp, SRC, "-lm"],
                   check=True)
    os.replace(tmp, LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, i32, u32, f32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_float, ctypes.c_longlong
        sig = {
            "orc_xs_next": (u32, [vp]),
            "orc_xs_float": (f32, [vp]),
            "orc_xs_seed": (None, [u32, vp]),
            "orc_slot_seed": (None, [u32, u32, vp]),
            "orc_wavelengths": (None, [f32, f32, vp, vp, vp]),
            "orc_spectrum_eval": (None, [f32, f32, i32, vp, vp, vp]),
            "orc_film_bin": (None, [vp, vp, vp]),
            "orc_sampling": (None, [i32, vp, vp, vp, vp]),
            "orc_bsdf": (i32, [i32, vp, vp, vp, i32, vp, vp, i32, vp, vp, vp, vp, vp]),
            "orc_scene_new": (vp, [i32, vp, vp, vp, i32, vp, vp, i32, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp,
                                   i32, vp, vp, f32, f32, f32, f32, f32]),
            "orc_scene_free": (None, [vp]),
            "orc_scene_compat": (None, [vp, i32]),
            "orc_intersect": (i32, [vp, i32, vp, vp, vp, vp, vp]),
            "orc_visibility": (i32, [vp, i32, vp, vp]),
            "orc_qbvh_trace": (i32, [vp, vp, vp, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp]),
            "orc_render": (i64, [vp, u32, u32, u32, u32, u32, u32, u32, u32, vp, vp]),
        }
        for k, (r, a) in sig.items():
            fn = getattr(L, k)
            fn.restype = r
            fn.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def xs_stream(seed, n):
    st = np.zeros(4, np.uint32)
    lib().orc_xs_seed(seed, _p(st))
    u = np.array([lib().orc_xs_next(_p(st)) for _ in range(n)], np.uint32)
    st2 = np.zeros(4, np.uint32)
    lib().orc_xs_seed(seed, _p(st2))
    f = np.array([lib().orc_xs_float(_p(st2)) for _ in range(n)], np.float32)
    return u, f


def wavelengths(off, u):
    l = np.zeros(16, np.float32)
    hero = np.zeros(1, np.int32)
    pdf = np.zeros(1, np.float32)
    lib().orc_wavelengths(off, u, _p(l), _p(hero), _p(pdf))
    return l, int(hero[0]), float(pdf[0])


def film_bin(lambdas, values):
    l = np.ascontiguousarray(lambdas, np.float32)
    v = np.ascontiguousarray(values, np.float32)
    out = np.zeros(16, np.float32)
    lib().orc_film_bin(_p(l), _p(v), _p(out))
    return out


def sampling(u):
    u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
    n = len(u)
    disk, hemi, tri = np.zeros((n, 2), np.float32), np.zeros((n, 3), np.float32), np.zeros((n, 2), np.float32)
    lib().orc_sampling(n, _p(u), _p(disk), _p(hemi), _p(tri))
    return disk, hemi, tri


def bsdf(kind, A, B, params, queries, wl, mode, u_or_dir):
    q = np.ascontiguousarray(queries, np.float32).reshape(-1, 6)
    n = len(q)
    A = np.ascontiguousarray(A, np.float32)
    B = None if B is None else np.ascontiguousarray(B, np.float32)
    p = np.ascontiguousarray(np.resize(np.asarray(params, np.float32), 2))
    wl = np.ascontiguousarray(np.broadcast_to(np.asarray(wl, np.int32), (n,)))
    u = np.ascontiguousarray(u_or_dir, np.float32).reshape(-1, 3)
    od, op = np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
    fs, ot = np.zeros((n, 16), np.float32), np.zeros(n, np.int32)
    lib().orc_bsdf(kind, _p(A), _p(B), _p(p), n, _p(q), _p(wl), mode, _p(u), _p(od), _p(op), _p(fs), _p(ot))
    return {"dir": od, "pdf": op, "fs": fs, "type": ot}


GRID = (360.0 + 5.0 * np.arange(95)).astype(np.float32)


def reference_spectra(desc):
    """Per-material spectra as regular tables: RGB colours upsampled by the
    compiled reference at its 95 table wavelengths (360..830 nm step 5),
    constants as 1-sample tables."""
    from oracle import refdrv
    specs = []
    mats = []
    for m in desc["materials"]:
        k = m["kind"]
        ent = {"kind": KIND.get(k, 4), "s0": -1, "s1": -1, "emit": -1, "p0": 0.0, "p1": 0.0}
        if k == "lambert" and "const" in m:
            specs.append((360.0, 830.0, np.array([m["const"]], np.float32)))
            ent["s0"] = len(specs) - 1
        elif k in ("lambert", "oren_nayar"):
            specs.append((360.0, 830.0, refdrv.upsample(0, m["rgb"], GRID)))
            ent["s0"] = len(specs) - 1
            if k == "oren_nayar":
                ent["p0"] = float(m["sigma"])
        elif k == "ggx_conductor":
            specs.append((360.0, 830.0, np.array([m["eta"]], np.float32)))
            specs.append((360.0, 830.0, np.array([m["k"]], np.float32)))
            ent.update(s0=len(specs) - 2, s1=len(specs) - 1, p0=m["alpha"], p1=m.get("alpha_y", m["alpha"]))
        elif k == "ggx_dielectric":
            specs.append((360.0, 830.0, np.array([m["eta_ext"]], np.float32)))
            specs.append((360.0, 830.0, np.array([m["eta_int"]], np.float32)))
            ent.update(s0=len(specs) - 2, s1=len(specs) - 1, p0=m["alpha"], p1=m.get("alpha_y", m["alpha"]))
        elif k == "emitter":
            base = m.get("base", -1)
            if base >= 0:
                ent = dict(mats[base])
            else:
                ent["kind"] = 4
            if "const" in m:
                specs.append((360.0, 830.0, np.array([m["const"]], np.float32)))
            else:
                specs.append((360.0, 830.0, refdrv.upsample(1, m["rgb"], GRID)))
            ent["emit"] = len(specs) - 1
        mats.append(ent)
    return mats, specs


class OracleScene:
    def __init__(self, desc, mats=None, specs=None):
        if mats is None:
            mats, specs = reference_spectra(desc)
        L = lib()
        pos, nrm, tan, idx, tmat = [], [], [], [], []
        base = 0
        for mesh in desc["meshes"]:
            pos.append(mesh["pos"]); nrm.append(mesh["nrm"]); tan.append(mesh["tan"])
            idx.append(np.asarray(mesh["idx"], np.uint32) + base)
            m = mesh["mat"]
            tmat.append(np.full(len(mesh["idx"]), m, np.uint32) if np.isscalar(m) else np.asarray(m, np.uint32))
            base += len(mesh["pos"])
        self.pos = np.ascontiguousarray(np.concatenate(pos), np.float32)
        self.nrm = np.ascontiguousarray(np.concatenate(nrm), np.float32)
        self.tan = np.ascontiguousarray(np.concatenate(tan), np.float32)
        self.idx = np.ascontiguousarray(np.concatenate(idx), np.uint32)
        self.tmat = np.ascontiguousarray(np.concatenate(tmat), np.uint32)
        arr = lambda key, dt: np.ascontiguousarray([m[key] for m in mats], dt)
        self.kind, self.s0, self.s1, self.emit = arr("kind", np.int32), arr("s0", np.int32), arr("s1", np.int32), \
            arr("emit", np.int32)
        self.p0, self.p1 = arr("p0", np.float32), arr("p1", np.float32)
        self.lo = np.ascontiguousarray([s[0] for s in specs], np.float32)
        self.hi = np.ascontiguousarray([s[1] for s in specs], np.float32)
        self.n = np.ascontiguousarray([len(s[2]) for s in specs], np.int32)
        self.off = np.ascontiguousarray(np.concatenate([[0], np.cumsum(self.n)[:-1]]), np.int32)
        self.pool = np.ascontiguousarray(np.concatenate([s[2] for s in specs]), np.float32)
        cam = desc["camera"]
        self.l2w = np.ascontiguousarray(np.asarray(cam["l2w"], np.float32).reshape(4, 4).T)
        self.h = L.orc_scene_new(len(self.pos), _p(self.pos), _p(self.nrm), _p(self.tan), len(self.idx), _p(self.idx),
                                 _p(self.tmat), len(mats), _p(self.kind), _p(self.s0), _p(self.s1), _p(self.emit),
                                 _p(self.p0), _p(self.p1), len(specs), _p(self.lo), _p(self.hi), _p(self.off),
                                 _p(self.n), len(self.pool), _p(self.pool), _p(self.l2w), cam["aspect"], cam["fov_y"],
                                 cam.get("lens_radius", 0.0), cam.get("img_dist", 1.0), cam.get("obj_dist", 1.0))
        self.L = L
        self.tangent_nee = bool(desc.get("compat_volpt_tangent_nee"))

    def intersect(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        t, prim = np.zeros(n, np.float32), np.zeros(n, np.int32)
        b0, b1 = np.zeros(n, np.float32), np.zeros(n, np.float32)
        self.L.orc_intersect(self.h, n, _p(rays), _p(t), _p(prim), _p(b0), _p(b1))
        return t, prim, b0, b1

    def qbvh_trace(self, dump, rays, any_hit=False):
        """QBVH::commonProcedure over a QBVH dump (Scene.qbvh_dump of the
        same scene: boxes, info, refs): (t, prim, b0, b1) for closest hits or
        the visibility for any hit, and the nodes popped and triangles
        tested over all rays."""
        boxes, info, refs = dump[0], dump[1], dump[2]
        boxes = np.ascontiguousarray(boxes, np.float32)
        info = np.ascontiguousarray(info, np.uint32)
        refs = np.ascontiguousarray(refs, np.int32)
        # every reference in exactly one leaf (leaf counts fit the 4-bit field)
        words = info[:, :4].ravel()
        leaf = (words != 0xFFFFFFFF) & ((words >> 31) == 1)
        assert int(((words[leaf] >> 27) & 0xF).sum()) == len(refs), "a leaf count does not fit the dump's 4 bits"
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        t, prim = np.zeros(n, np.float32), np.zeros(n, np.int32)
        b0, b1, vis = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        counts = np.zeros(2)
        r = self.L.orc_qbvh_trace(self.h, _p(boxes), _p(info), _p(refs), n, _p(rays), 1 if any_hit else 0, _p(t),
                                  _p(prim), _p(b0), _p(b1), _p(vis), _p(counts))
        if r != 0:
            raise RuntimeError("QBVH traversal stack overflow")
        if any_hit:
            return vis, counts[0], counts[1]
        return (t, prim, b0, b1), counts[0], counts[1]

    def visibility(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        vis = np.zeros(len(rays), np.float32)
        self.L.orc_visibility(self.h, len(rays), _p(rays), _p(vis))
        return vis

    def render(self, W, H, spp, seed, K=1, max_depth=0, slots=None, threads=1, volumetric=False):
        """Film (H, W, 16) float64 of raw sums, rays traced. threads > 1 splits
        the slots over a thread pool (the C call releases the GIL).
        volumetric: VolumetricPTRenderer over a scene without media (no path
        length cap; with the scene's compat_volpt_tangent_nee, the reference's
        NaN for a light sample in the shading plane)."""
        self.L.orc_scene_compat(self.h, 1 if volumetric and self.tangent_nee else 0)
        if volumetric and max_depth == 0:
            max_depth = 0xFFFFFFFF
        film = np.zeros((H, W, 16), np.float64)
        nonfin = np.zeros(1, np.int64)
        total = K * W * H
        b, e = (0, total) if slots is None else slots
        if threads <= 1:
            rays = self.L.orc_render(self.h, W, H, spp, seed & 0xFFFFFFFF, K, max_depth, b, e, _p(film), _p(nonfin))
            return film, rays
        from concurrent.futures import ThreadPoolExecutor
        # split on pixel boundaries of each plane so no two threads share a film entry
        P = W * H
        chunks = []
        step = max(1, (e - b + threads - 1) // threads)
        films = []
        for c0 in range(b, e, step):
            chunks.append((c0, min(e, c0 + step)))
            films.append(np.zeros((H, W, 16), np.float64))

        def run(i):
            c0, c1 = chunks[i]
            nf = np.zeros(1, np.int64)
            return self.L.orc_render(self.h, W, H, spp, seed & 0xFFFFFFFF, K, max_depth, c0, c1, _p(films[i]),
                                     _p(nf))

        with ThreadPoolExecutor(threads) as ex:
            rays = sum(ex.map(run, range(len(chunks))))
        for f in films:
            film += f
        return film, rays

    def close(self):
        if self.h:
            self.L.orc_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
