"""Seeded synthetic scenes of the benchmark configurations.

Every generator is deterministic for a given seed and returns a plain
description dict that slr_amd.build_scene (and the test harness around the
reference renderer) turns into a scene:

    {"materials": [{"kind": ..., ...}, ...],
     "meshes": [{"pos": (n,3) f32, "nrm": (n,3), "tan": (n,3), "idx": (m,3) u32,
                 "mat": int or (m,) material indices}, ...],
     "camera": {"l2w": 4x4, "aspect", "fov_y", "lens_radius", "img_dist", "obj_dist"}}

Geometry is in the reference test scene's units scaled by 0.01
(TestScenes/Original_Cornell_Box.txt:152-156): the box spans 5.5 units.
"""
import math

import numpy as np

SCALE = 0.01


def _rot_y(deg):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], np.float64)


def _tangent_of(n):
    # makeCoordinateSystem (libSLR/BasicTypes/Vector3D.h:61-67)
    n = np.asarray(n, np.float64)
    sign = np.where(n[..., 2] >= 0, 1.0, -1.0)
    a = -1.0 / (sign + n[..., 2])
    b = n[..., 0] * n[..., 1] * a
    return np.stack([1.0 + sign * n[..., 0] ** 2 * a, sign * b, -sign * n[..., 0]], -1)


def _quad(p0, p1, p2, p3, normal):
    pos = np.array([p0, p1, p2, p3], np.float64)
    nrm = np.tile(np.asarray(normal, np.float64), (4, 1))
    return pos, nrm, np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def _box(center, size, rot_deg):
    """Axis-aligned box rotated about y; 6 faces, 12 triangles, flat normals."""
    hx, hy, hz = [s * 0.5 for s in size]
    R = _rot_y(rot_deg)
    faces = [
        ((+1, 0, 0), [(hx, -hy, -hz), (hx, hy, -hz), (hx, hy, hz), (hx, -hy, hz)]),
        ((-1, 0, 0), [(-hx, -hy, hz), (-hx, hy, hz), (-hx, hy, -hz), (-hx, -hy, -hz)]),
        ((0, +1, 0), [(-hx, hy, -hz), (-hx, hy, hz), (hx, hy, hz), (hx, hy, -hz)]),
        ((0, -1, 0), [(-hx, -hy, hz), (-hx, -hy, -hz), (hx, -hy, -hz), (hx, -hy, hz)]),
        ((0, 0, +1), [(hx, -hy, hz), (hx, hy, hz), (-hx, hy, hz), (-hx, -hy, hz)]),
        ((0, 0, -1), [(-hx, -hy, -hz), (-hx, hy, -hz), (hx, hy, -hz), (hx, -hy, -hz)]),
    ]
    pos, nrm, idx = [], [], []
    for k, (n, vs) in enumerate(faces):
        for v in vs:
            pos.append(R @ np.array(v) + np.asarray(center))
            nrm.append(R @ np.array(n, np.float64))
        b = 4 * k
        idx += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
    return np.array(pos), np.array(nrm), np.array(idx, np.uint32)


def _mesh(pos, nrm, idx, mat, scale=SCALE):
    pos = np.asarray(pos, np.float64) * scale
    nrm = np.asarray(nrm, np.float64)
    nrm = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    return {"pos": pos.astype(np.float32), "nrm": nrm.astype(np.float32),
            "tan": _tangent_of(nrm).astype(np.float32), "idx": np.asarray(idx, np.uint32), "mat": mat}


def cornell_box(width=512, height=512, light_rgb=(17.0, 12.0, 4.0)):
    """Cornell box of 5 walls, two boxes rotated +-18 degrees, one 130x105
    quad light with diffuse emission, Lambert albedos white 0.75, red
    (0.63, 0.06, 0.04), green (0.15, 0.48, 0.09); pinhole camera."""
    W, R, G, L = 0, 1, 2, 3
    materials = [
        {"kind": "lambert", "rgb": (0.75, 0.75, 0.75)},
        {"kind": "lambert", "rgb": (0.63, 0.06, 0.04)},
        {"kind": "lambert", "rgb": (0.15, 0.48, 0.09)},
        # the emitter scatters like the white walls: a path that hits the light
        # keeps going (EmitterSurfaceMaterial over a base material)
        {"kind": "emitter", "base": 0, "rgb": tuple(light_rgb)},
    ]
    S = 550.0
    meshes = []
    walls = [
        (_quad((0, 0, 0), (0, 0, S), (S, 0, S), (S, 0, 0), (0, 1, 0)), W),        # floor
        (_quad((0, S, 0), (S, S, 0), (S, S, S), (0, S, S), (0, -1, 0)), W),       # ceiling
        (_quad((0, 0, S), (0, S, S), (S, S, S), (S, 0, S), (0, 0, -1)), W),       # back
        (_quad((S, 0, 0), (S, 0, S), (S, S, S), (S, S, 0), (-1, 0, 0)), R),       # left (red)
        (_quad((0, 0, 0), (0, S, 0), (0, S, S), (0, 0, S), (1, 0, 0)), G),        # right (green)
    ]
    for (pos, nrm, idx), m in walls:
        meshes.append(_mesh(pos, nrm, idx, m))
    ly = S - 1.0
    pos, nrm, idx = _quad((213, ly, 227), (343, ly, 227), (343, ly, 332), (213, ly, 332), (0, -1, 0))
    meshes.append(_mesh(pos, nrm, idx, L))
    for center, size, rot in (((185.5, 82.5, 169.0), (165, 165, 165), -18.0),
                              ((368.5, 165.0, 351.25), (165, 330, 165), 18.0)):
        pos, nrm, idx = _box(center, size, rot)
        meshes.append(_mesh(pos, nrm, idx, W))
    l2w = np.eye(4)
    l2w[:3, 3] = np.array([278.0, 273.0, -800.0]) * SCALE
    camera = {"l2w": l2w, "aspect": width / height, "fov_y": 0.6860478808, "lens_radius": 0.0,
              "img_dist": 1.0, "obj_dist": 4.5, "sensitivity": 1.0, "width": width, "height": height}
    return {"materials": materials, "meshes": meshes, "camera": camera}


def furnace(width=32, height=32, albedo=0.5, emittance=1.0):
    """Closed cube, every face a diffuse emitter of constant spectrum
    `emittance` over a Lambert base of constant `albedo`, camera inside.
    Radiance everywhere is (emittance / pi) / (1 - albedo)."""
    materials = [{"kind": "lambert", "const": albedo},
                 {"kind": "emitter", "base": 0, "const": emittance}]
    meshes = []
    S = 100.0
    faces = [
        ((0, 0, 0), (0, S, 0), (S, S, 0), (S, 0, 0), (0, 0, 1)),
        ((0, 0, S), (S, 0, S), (S, S, S), (0, S, S), (0, 0, -1)),
        ((0, 0, 0), (0, 0, S), (0, S, S), (0, S, 0), (1, 0, 0)),
        ((S, 0, 0), (S, S, 0), (S, S, S), (S, 0, S), (-1, 0, 0)),
        ((0, 0, 0), (S, 0, 0), (S, 0, S), (0, 0, S), (0, 1, 0)),
        ((0, S, 0), (0, S, S), (S, S, S), (S, S, 0), (0, -1, 0)),
    ]
    for f in faces:
        pos, nrm, idx = _quad(*f)
        meshes.append(_mesh(pos, nrm, idx, 1))
    l2w = np.eye(4)
    l2w[:3, 3] = np.array([50.0, 50.0, 20.0]) * SCALE
    camera = {"l2w": l2w, "aspect": width / height, "fov_y": 1.0, "lens_radius": 0.0, "img_dist": 1.0,
              "obj_dist": 1.0, "sensitivity": 1.0, "width": width, "height": height}
    return {"materials": materials, "meshes": meshes, "camera": camera}


def pixel_cos4(desc, width, height, n=64):
    """Per-pixel mean of cos^4 of the pinhole ray against the optical axis
    times imgPlaneArea / imgDist^2: the film value of a constant unit radiance
    per sample and stratum (PTRenderer.cpp:129 with PerspectiveIDF's dirPDF,
    PerspectiveCamera.cpp:71-82)."""
    cam = desc["camera"]
    opH = 2.0 * cam["obj_dist"] * math.tan(cam["fov_y"] * 0.5)
    opW = opH * cam["aspect"]
    area = opW * opH * (cam["img_dist"] / cam["obj_dist"]) ** 2
    g = (np.arange(n) + 0.5) / n
    ux = (np.arange(width)[:, None] + g[None, :]) / width
    uy = (np.arange(height)[:, None] + g[None, :]) / height
    x = opW * (0.5 - ux)
    y = opH * (0.5 - uy)
    z = cam["obj_dist"]
    r2 = x[None, :, :, None] ** 2 + y[:, None, None, :] ** 2 + z * z
    cos4 = (z * z / r2) ** 2
    return cos4.mean(axis=(2, 3)) * area / cam["img_dist"] ** 2


# ------------------------------------------------------------------ icosphere

def icosphere(subdiv):
    """Unit icosphere: 20 * 4^subdiv triangles, shared vertices."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        a = base + inv[:nf]
        b = base + inv[nf:2 * nf]
        c = base + inv[2 * nf:]
        f = np.concatenate([np.stack([f[:, 0], a, c], 1), np.stack([a, f[:, 1], b], 1),
                            np.stack([c, b, f[:, 2]], 1), np.stack([a, b, c], 1)])
    return v, f.astype(np.uint32)


def _perlin3(p, perm):
    """Classic improved Perlin noise (Perlin 2002) with a seeded permutation."""
    pi = np.floor(p).astype(np.int64)
    pf = p - pi
    xi, yi, zi = (pi[:, 0] & 255), (pi[:, 1] & 255), (pi[:, 2] & 255)
    xf, yf, zf = pf[:, 0], pf[:, 1], pf[:, 2]

    def fade(t):
        return t * t * t * (t * (t * 6 - 15) + 10)

    u, v, w = fade(xf), fade(yf), fade(zf)
    P = np.concatenate([perm, perm])

    def grad(h, x, y, z):
        h = h & 15
        uu = np.where(h < 8, x, y)
        vv = np.where(h < 4, y, np.where((h == 12) | (h == 14), x, z))
        return np.where(h & 1, -uu, uu) + np.where(h & 2, -vv, vv)

    A = P[xi] + yi
    AA, AB = P[A] + zi, P[A + 1] + zi
    B = P[xi + 1] + yi
    BA, BB = P[B] + zi, P[B + 1] + zi

    def lerp(t, a, b):
        return a + t * (b - a)

    x1 = lerp(u, grad(P[AA], xf, yf, zf), grad(P[BA], xf - 1, yf, zf))
    x2 = lerp(u, grad(P[AB], xf, yf - 1, zf), grad(P[BB], xf - 1, yf - 1, zf))
    y1 = lerp(v, x1, x2)
    x1 = lerp(u, grad(P[AA + 1], xf, yf, zf - 1), grad(P[BA + 1], xf - 1, yf, zf - 1))
    x2 = lerp(u, grad(P[AB + 1], xf, yf - 1, zf - 1), grad(P[BB + 1], xf - 1, yf - 1, zf - 1))
    y2 = lerp(v, x1, x2)
    return lerp(w, y1, y2)


def fbm(p, seed, octaves=4, lacunarity=2.0, gain=0.5, freq=2.0):
    rng = np.random.RandomState(seed)
    perm = rng.permutation(256).astype(np.int64)
    total = np.zeros(len(p))
    amp, f = 1.0, freq
    norm = 0.0
    for _ in range(octaves):
        total += amp * _perlin3(p * f, perm)
        norm += amp
        amp *= gain
        f *= lacunarity
    return total / norm


def smooth_normals(pos, idx):
    p = pos.astype(np.float64)
    fn = np.cross(p[idx[:, 1]] - p[idx[:, 0]], p[idx[:, 2]] - p[idx[:, 0]])
    ln = np.linalg.norm(fn, axis=1, keepdims=True)
    fn = np.where(ln > 0, fn / np.maximum(ln, 1e-30), 0)
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, idx[:, k], fn)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    return np.where(ln > 0, n / np.maximum(ln, 1e-30), np.array([0.0, 0.0, 1.0]))


def displaced_icosphere(subdiv, radius, center, seed=7, amplitude=0.05, octaves=4):
    """Icosphere with radial displacement amplitude * r * fBm(p) (seeded Perlin)."""
    v, f = icosphere(subdiv)
    d = fbm(v, seed, octaves=octaves)
    p = v * (radius * (1.0 + amplitude * d))[:, None] + np.asarray(center, np.float64)
    return p, f


def cornell_icosphere(width=1920, height=1080, subdiv=8, seed=7):
    """Cornell box plus a displaced icosphere (GGX conductor, alpha 0.2,
    eta 0.2, k 3.9), the 1.31M-triangle configuration at subdiv 8."""
    desc = cornell_box(width, height)
    desc["materials"].append({"kind": "ggx_conductor", "eta": 0.2, "k": 3.9, "alpha": 0.2})
    mat = len(desc["materials"]) - 1
    p, f = displaced_icosphere(subdiv, 100.0, (150.0, 320.0, 380.0), seed=seed)
    p *= SCALE
    n = smooth_normals(p, f)
    desc["meshes"].append({"pos": p.astype(np.float32), "nrm": n.astype(np.float32),
                           "tan": _tangent_of(n).astype(np.float32), "idx": f, "mat": mat})
    return desc


def cornell_sphere(width=32, height=32, material=None, subdiv=3, amplitude=0.05):
    """Cornell box plus a displaced icosphere (radius 100, smooth vertex
    normals) of the given material dict, for per-material parity: its
    shading normals differ from the geometric ones, so the BSDF front end's
    shading-normal correction is exercised."""
    desc = cornell_box(width, height)
    desc["materials"].append(dict(material or {"kind": "lambert", "rgb": (0.75, 0.75, 0.75)}))
    mat = len(desc["materials"]) - 1
    p, f = displaced_icosphere(subdiv, 100.0, (300.0, 300.0, 300.0), seed=3, amplitude=amplitude)
    p *= SCALE
    n = smooth_normals(p, f)
    desc["meshes"].append({"pos": p.astype(np.float32), "nrm": n.astype(np.float32),
                           "tan": _tangent_of(n).astype(np.float32), "idx": f, "mat": mat})
    return desc


def _rotation(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    x, y, z = axis
    c, s = math.cos(angle), math.sin(angle)
    C = 1 - c
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


def random_affine(rng, center_lo, center_hi, scale_lo=0.6, scale_hi=1.3, uniform=True):
    """Seeded object-to-world affine transform (3x4): a random rotation, a
    scale (uniform or per axis) and a translation in [center_lo, center_hi]."""
    R = _rotation(rng.normal(size=3), rng.uniform(0, 2 * math.pi))
    sc = rng.uniform(scale_lo, scale_hi, size=1 if uniform else 3) * np.ones(3)
    m = np.zeros((3, 4))
    m[:, :3] = R @ np.diag(sc)
    m[:, 3] = rng.uniform(center_lo, center_hi)
    return m


def flatten_instances(desc):
    """The same scene with every instance's triangles transformed into world
    space as plain meshes (for checking two-level traversal against the
    single-level one). Returns (desc, prim_map): prim_map[(inst, proto prim)]
    -> flattened prim id is given as an array per instance."""
    out = {k: v for k, v in desc.items() if k not in ("prototypes", "instances")}
    meshes = list(desc["meshes"])
    base = sum(len(m["idx"]) for m in meshes)
    proto_first = []
    first = base
    for pr in desc.get("prototypes", []):
        proto_first.append(first)
        first += len(pr["idx"])
    maps = []
    nflat = base
    for inst in desc.get("instances", []):
        pr = desc["prototypes"][inst["proto"]]
        m = np.asarray(inst["l2w"], np.float64).reshape(-1, 4)[:3]
        L, t = m[:, :3], m[:, 3]
        pos = np.asarray(pr["pos"], np.float64) @ L.T + t
        N = np.linalg.inv(L).T
        nrm = np.asarray(pr["nrm"], np.float64) @ N.T
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        tan = np.asarray(pr["tan"], np.float64) @ L.T
        tan /= np.linalg.norm(tan, axis=1, keepdims=True)
        mat = inst["mat"] if inst.get("mat", -1) >= 0 else pr["mat"]
        meshes.append({"pos": pos.astype(np.float32), "nrm": nrm.astype(np.float32), "tan": tan.astype(np.float32),
                       "idx": pr["idx"], "mat": mat})
        maps.append(nflat + np.arange(len(pr["idx"])))
        nflat += len(pr["idx"])
    out["meshes"] = meshes
    return out, maps, proto_first


def instanced_cornell(width=64, height=64, n_instances=8, seed=11, uniform_scale=True, instanced_light=False):
    """The Cornell box walls and light with the two boxes replaced by
    instances of two prototypes, a box (12 triangles, white Lambert) and an
    icosphere of subdivision 2 (320 triangles, GGX conductor), under seeded
    random rotations, scales and positions inside the box. instanced_light:
    the ceiling light is a third prototype placed by a rigid instance."""
    desc = cornell_box(width, height)
    light = desc["meshes"][5]
    desc["meshes"] = desc["meshes"][:5 if instanced_light else 6]
    desc["materials"].append({"kind": "ggx_conductor", "eta": 0.2, "k": 3.9, "alpha": 0.2})
    pos, nrm, idx = _box((0.0, 0.0, 0.0), (100.0, 100.0, 100.0), 0.0)
    box = _mesh(pos, nrm, idx, 0)
    v, f = icosphere(2)
    sph = _mesh(v * 60.0, v, f, len(desc["materials"]) - 1)
    desc["prototypes"] = [box, sph]
    rng = np.random.default_rng(seed)
    insts = []
    for i in range(n_instances):
        m = random_affine(rng, np.array([100.0, 60.0, 100.0]) * SCALE, np.array([450.0, 300.0, 450.0]) * SCALE,
                          uniform=uniform_scale)
        insts.append({"proto": i % 2, "l2w": m})
    if instanced_light:
        # the light quad about its centre, placed back by a quarter turn
        # about y and a translation (its footprint is symmetric enough to
        # stay under the ceiling)
        c = np.asarray(light["pos"], np.float64).mean(0)
        desc["prototypes"].append({"pos": (np.asarray(light["pos"], np.float64) - c).astype(np.float32),
                                   "nrm": light["nrm"], "tan": light["tan"], "idx": light["idx"],
                                   "mat": light["mat"]})
        m = np.zeros((3, 4))
        m[:, :3] = _rot_y(90.0)
        m[:, 3] = c
        insts.append({"proto": 2, "l2w": m})
    desc["instances"] = insts
    return desc


def fbm_density(n, seed=5, octaves=5, freq=2.0, bias=0.0, scale=1.0, chunk_points=1 << 20):
    """(n, n, n) lattice (z, y, x) of max(0, scale * (fBm(p) + bias)) over the
    unit cube: a seeded Perlin fBm cloud. The defaults give the volumetric
    configuration's density max(0, fBm) (about 60 % of the lattice non-zero,
    mean 0.09, maximum 0.6 for 5 octaves), which the medium's sigma_t scales.
    Evaluated a few z slices at a time to bound the temporaries."""
    g = (np.arange(n) + 0.0) / (n - 1)
    y, x = np.meshgrid(g, g, indexing="ij")
    out = np.empty((n, n, n), np.float32)
    step = max(1, chunk_points // (n * n))
    for z0 in range(0, n, step):
        zs = g[z0:z0 + step]
        p = np.stack([np.tile(x.ravel(), len(zs)), np.tile(y.ravel(), len(zs)), np.repeat(zs, n * n)], -1)
        d = np.maximum(0.0, scale * (fbm(p, seed, octaves=octaves, freq=freq) + bias))
        out[z0:z0 + len(zs)] = d.reshape(len(zs), n, n)
    return out


def terrain(n=316, size=4.0, height=0.6, seed=13, octaves=5):
    """Seeded height field: an (n+1)^2 vertex grid over a size x size square
    in xz centred at the origin, heights height * fBm; 2 n^2 triangles
    (199,712 at n = 316), smooth normals."""
    g = (np.arange(n + 1) / n - 0.5) * size
    z, x = np.meshgrid(g, g, indexing="ij")
    q = np.stack([x.ravel() / size * 3.0, np.zeros(x.size), z.ravel() / size * 3.0], -1)
    y = height * fbm(q + 0.37, seed, octaves=octaves, freq=1.0)
    pos = np.stack([x.ravel(), y, z.ravel()], -1)
    i = np.arange(n)
    a = (i[:, None] * (n + 1) + i[None, :]).ravel()
    b, c, d = a + 1, a + n + 1, a + n + 2
    idx = np.concatenate([np.stack([a, c, b], 1), np.stack([b, c, d], 1)]).astype(np.uint32)
    return pos, idx


def torus_knot(p=2, q=3, segments=500, sides=50, radius=1.0, tube=0.15):
    """(p, q) torus knot swept by a circular tube: segments * sides * 2
    triangles (50,000 by default), smooth normals from the sweep frame."""
    t = np.arange(segments) / segments * 2 * np.pi

    def curve(t):
        r = radius * (2 + np.cos(q * t)) / 3
        return np.stack([r * np.cos(p * t), radius * np.sin(q * t) / 3, r * np.sin(p * t)], -1)

    c = curve(t)
    tg = curve(t + 1e-4) - curve(t - 1e-4)
    tg /= np.linalg.norm(tg, axis=1, keepdims=True)
    nb = np.cross(tg, c)
    nb /= np.linalg.norm(nb, axis=1, keepdims=True)
    nn = np.cross(nb, tg)
    a = np.arange(sides) / sides * 2 * np.pi
    nrm = np.cos(a)[None, :, None] * nn[:, None, :] + np.sin(a)[None, :, None] * nb[:, None, :]
    pos = c[:, None, :] + tube * nrm
    s_, k_ = np.meshgrid(np.arange(segments), np.arange(sides), indexing="ij")
    v00 = s_ * sides + k_
    v10 = ((s_ + 1) % segments) * sides + k_
    v01 = s_ * sides + (k_ + 1) % sides
    v11 = ((s_ + 1) % segments) * sides + (k_ + 1) % sides
    idx = np.concatenate([np.stack([v00.ravel(), v01.ravel(), v10.ravel()], 1),
                          np.stack([v10.ravel(), v01.ravel(), v11.ravel()], 1)]).astype(np.uint32)
    return pos.reshape(-1, 3), nrm.reshape(-1, 3), idx


def _proto(pos, idx, nrm=None, mat=0):
    pos = np.asarray(pos, np.float64)
    nrm = smooth_normals(pos, idx) if nrm is None else np.asarray(nrm, np.float64)
    nrm = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    return {"pos": pos.astype(np.float32), "nrm": nrm.astype(np.float32), "tan": _tangent_of(nrm).astype(np.float32),
            "idx": np.asarray(idx, np.uint32), "mat": mat}


def random_material(rng):
    """One of Lambert (grey albedo U[0.1, 0.9]), GGX conductor (alpha U[0.05,
    0.5]), GGX dielectric (IOR 1.5, alpha 0.1) and Oren-Nayar (sigma 0.3),
    drawn uniformly."""
    k = rng.randint(4)
    if k == 0:
        a = float(rng.uniform(0.1, 0.9))
        return {"kind": "lambert", "rgb": (a, a, a)}
    if k == 1:
        return {"kind": "ggx_conductor", "eta": 0.2, "k": 3.9, "alpha": float(rng.uniform(0.05, 0.5))}
    if k == 2:
        return {"kind": "ggx_dielectric", "eta_ext": 1.0, "eta_int": 1.5, "alpha": 0.1}
    return {"kind": "oren_nayar", "rgb": (0.6, 0.5, 0.4), "sigma": 0.3}


def instanced_field(width=1920, height=1080, n_instances=256, seed=21, subdiv=8, terrain_n=316,
                    knot=(500, 50), env_size=(2048, 1024), sun_radiance=1e4, materials_per_instance=True):
    """Instances under an image-based sky: n_instances seeded random affine
    instances of four prototypes, a displaced icosphere (subdivision subdiv,
    1,310,720 triangles at 8), a height-field terrain (199,712 triangles at
    316), a torus knot (50,000 triangles at 500 x 50) and a 12-triangle box,
    each instance with a material drawn by random_material, lit by
    sky_environment at env_size with a sun of sun_radiance. Without
    materials_per_instance every (prototype, material) pair is its own
    prototype, the form a scene graph without per-instance materials takes."""
    rng = np.random.RandomState(seed)
    sph_p, sph_f = displaced_icosphere(subdiv, 1.0, (0.0, 0.0, 0.0), seed=seed)
    ter_p, ter_f = terrain(terrain_n, seed=seed + 1)
    kp, kn, kf = torus_knot(segments=knot[0], sides=knot[1])
    bp, bn, bf = _box((0.0, 0.0, 0.0), (1.6, 1.6, 1.6), 0.0)
    shapes = [(sph_p, sph_f, None), (ter_p, ter_f, None), (kp, kf, kn), (bp, bf, bn)]
    materials, insts = [], []
    for i in range(n_instances):
        m = random_affine(rng, np.array([-12.0, 0.0, 4.0]), np.array([12.0, 3.0, 30.0]), scale_lo=0.5,
                          scale_hi=1.5)
        materials.append(random_material(rng))
        insts.append({"proto": int(rng.randint(4)), "l2w": m, "mat": i})
    if materials_per_instance:
        protos = [_proto(p, f, n, 0) for p, f, n in shapes]
    else:
        protos, key = [], {}
        for inst in insts:
            k = (inst["proto"], inst["mat"])
            if k not in key:
                key[k] = len(protos)
                p, f, n = shapes[inst["proto"]]
                protos.append(_proto(p, f, n, inst["mat"]))
            inst["proto"] = key[k]
            inst.pop("mat")
    l2w = np.eye(4)
    t = math.radians(-14.0)
    l2w[:3, :3] = np.array([[1, 0, 0], [0, math.cos(t), -math.sin(t)], [0, math.sin(t), math.cos(t)]])
    l2w[:3, 3] = np.array([0.0, 6.0, -8.0])
    camera = {"l2w": l2w, "aspect": width / height, "fov_y": 0.9, "lens_radius": 0.0, "img_dist": 1.0,
              "obj_dist": 1.0, "sensitivity": 1.0, "width": width, "height": height}
    env = sky_environment(env_size[0], env_size[1], sun_radiance=sun_radiance)
    return {"materials": materials, "meshes": [], "prototypes": protos, "instances": insts, "camera": camera,
            "environment": {"rgb": env, "scale": 1.0}}


def cornell_medium(width=64, height=64, n=32, sigma_t=4.0, albedo=0.9, g=0.7, seed=5, octaves=5):
    """The Cornell box walls and light enclosing a heterogeneous medium on an
    n^3 lattice (seeded fBm densities times sigma_t, albedo, Henyey-Greenstein
    g) that fills most of the box."""
    desc = cornell_box(width, height)
    desc["meshes"] = desc["meshes"][:6]          # walls and light
    lo = np.array([30.0, 10.0, 30.0]) * SCALE
    hi = np.array([520.0, 500.0, 520.0]) * SCALE
    desc["medium"] = {"bounds": np.concatenate([lo, hi]).astype(np.float32),
                      "density": fbm_density(n, seed=seed, octaves=octaves),
                      "sigma_s": albedo * sigma_t, "sigma_e": sigma_t, "g": g}
    return desc


def absorbing_slab(width=16, height=16, sigma=1.5, box=((-1.0, -1.0, 1.0), (1.0, 1.0, 3.0))):
    """A uniform white environment seen through a box of constant density
    that absorbs (sigma_s = 0, sigma_e = sigma): each camera ray keeps
    exp(-sigma * length inside the box) of the sky (see pixel_transmittance)."""
    desc = env_furnace(width, height, quad_in_view=False)
    lo, hi = np.asarray(box[0], np.float32), np.asarray(box[1], np.float32)
    desc["medium"] = {"bounds": np.concatenate([lo, hi]).astype(np.float32),
                      "density": np.ones((4, 4, 4), np.float32), "sigma_s": 0.0, "sigma_e": sigma, "g": 0.0}
    return desc


def pixel_transmittance(desc, width, height, n=32):
    """pixel_cos4's per-pixel weight with each sub-pixel ray attenuated by
    exp(-sigma_e * density * chord length) through the medium box (constant
    density): the film of absorbing_slab under a unit sky."""
    cam = desc["camera"]
    med = desc["medium"]
    sig = med["sigma_e"] * float(np.asarray(med["density"]).max())
    lo, hi = np.asarray(med["bounds"][:3], np.float64), np.asarray(med["bounds"][3:], np.float64)
    opH = 2.0 * cam["obj_dist"] * math.tan(cam["fov_y"] * 0.5)
    opW = opH * cam["aspect"]
    area = opW * opH * (cam["img_dist"] / cam["obj_dist"]) ** 2
    l2w = np.asarray(cam["l2w"], np.float64).reshape(4, 4)
    g = (np.arange(n) + 0.5) / n
    out = np.zeros((height, width))
    for py in range(height):
        for px in range(width):
            ux = (px + g[:, None] + 0 * g[None, :]) / width
            uy = (py + g[None, :] + 0 * g[:, None]) / height
            x = opW * (0.5 - ux)
            y = opH * (0.5 - uy)
            z = np.full_like(x, cam["obj_dist"])
            dl = np.stack([x, y, z], -1).reshape(-1, 3)
            cos4 = (dl[:, 2] ** 2 / (dl ** 2).sum(1)) ** 2
            d = dl @ l2w[:3, :3].T
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            o = l2w[:3, 3]
            with np.errstate(divide="ignore", invalid="ignore"):
                t0 = (lo - o) / d
                t1 = (hi - o) / d
            tn = np.maximum(np.minimum(t0, t1).max(1), 0.0)
            tf = np.maximum(t0, t1).min(1)
            L = np.maximum(tf - tn, 0.0)
            out[py, px] = np.mean(cos4 * np.exp(-sig * L)) * area / cam["img_dist"] ** 2
    return out


def sky_environment(width=256, height=128, sun_dir=(0.4, 0.8, 0.45), sun_radiance=1e4, sun_radius_deg=1.5,
                    zenith=(0.25, 0.45, 1.0), horizon=(1.0, 0.95, 0.85), ground=(0.2, 0.18, 0.15)):
    """Equirectangular sky: a zenith-to-horizon gradient, a darker ground and a
    small sun disc of the given relative radiance. Rows run from the +y pole
    (theta = 0) down; phi = 0 looks along +z (Vector3D::toPolarYUp)."""
    theta = (np.arange(height) + 0.5) / height * np.pi
    phi = (np.arange(width) + 0.5) / width * 2 * np.pi
    th, ph = np.meshgrid(theta, phi, indexing="ij")
    d = np.stack([-np.sin(ph) * np.sin(th), np.cos(th), np.cos(ph) * np.sin(th)], -1)
    up = np.clip(d[..., 1], 0, 1)[..., None]
    sky = np.array(horizon) * (1 - up) + np.array(zenith) * up
    img = np.where(d[..., 1:2] >= 0, sky, np.array(ground))
    s = np.asarray(sun_dir, np.float64)
    s /= np.linalg.norm(s)
    cosang = d @ s
    img = img + (cosang > math.cos(math.radians(sun_radius_deg)))[..., None] * sun_radiance
    return img.astype(np.float32)


def constant_environment(value=1.0, width=32, height=16):
    return np.full((height, width, 3), value, np.float32)


def env_furnace(width=32, height=32, albedo=0.5, env_value=1.0, quad_in_view=True):
    """A large Lambert quad of constant albedo under a uniform environment.
    In view (quad_in_view) it fills the image and, with nothing shadowing its
    hemisphere, reflects albedo times the environment's radiance; with the
    quad behind the camera the image is the environment itself. The two
    films therefore differ by the factor albedo."""
    materials = [{"kind": "lambert", "const": albedo}]
    z = 500.0 if quad_in_view else -500.0
    n = (0, 0, -1) if quad_in_view else (0, 0, 1)
    S = 5000.0
    if quad_in_view:
        pos, nrm, idx = _quad((-S, -S, z), (-S, S, z), (S, S, z), (S, -S, z), n)
    else:
        pos, nrm, idx = _quad((-S, -S, z), (S, -S, z), (S, S, z), (-S, S, z), n)
    meshes = [_mesh(pos, nrm, idx, 0)]
    l2w = np.eye(4)
    camera = {"l2w": l2w, "aspect": width / height, "fov_y": 1.0, "lens_radius": 0.0, "img_dist": 1.0,
              "obj_dist": 1.0, "sensitivity": 1.0, "width": width, "height": height}
    return {"materials": materials, "meshes": meshes, "camera": camera,
            "environment": {"rgb": constant_environment(env_value), "scale": 1.0}}


def sky_scene(width=64, height=48, env_width=64, env_height=32, sun_radiance=50.0, sun_radius_deg=6.0):
    """Open scene under the equirectangular sky of sky_environment: a ground
    quad, a rotated box and a GGX conductor box, camera looking slightly
    down. Every material is one the reference test harness also builds."""
    materials = [{"kind": "lambert", "rgb": (0.6, 0.6, 0.55)},
                 {"kind": "lambert", "rgb": (0.63, 0.06, 0.04)},
                 {"kind": "ggx_conductor", "eta": 0.2, "k": 3.9, "alpha": 0.25}]
    G = 2000.0
    pos, nrm, idx = _quad((-G, 0, -G), (-G, 0, G), (G, 0, G), (G, 0, -G), (0, 1, 0))
    meshes = [_mesh(pos, nrm, idx, 0)]
    for (center, size, rot), m in ((((-120.0, 80.0, 300.0), (160, 160, 160), 25.0), 1),
                                   (((150.0, 120.0, 420.0), (140, 240, 140), -15.0), 2)):
        pos, nrm, idx = _box(center, size, rot)
        meshes.append(_mesh(pos, nrm, idx, m))
    l2w = np.eye(4)
    t = math.radians(-12.0)   # pitched down about x
    l2w[:3, :3] = np.array([[1, 0, 0], [0, math.cos(t), -math.sin(t)], [0, math.sin(t), math.cos(t)]])
    l2w[:3, 3] = np.array([0.0, 220.0, -350.0]) * SCALE
    camera = {"l2w": l2w, "aspect": width / height, "fov_y": 0.9, "lens_radius": 0.0, "img_dist": 1.0,
              "obj_dist": 1.0, "sensitivity": 1.0, "width": width, "height": height}
    env = sky_environment(env_width, env_height, sun_radiance=sun_radiance, sun_radius_deg=sun_radius_deg)
    return {"materials": materials, "meshes": meshes, "camera": camera,
            "environment": {"rgb": env, "scale": 1.0}}


def camera_pose(desc, yaw_deg=0.0, dx=0.0, dy=0.0, width=None, height=None):
    """desc's camera turned by yaw_deg about the vertical axis through the
    scene's centre and shifted by (dx, dy) in its own plane; aspect from
    width/height when given."""
    cam = dict(desc["camera"])
    pts = [np.asarray(m["pos"], np.float64) for m in desc["meshes"]]
    pts += [np.asarray(i["l2w"], np.float64).reshape(-1, 4)[:3, 3][None] for i in desc.get("instances", [])]
    allp = np.concatenate(pts)
    centre = 0.5 * (allp.min(0) + allp.max(0))
    l2w = np.array(cam["l2w"], dtype=np.float64).reshape(4, 4).copy()
    t = math.radians(yaw_deg)
    r = np.array([[math.cos(t), 0.0, math.sin(t)], [0.0, 1.0, 0.0], [-math.sin(t), 0.0, math.cos(t)]])
    pos = l2w[:3, 3] + l2w[:3, 0] * dx + l2w[:3, 1] * dy
    l2w[:3, :3] = r @ l2w[:3, :3]
    l2w[:3, 3] = centre + r @ (pos - centre)
    cam["l2w"] = l2w
    if width and height:
        cam["aspect"] = width / height
    return cam


def random_requests(desc, n, seed=3, sizes=((256, 256), (512, 512), (1024, 1024), (1920, 1080)),
                    spps=(4, 16, 64)):
    """Seeded serving workload: per request a camera pose, a resolution and
    a sample count drawn uniformly from the lists."""
    rng = np.random.RandomState(seed)
    out = []
    for i in range(n):
        w, h = sizes[rng.randint(len(sizes))]
        spp = int(spps[rng.randint(len(spps))])
        cam = camera_pose(desc, yaw_deg=rng.uniform(-12, 12), dx=rng.uniform(-0.3, 0.3), dy=rng.uniform(-0.2, 0.2),
                          width=w, height=h)
        out.append({"width": w, "height": h, "spp": spp, "camera": cam, "seed": int(rng.randint(1, 2 ** 31))})
    return out


def random_rays(desc, n, seed=1, mode="mixed"):
    """Seeded ray batch (n, 8) {org, dir, tmin, tmax} through the scene
    bounds: origins inside the box, directions uniform on the sphere."""
    rng = np.random.RandomState(seed)
    allp = np.concatenate([m["pos"] for m in desc["meshes"]]).astype(np.float64)
    lo, hi = allp.min(0), allp.max(0)
    org = lo + (hi - lo) * rng.uniform(0.02, 0.98, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    if mode == "axis":
        # a share of exactly axis-aligned and signed-zero directions
        k = n // 4
        ax = rng.randint(0, 3, size=k)
        d[:k] = 0.0
        d[np.arange(k), ax] = rng.choice([-1.0, 1.0], size=k)
        d[:k][d[:k] == 0.0] = -0.0 * (rng.uniform(size=(d[:k] == 0.0).sum()) < 0.5)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = org
    r[:, 3:6] = d
    r[:, 6] = 0.0
    r[:, 7] = np.inf
    return r


def to_scene_text(desc, width, height, spp, method="PT", brightness=1.0):
    """The scene of `desc` written in SLR's scene language (the subset
    slrh_scene_load_text reads; libSLRSceneGraph/API.cpp builtins): sRGB
    spectra, matte and emitter materials, one createMesh per mesh with its
    vertices in the reference scene's units under setTransform(root,
    scale(0.01)) as TestScenes/Original_Cornell_Box.txt does, the camera
    under its own node, renderer and settings. Lambert and emitter materials
    with RGB values only (the Cornell boxes)."""
    out = ['setRenderer("method": "%s", ("samples": %d,));' % (method, spp),
           'setRenderSettings("width": %d, "height": %d, "brightness": %r);' % (width, height, brightness)]
    for i, m in enumerate(desc["materials"]):
        if m["kind"] == "lambert" and "rgb" in m:
            out.append('mat%d = createSurfaceMaterial("matte", (SpectrumTexture(Spectrum("Reflectance", "sRGB", %r, %r, '
                       '%r)),));' % ((i,) + tuple(float(c) for c in m["rgb"])))
        elif m["kind"] == "emitter" and "rgb" in m:
            out.append('mat%d = createSurfaceMaterial("emitter", (mat%d, createEmitterSurfaceProperty("diffuse", '
                       '(SpectrumTexture(Spectrum("LightSource", "sRGB", %r, %r, %r)),))));'
                       % ((i, m["base"]) + tuple(float(c) for c in m["rgb"])))
        else:
            raise ValueError("to_scene_text: material %r" % (m,))
    for mesh in desc["meshes"]:
        pos = np.asarray(mesh["pos"], np.float64) / SCALE
        nrm = np.asarray(mesh["nrm"], np.float64)
        tan = np.asarray(mesh["tan"], np.float64)
        idx = np.asarray(mesh["idx"], np.int64)
        mat = mesh["mat"]
        verts = ",\n".join("((%.9g, %.9g, %.9g), (%.9g, %.9g, %.9g), (%.9g, %.9g, %.9g), (0, 0))"
                           % (tuple(p) + tuple(n) + tuple(t)) for p, n, t in zip(pos, nrm, tan))
        mats = np.full(len(idx), mat) if np.isscalar(mat) else np.asarray(mat)
        groups = []
        for mid in sorted(set(int(x) for x in mats)):
            tris = ", ".join("(%d, %d, %d)" % tuple(t) for t in idx[mats == mid])
            groups.append("(mat%d, (%s,))" % (mid, tris))
        out.append("addChild(root, createMesh((%s), (%s,)));" % (verts, ", ".join(groups)))
    cam = desc["camera"]
    l2w = np.asarray(cam["l2w"], np.float64)
    if not np.allclose(l2w[:3, :3], np.eye(3)):
        raise ValueError("to_scene_text: camera rotation")
    t = l2w[:3, 3] / SCALE
    out.append('cameraNode = createNode();')
    out.append('addChild(cameraNode, createPerspectiveCamera("aspect": %r, "fovY": %r, "radius": %r, "imgDist": %r, '
               '"objDist": %r));' % (float(cam["aspect"]), float(cam["fov_y"]), float(cam.get("lens_radius", 0.0)),
                                     float(cam.get("img_dist", 1.0)), float(cam.get("obj_dist", 1.0))))
    out.append('setTransform(cameraNode, translate(%.9g, %.9g, %.9g));' % tuple(t))
    out.append('addChild(root, cameraNode);')
    out.append('setTransform(root, scale(%r));' % SCALE)
    return "\n".join(out) + "\n"
