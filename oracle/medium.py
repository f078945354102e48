"""CPU restatement of the density-grid medium, for tests only (a checker,
never the product path).

* density(med, params): DensityGridMediumDistribution::calcDensity
  (libSLR/MediumDistribution/DensityGridMediumDistribution.cpp:369-395) in
  FP32 with the reference's operation order; pinned bit-for-bit against the
  compiled reference (oracle/_ref, refd_medium_density).
* optical_depth / transmittance: the line integral of sigma_e * density
  along a segment by composite midpoint quadrature in FP64, the quantity
  whose exponential ratio tracking (:577-745) and delta tracking (:402-516)
  estimate and sample. Pinned by closed forms: a constant density gives
  sigma * chord length; a density linear along the ray gives the trapezoid
  value exactly.

`med` is a scene description's "medium" dict: bounds (min xyz, max xyz),
density (nz, ny, nx) float32, sigma_e.
"""
import numpy as np

f32 = np.float32


def local_coords(med, p):
    """BoundingBox3D::calculateLocalCoordinates: (p - min) / (max - min), FP32."""
    b = np.asarray(med["bounds"], f32)
    lo, hi = b[:3], b[3:]
    return (np.asarray(p, f32) - lo) / (hi - lo)


def density(med, params):
    """calcDensity at (n, 3) local coordinates (FP32, reference order)."""
    g = np.asarray(med["density"], f32)
    nz, ny, nx = g.shape
    q = np.asarray(params, f32).reshape(-1, 3)
    out = np.zeros(len(q), f32)
    inside = ~((q < 0).any(1) | (q > 1).any(1))
    x, y, z = q[inside, 0], q[inside, 1], q[inside, 2]
    lx = np.minimum((x * f32(nx - 1)).astype(np.uint32), nx - 1)
    ly = np.minimum((y * f32(ny - 1)).astype(np.uint32), ny - 1)
    lz = np.minimum((z * f32(nz - 1)).astype(np.uint32), nz - 1)
    ux, uy, uz = np.minimum(lx + 1, nx - 1), np.minimum(ly + 1, ny - 1), np.minimum(lz + 1, nz - 1)
    wux = x * f32(nx - 1) - lx.astype(f32)
    wuy = y * f32(ny - 1) - ly.astype(f32)
    wuz = z * f32(nz - 1) - lz.astype(f32)
    wlx, wly, wlz = f32(1) - wux, f32(1) - wuy, f32(1) - wuz
    terms = [g[lz, ly, lx] * wlz * wly * wlx, g[lz, ly, ux] * wlz * wly * wux,
             g[lz, uy, lx] * wlz * wuy * wlx, g[lz, uy, ux] * wlz * wuy * wux,
             g[uz, ly, lx] * wuz * wly * wlx, g[uz, ly, ux] * wuz * wly * wux,
             g[uz, uy, lx] * wuz * wuy * wlx, g[uz, uy, ux] * wuz * wuy * wux]
    acc = terms[0]
    for t in terms[1:]:
        acc = acc + t
    out[inside] = acc
    return out


def optical_depth(med, o, d, t0, t1, steps=4096):
    """sigma_e times the integral of the density over [t0, t1] (midpoint rule)."""
    o = np.asarray(o, np.float64)
    d = np.asarray(d, np.float64)
    if not np.isfinite(t1):
        raise ValueError("finite segment required")
    ts = t0 + (np.arange(steps) + 0.5) * (t1 - t0) / steps
    pts = o[None, :] + ts[:, None] * d[None, :]
    dens = density(med, local_coords(med, pts)).astype(np.float64)
    return float(med["sigma_e"]) * dens.sum() * (t1 - t0) / steps


def transmittance(med, o, d, t0, t1, steps=4096):
    return float(np.exp(-optical_depth(med, o, d, t0, t1, steps)))
