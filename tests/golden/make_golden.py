"""Writes the golden fixtures of this directory from the compiled reference
(oracle/_ref, built from the reference sources by oracle/ref.mk):

* cornell_hits.npz: 2048 seeded rays through the Cornell box (half through
  the box interior, half axis-aligned) and the reference QBVH's closest hits
  for them (QBVH::intersect over TriangleSurfaceShape::intersect): t, global
  triangle id, b0 = 1 - u - v, b1 = u.
* cornell_16x16_image.npz: the reference PTRenderer's image of the Cornell
  box at 16 x 16: per pixel CIE XYZ, the mean and the standard error over
  32 independent batches of 512 samples per pixel (16384 spp in all).
* ref_<case>.npz for each case of tests/refcases.py: the reference
  renderer's image statistics (PTRenderer, VolumetricPTRenderer for the
  volumetric cases, BPTRenderer for the bpt_ cases) that the 3-sigma parity
  tests compare against.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import refdrv  # noqa: E402
from slr_amd import scenes  # noqa: E402
from tests import refcases  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def hits():
    desc = scenes.cornell_box(16, 16)
    rays = np.concatenate([scenes.random_rays(desc, 1024, seed=21),
                           scenes.random_rays(desc, 1024, seed=22, mode="axis")]).astype(np.float32)
    r = refdrv.RefScene(desc)
    r.build_accel()
    t, prim, b0, b1 = r.intersect(rays)[:4]
    r.close()
    np.savez_compressed(os.path.join(HERE, "cornell_hits.npz"), rays=rays, t=t, prim=prim, b0=b0, b1=b1)


def image(W=16, H=16, B=32, spp=512):
    desc = scenes.cornell_box(W, H)
    x, y, z, integ = refdrv.cmf_strata()
    threads = max(1, min(8, len(os.sched_getaffinity(0))))
    out = []
    for b in range(B):
        r = refdrv.RefScene(desc)
        f, _, _ = r.render(W, H, spp, 10000 + b, threads=threads)
        r.close()
        f = f.astype(np.float64)
        out.append(np.stack([f @ x, f @ y, f @ z], -1) / integ / spp)
    out = np.array(out)
    np.savez_compressed(os.path.join(HERE, "cornell_16x16_image.npz"), mean=out.mean(0),
                        se=out.std(0, ddof=1) / np.sqrt(B), batches=B, spp=spp)


def case(name):
    make, W, H, B, spp, seed, vol = refcases.CASES[name]
    desc = make()
    cmf = refdrv.cmf_strata()
    threads = max(1, min(8, len(os.sched_getaffinity(0))))
    # all batches on one scene build (seeds seed .. seed + B - 1)
    r = refdrv.RefScene(desc)
    films, sec, build_s = r.render_batches(W, H, spp, seed, B, threads=threads, volumetric=vol is True,
                                           bpt=vol == "bpt")
    r.close()
    print("%s: %d batches of %dx%d x %d spp in %.1fs (build %.1fs)" % (name, B, W, H, spp, sec, build_s))
    out = np.array([refcases.xyz(f.astype(np.float64), cmf) / spp for f in films])
    tot = out.sum(axis=(1, 2, 3))
    np.savez_compressed(os.path.join(HERE, "ref_%s.npz" % name), mean=out.mean(0), se=out.std(0, ddof=1) / np.sqrt(B),
                        tot_mean=tot.mean(), tot_se=tot.std(ddof=1) / np.sqrt(B), batches=B, spp=spp)


if __name__ == "__main__":
    if not refdrv.available():
        sys.exit("the compiled reference is missing: make -f oracle/ref.mk")
    which = sys.argv[1:]
    if not which or "hits" in which:
        hits()
    if not which or "image" in which:
        image()
    for name in refcases.CASES:
        if not which or name in which:
            case(name)
