"""The SBVH build's device half (sbvh_gpu.hip): with a device context the
nodes of at least 2^16 fragments at the top of a triangle tree are binned and
partitioned on the device. The trees must equal the host builder's (which
equals the reference's, tests/test_host.py) node for node: boxes, children,
leaf references and their order, SBVH and QBVH, compared bit for bit
against the same scene built in a host-only context."""
import numpy as np
import pytest

from slr_amd import build_scene, scenes

pytestmark = pytest.mark.gpu


def _terrain_scene():
    desc = scenes.cornell_box(16, 16)
    p, f = scenes.terrain(n=316, size=4.0, height=0.6, seed=13)
    n = scenes.smooth_normals(p, f)
    desc["meshes"].append({"pos": p.astype(np.float32), "nrm": n.astype(np.float32),
                           "tan": scenes._tangent_of(n).astype(np.float32), "idx": f, "mat": 0})
    return desc


def _origin_sphere_scene():
    # a displaced sphere about the origin: coordinates of both signs and
    # exact zeros, as the prototypes of the instanced configuration have
    desc = scenes.cornell_box(16, 16)
    p, f = scenes.displaced_icosphere(7, 1.0, (0.0, 0.0, 0.0), seed=3)
    n = scenes.smooth_normals(p, f)
    desc["meshes"].append({"pos": p.astype(np.float32), "nrm": n.astype(np.float32),
                           "tan": scenes._tangent_of(n).astype(np.float32), "idx": f, "mat": 0})
    return desc


SCENES = {
    "origin_sphere": _origin_sphere_scene,
    "c2": lambda: scenes.cornell_icosphere(64, 36, subdiv=8),
    "sphere7": lambda: scenes.cornell_sphere(16, 16, subdiv=7),
    "terrain": _terrain_scene,
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_device_binned_tree_equals_host_tree(gpu_ctx, host_ctx, name):
    desc = SCENES[name]()
    dev = build_scene(gpu_ctx, desc)
    host = build_scene(host_ctx, desc)
    bd, bh = dev.build_stats(), host.build_stats()
    assert bd["device_nodes"] > 0 and bh["device_nodes"] == 0
    for dump in ("sbvh_dump", "qbvh_dump"):
        for a, b in zip(getattr(dev, dump)(), getattr(host, dump)()):
            assert a.shape == b.shape
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), dump
    sd, sh = dev.accel_stats(), host.accel_stats()
    for k in ("sbvh_nodes", "sbvh_refs", "sbvh_depth", "sbvh_cost", "qbvh_nodes", "qbvh_depth", "qbvh_cost"):
        assert sd[k] == sh[k], k
    print("build %s: %d device nodes, %.3f s with the device, %.3f s host only" % (
        name, bd["device_nodes"], bd["build_seconds"], bh["build_seconds"]))
