"""Device BVH: structure and exact boxes after builds and refits.

Checks, on the node array the device traverses:
* every internal node is reached exactly once from the root and every
  primitive appears as exactly one leaf;
* a leaf's box equals the primitive's bounds (spheres: position -/+ radius in
  float32, the reference's GeometrySphere.cc:60-73 bounds);
* an internal child's box equals the union (min/max, exact in float) of its
  two children's boxes.
Refitting moved primitives must satisfy the same invariants and render
bit-identically to a scene rebuilt from scratch: the closest hit is the
smallest (t, geometry, primitive), whatever the hierarchy.
"""

import numpy
import pytest

import fresnel_amd as fresnel

pytestmark = pytest.mark.gpu

LEAF = 0x80000000
EMPTY = 0xffffffff


def check_bvh(nodes, n_prims, sphere_bounds=None):
    """Walk the dumped hierarchy; returns {(slot, prim): box}."""
    n = nodes.shape[0]
    boxes = nodes[:, :12].copy().view(numpy.float32).reshape(n, 2, 6)
    links = nodes[:, 12:16]
    seen = numpy.zeros(n, bool)
    leaves = {}

    def child_union(c):
        return numpy.concatenate([numpy.minimum(boxes[c, 0, :3], boxes[c, 1, :3]),
                                  numpy.maximum(boxes[c, 0, 3:], boxes[c, 1, 3:])])

    stack = [0]
    while stack:
        i = stack.pop()
        assert not seen[i], f"node {i} reached twice"
        seen[i] = True
        for side in range(2):
            c = int(links[i, side])
            if c == EMPTY:
                continue
            if c & LEAF:
                key = (int(links[i, 2 + side]), c & ~LEAF)
                assert key not in leaves, f"primitive {key} appears twice"
                leaves[key] = boxes[i, side]
            else:
                assert 0 < c < n
                assert numpy.array_equal(boxes[i, side], child_union(c)), f"box of node {c} is not its children's union"
                stack.append(c)
    assert seen.all(), "unreachable nodes"
    assert len(leaves) == n_prims
    if sphere_bounds is not None:
        lo, hi = sphere_bounds
        for (slot, prim), b in leaves.items():
            assert numpy.array_equal(b, numpy.concatenate([lo[prim], hi[prim]])), f"leaf {prim} bounds"
    return leaves


def sphere_scene(device, pos, rad):
    s = fresnel.Scene(device, lights=fresnel.light.cloudy())
    g = fresnel.geometry.Sphere(s, position=pos, radius=rad,
                                material=fresnel.material.Material(color=fresnel.color.linear([0.7, 0.3, 0.2]),
                                                                   roughness=0.4))
    s.camera = fresnel.camera.Orthographic.fit(s)
    return s, g


@pytest.mark.parametrize('n', [2, 3, 1000, 100000])
def test_built_bvh_is_exact(n, gpu_device):
    rng = numpy.random.default_rng(n)
    pos = rng.uniform(-20, 20, (n, 3)).astype(numpy.float32)
    rad = rng.uniform(0.1, 0.9, n).astype(numpy.float32)
    s, _ = sphere_scene(gpu_device, pos, rad)
    nodes = s._scene.debugBVH()
    assert nodes.shape == (max(n - 1, 1), 16)
    check_bvh(nodes, n, (pos - rad[:, None], pos + rad[:, None]))


def test_bvh_with_duplicate_positions(gpu_device):
    # equal Morton codes everywhere: the hierarchy falls back to index order
    n = 5000
    pos = numpy.zeros((n, 3), numpy.float32)
    pos[: n // 2] = [1, 2, 3]
    rad = numpy.full(n, 0.5, numpy.float32)
    s, _ = sphere_scene(gpu_device, pos, rad)
    check_bvh(s._scene.debugBVH(), n, (pos - rad[:, None], pos + rad[:, None]))


def test_mixed_geometry_bvh(gpu_device):
    import scenes
    s = scenes.eight_polyhedra(gpu_device)
    fresnel.geometry.Sphere(s, position=[[0, 0, 5], [1, 1, 6]], radius=0.3)
    nodes = s._scene.debugBVH()
    leaves = check_bvh(nodes, s._scene.bvhInfo()[2])
    assert len({slot for slot, _ in leaves}) == 2


def test_refit_matches_rebuild(gpu_device):
    """Refit after a move: exact boxes, and the image equals a fresh build bit for bit."""
    n = 50000
    rng = numpy.random.default_rng(11)
    pos = rng.uniform(-15, 15, (n, 3)).astype(numpy.float32)
    rad = rng.uniform(0.2, 0.6, n).astype(numpy.float32)
    s, g = sphere_scene(gpu_device, pos, rad)
    s.set_bvh_refit(True, limit=100.0)
    t = fresnel.tracer.Path(gpu_device, w=96, h=64)
    t.sample(s, samples=2, light_samples=2)
    b0, r0, _ = s._scene.bvhInfo()
    for frame in range(3):
        pos = (pos + rng.normal(0, 0.3, pos.shape)).astype(numpy.float32)
        g.position[:] = pos
        t.sample(s, samples=2, light_samples=2)
        refit_img = numpy.array(t.linear_output[:])
        check_bvh(s._scene.debugBVH(), n, (pos - rad[:, None], pos + rad[:, None]))
        # same frame on a scene built from scratch, same seed
        s2, _ = sphere_scene(gpu_device, pos, rad)
        s2.camera = s.camera
        t2 = fresnel.tracer.Path(gpu_device, w=96, h=64)
        t2.seed = t.seed - 1
        t2.sample(s2, samples=2, light_samples=2)
        assert numpy.array_equal(refit_img, numpy.array(t2.linear_output[:])), f"frame {frame}"
    b1, r1, _ = s._scene.bvhInfo()
    assert b1 == b0 and r1 == r0 + 3


def test_refit_policy_rebuilds_degraded_hierarchy(gpu_device):
    """A refit that inflates the surface-area cost beyond the limit is replaced by a rebuild."""
    n = 20000
    rng = numpy.random.default_rng(12)
    pos = rng.uniform(-15, 15, (n, 3)).astype(numpy.float32)
    rad = numpy.full(n, 0.3, numpy.float32)
    s, g = sphere_scene(gpu_device, pos, rad)
    s.set_bvh_refit(True, limit=1.5)
    s._scene.debugBVH()  # build
    b0, r0, _ = s._scene.bvhInfo()
    built, cur = s._scene.bvhCost()
    assert built > 0 and cur == built
    # small move: refit kept
    g.position[:] = (pos + rng.normal(0, 0.01, pos.shape)).astype(numpy.float32)
    s._scene.debugBVH()
    b1, r1, _ = s._scene.bvhInfo()
    assert (b1, r1) == (b0, r0 + 1)
    assert s._scene.bvhCost()[1] < 1.5 * built
    # scramble: the refit tree's cost explodes, so the scene is rebuilt
    g.position[:] = rng.permutation(pos)
    nodes = s._scene.debugBVH()
    b2, r2, _ = s._scene.bvhInfo()
    assert b2 == b1 + 1 and r2 == r1 + 1
    new_pos = numpy.array(g.position[:])
    check_bvh(nodes, n, (new_pos - rad[:, None], new_pos + rad[:, None]))
