"""Scenes packed from cooked geometry on the GPU: traversal over a mix of a near-field mesh with
masks, a cooked cluster BVH and a triangle list equals the oracle's (REF bit for bit, PLAIN bit for
bit); gi_near_blocked over the packed masks (rq_near_blocked) equals the oracle's restatement on
random queries, with and without masks, both sides, every bin count."""
import numpy as np
import pytest

import oracle
from helpers import compare_exact, gpu_trace, to_device
from renderlab_amd import rq
from test_cooked import grid, rnf_bytes

pytestmark = pytest.mark.gpu


def cooked_scene(bins, g):
    p, i = grid(40)
    p = (p * np.float32(8.0)).astype(np.float32)
    field_bvh = rq.ClusterBvh.from_mesh(p, i, sah=False)
    words = bins * bins // 32
    place = []
    models = []
    for k in range(2):
        m = np.eye(4, dtype=np.float32)
        m[3, :3] = [10.0 * k - 5.0, 1.0 * k, -3.0]
        masks = g.integers(0, 1 << 32, size=len(p) * 2 * words, dtype=np.uint64).astype(np.uint32)
        masks[: len(p) * words] |= np.uint32(0xF0F0F0F0)  # more blocked bins on some front masks
        place.append((0, m, masks))
        models.append(m)
    field = rq.NearField.decode(rnf_bytes((0.5, 1.0, 0.01, bins), [3] * 32, [field_bvh.encode()], place))
    soup = g.normal(size=(600, 3)).astype(np.float32) * 3
    cluster = rq.ClusterBvh.from_mesh(soup, np.arange(600, dtype=np.uint32), sah=True)
    tris = g.normal(size=(900, 3)).astype(np.float32) * 2 + [0, 4, 0]
    meshes = [("field", field, 0), ("bvh", cluster, 0), ("mesh", (tris, np.arange(900, dtype=np.uint32), 1))]
    inst = [dict(model=models[0].T, mesh=0), dict(model=models[1].T, mesh=0),
            dict(model=np.eye(4, dtype=np.float32), mesh=0)]  # same mesh, no cooked bits: no masks
    for k in range(40):
        m = np.eye(4, dtype=np.float32)
        m[:3, 3] = g.uniform(-15, 15, 3)
        inst.append(dict(model=m, mesh=1 + k % 2))
    scene = rq.Scene.build_cooked(meshes, inst)
    return scene, p, i, len(inst)


@pytest.fixture(scope="module", params=[8, 16, 32])
def setup(request):
    g = np.random.default_rng(request.param)
    print(f"workload: cooked scene, bins {request.param}", flush=True)
    scene, p, i, n_inst = cooked_scene(request.param, g)
    scene.upload(("ref", "plain"))
    return request.param, g, scene, p, i, n_inst


def test_cooked_scene_traversal(setup):
    bins, g, scene, p, i, n_inst = setup
    rays = np.zeros(5000, oracle.RAY_DTYPE)
    rays["origin"] = g.uniform(-20, 20, (5000, 3))
    rays["direction"] = g.normal(size=(5000, 3))
    rays["direction"][::7] = [0, -1, 0]
    rays["tmax"] = 1e20
    words = scene.words
    print(f"case: cooked traversal bins {bins}", flush=True)
    o, _, _ = oracle.traverse(words, rays, 0)
    assert ((o["flags"] & 1) == 1).sum() > 500
    for layout in ("ref", "plain"):
        h, _ = gpu_trace(scene, layout, rays, rq.CLOSEST)
        assert len(compare_exact(h, o)) == 0, layout


def test_near_blocked_matches_oracle(setup):
    bins, g, scene, p, i, n_inst = setup
    n = 20000
    q = np.zeros(n, rq.NEAR_QUERY_DTYPE)
    q["inst"] = g.integers(0, 4, n)  # 0, 1 have masks; 2 has none; 3 is a cluster-BVH instance
    q["inst"][::97] = n_inst + 5     # past the instance records
    t = g.integers(0, len(i) // 3, n)
    q["tri"] = i.reshape(-1, 3)[t]
    u, v = g.random(n), g.random(n)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    q["bary"][:, 0], q["bary"][:, 1] = u, v
    q["bary"][::11] = [0.0, 0.0]     # all weight on the first vertex
    q["back"] = g.integers(0, 2, n)
    q["d"] = g.normal(size=(n, 3))
    q["d"][::13] = [0, 0, -1]
    print(f"case: near_blocked bins {bins} n={n}", flush=True)
    dq = to_device(q)
    bits = rq.DeviceBuffer(((n + 31) // 32) * 4)
    rq.near_blocked(scene, dq, n, bits)
    rq.synchronize()
    got = np.unpackbits(bits.download(np.uint32, (n + 31) // 32).view(np.uint8), bitorder="little")[:n].astype(bool)
    want = oracle.near_blocked(scene.words, bins, q)
    assert np.array_equal(got, want)
    assert 0.01 < want.mean() < 0.9 and not want[q["inst"] >= 2].any()
