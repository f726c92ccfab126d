"""GPU refit of a resident scene's mesh (rq_scene_refit_mesh) against the CPU path it replaces:
near_field::refit_bvh (rq_bvh_refit, whose restatement the reference's own refit test pins),
the instance boxes a fresh pack of the new positions computes, and instance_bvh::place over
the previous words. The whole packed buffer after a GPU refit must equal that CPU result word
for word, and the PLAIN device image refreshed in place must equal a host relayout of those
words; traversal over them (REF and PLAIN refit on the device, CW8 uploaded again) must
equal the oracle's bit for bit, and the oracle must agree with the f64 brute force on hit or
miss over the moved geometry (the refit boxes enclose what they must)."""
import numpy as np
import pytest

import oracle
from helpers import compare_exact, gpu_trace, to_device
from renderlab_amd import rq

pytestmark = pytest.mark.gpu

HEADER, STRIDE = 16, 32


def grid(n):
    pos = np.array([[x, y, 0.0] for y in range(n + 1) for x in range(n + 1)], np.float32) / n
    idx = []
    for y in range(n):
        for x in range(n):
            i, w = y * (n + 1) + x, n + 1
            idx += [i, i + 1, i + w + 1, i, i + w + 1, i + w]
    return pos, np.array(idx, np.uint32)


def model(g):
    q = g.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r * g.uniform(1.0, 3.0)
    m[:3, 3] = g.uniform(-6, 6, 3)
    return m.astype(np.float32)


def pose(pos, k, g):
    """A skinned-looking pose: a travelling wave, a twist and some jitter, plus a shift."""
    p = pos.astype(np.float64).copy()
    p[:, 2] += 0.15 * np.sin(6.0 * p[:, 0] + k) * np.cos(4.0 * p[:, 1])
    a = 0.3 * k * p[:, 1]
    x = p[:, 0] * np.cos(a) - p[:, 2] * np.sin(a)
    z = p[:, 0] * np.sin(a) + p[:, 2] * np.cos(a)
    p[:, 0], p[:, 2] = x, z
    p += g.normal(scale=0.004, size=p.shape)
    p += [0.2 * k, -0.1 * k, 0.05 * k]
    return p.astype(np.float32)


@pytest.fixture(scope="module")
def setup():
    print("workload: refit scene (grid 48x48 x 40 instances + soup x 24)", flush=True)
    g = np.random.default_rng(5)
    gpos, gidx = grid(48)
    soup = g.normal(size=(900, 3)).astype(np.float32)
    meshes = [(gpos, gidx, 1), (soup, np.arange(900, dtype=np.uint32), 0)]
    inst = [dict(model=model(g), mesh=0 if k < 40 else 1, flags=int(k % 7 == 3)) for k in range(64)]
    return g, meshes, inst


def expected(old, meshes, inst, mesh, orig, new):
    """The CPU refit and repack of `old` (packed words) with mesh `mesh` moved to `new`."""
    w = old.copy()
    k = next(i for i, x in enumerate(inst) if x["mesh"] == mesh)
    r = HEADER + k * STRIDE
    nodes, leaves, positions, indices = (int(v) for v in old[r + 18:r + 22])
    bvh = rq.build_bvh_sah(orig, meshes[mesh][1])
    rq.refit_bvh(bvh, new)  # same tree as the pack's: refit of the built topology
    w[nodes:leaves] = bvh.records()["nodes"].reshape(-1)
    w[positions:indices] = new.reshape(-1).view(np.uint32)
    moved = list(meshes)
    moved[mesh] = (new, meshes[mesh][1], meshes[mesh][2])
    fresh = rq.Scene.build(moved, inst).words  # instance boxes depend on the mesh bounds only
    n = len(inst)
    for i in range(n):
        q = HEADER + i * STRIDE
        w[q + 12:q + 18] = fresh[q + 12:q + 18]
    base, reserve = HEADER + n * STRIDE, rq.instance_bvh.reserve(n)
    w[:base + reserve] = rq.instance_bvh.place(w[:base], HEADER, STRIDE, n, previous=old[:base + reserve])
    return w


def rays_at(words, inst, g, n=6000):
    lo = np.array([words[HEADER + i * STRIDE + 12:HEADER + i * STRIDE + 15].view(np.float32) for i in range(len(inst))])
    hi = np.array([words[HEADER + i * STRIDE + 15:HEADER + i * STRIDE + 18].view(np.float32) for i in range(len(inst))])
    pick = g.integers(len(inst), size=n)
    target = lo[pick] + (hi[pick] - lo[pick]) * g.random((n, 3))
    r = np.zeros(n, oracle.RAY_DTYPE)
    r["origin"] = g.uniform(-14, 14, (n, 3))
    r["direction"] = target - r["origin"]
    r["direction"][::9] = [0, 0, -1]
    r["tmax"] = 1e20
    r["tmin"][1::11] = 0.5
    return r


def check_traversal(scene, words, rays, layouts):
    o, _, _ = oracle.traverse(words, rays, 0)
    hit = (o["flags"] & 1) == 1
    assert hit.sum() > len(rays) // 5
    b = oracle.brute_f64(words, rays[:1500], 0)
    assert np.array_equal(hit[:1500], b["hit"] == 1), "a refit box stopped enclosing its geometry"
    _, ov, _ = oracle.traverse(words, rays, 1, want_hits=False, want_vis=True)
    for layout in layouts:
        h, _ = gpu_trace(scene, layout, rays, rq.CLOSEST)
        assert len(compare_exact(h, o)) == 0, layout
        _, v = gpu_trace(scene, layout, rays, rq.ANY)
        assert np.array_equal(v, ov), layout


def test_refit_equals_cpu_refit_and_repack(setup):
    g, meshes, inst = setup
    scene = rq.Scene.build(meshes, inst)
    scene.upload(("ref",))
    orig = meshes[0][0]
    old = scene.words
    r = HEADER  # instance 0 draws mesh 0
    nodes, leaves = int(old[r + 18]), int(old[r + 19])
    packed = rq.build_bvh_sah(orig, meshes[0][1]).records()["nodes"].reshape(-1)
    assert np.array_equal(packed, old[nodes:leaves]), "packed BLAS differs from rq_bvh_build"
    for k in (1, 2):  # two poses in a row: the cached plan, the top level's refit count
        new = pose(orig, k, g)
        print(f"case: refit pose {k}", flush=True)
        scene.refit_mesh(0, to_device(new), len(new))
        want = expected(old, meshes, inst, 0, orig, new)
        got = scene.words
        assert got.size == want.size
        diff = np.nonzero(got != want)[0]
        assert diff.size == 0, f"{diff.size} words differ, first at {diff[:8]}"
        assert got[3] == k  # top level refit, not rebuilt
        check_traversal(scene, got, rays_at(got, inst, g), ("ref",))
        old = got


def test_refit_keeps_plain_in_step(setup):
    g, meshes, inst = setup
    scene = rq.Scene.build(meshes, inst)
    scene.upload(("ref", "plain", "cw8"))
    for k in (3, 4):
        new = pose(meshes[0][0], k, g)
        print(f"case: refit pose {k} with PLAIN resident", flush=True)
        scene.refit_mesh(0, to_device(new), len(new))
        assert scene.device_bytes("plain") > 0 and scene.device_bytes("cw8") == 0
        # the PLAIN device image refreshed in place equals a host relayout of the refit words
        for part in (0, 1, 2, 3, 4, 6):
            host = scene.layout_image("plain", part)
            dev = scene.download_part("plain", part)
            assert np.array_equal(dev[:host.size], host), f"PLAIN part {part}"
        words = scene.words
        check_traversal(scene, words, rays_at(words, inst, g), ("ref", "plain"))
    scene.upload(("cw8",))
    check_traversal(scene, words, rays_at(words, inst, g), ("cw8",))


def test_refit_rejects_bad_input(setup):
    g, meshes, inst = setup
    scene = rq.Scene.build(meshes, inst)
    before = scene.words
    with pytest.raises(rq.InvalidInput):  # REF not uploaded
        scene.refit_mesh(0, to_device(meshes[0][0]), len(meshes[0][0]))
    scene.upload(("ref",))
    bad = meshes[0][0].copy()
    bad[17, 1] = np.nan
    print("case: refit non-finite / wrong count / no such mesh", flush=True)
    with pytest.raises(rq.InvalidInput):
        scene.refit_mesh(0, to_device(bad), len(bad))
    with pytest.raises(rq.InvalidInput):
        scene.refit_mesh(0, to_device(meshes[0][0][:-1]), len(meshes[0][0]) - 1)
    with pytest.raises(rq.InvalidInput):
        scene.refit_mesh(2, to_device(meshes[0][0]), len(meshes[0][0]))
    assert np.array_equal(scene.words, before)
    # the device words are unchanged too: traversal still matches the oracle on the old words
    check_traversal(scene, before, rays_at(before, inst, g, 2000), ("ref",))
    words = rq.Scene.from_gi_words(before)
    words.upload(("ref",))
    with pytest.raises(rq.InvalidInput):  # adopted words carry no instance transforms
        words.refit_mesh(0, to_device(meshes[0][0]), len(meshes[0][0]))
