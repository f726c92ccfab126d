"""The reflection rays' geometry and traversal (rl_geometry, rs_trace): the packer
(src/reflection_software.rs) and the oracle's rs_trace restatement
(shaders/reflections_software_bvh4.wgsl), pinned by the reference's own tests
(reflection_software.rs:476-545, restated), the f64 brute force, the tie rules
and the input checks."""
import numpy as np
import pytest

import oracle
from renderlab_amd import rq

RS_HEADER, RS_REC, RS_TRI = 8, 24, 12


def grid(n=23):
    pos = np.array([[x, y, ((x * 7 + y * 3) % 5) * 0.01] for y in range(n + 1) for x in range(n + 1)], np.float32)
    idx = []
    for y in range(n):
        for x in range(n):
            i, w = y * (n + 1) + x, n + 1
            idx += [i, i + 1, i + w + 1, i, i + w + 1, i + w]
    return pos, np.array(idx, np.uint32)


def rays_of(o, d, tmin=0.0, tmax=1e20):
    o = np.atleast_2d(np.asarray(o, np.float32))
    d = np.atleast_2d(np.asarray(d, np.float32))
    r = np.zeros(len(o), oracle.RAY_DTYPE)
    r["origin"], r["direction"], r["tmin"], r["tmax"] = o, d, tmin, tmax
    return r


def triangles_of(words, record):
    """(a, e1, e2, primitive) records of one instance's mesh."""
    r = RS_HEADER + RS_REC * record
    tris = int(words[r + 19])
    ends = [int(x) for i in range(int(words[0])) for x in words[RS_HEADER + RS_REC * i + 18:RS_HEADER + RS_REC * i + 20]
            if int(x) > tris]
    end = min(ends) if ends else len(words)
    return words[tris:end].reshape(-1, RS_TRI)


def test_reflection_bvh_keeps_primitive_identity():
    """reflection_software.rs:476-511: every triangle once, found by a ray through its
    off-centre point with its own primitive, (u, v) = (0.3, 0.2) and t = 1; a rebuild is
    word-identical."""
    pos, idx = grid()
    sc = rq.Scene.build([(pos, idx, 0)], [dict(model=np.eye(4), mesh=0, placement=3)], reflection=True)
    w = sc.words
    t = triangles_of(w, 0)
    assert sorted(t[:, 9].tolist()) == list(range(len(idx) // 3))
    p = pos[idx.reshape(-1, 3)]
    target = p[:, 0] * 0.5 + p[:, 1] * 0.3 + p[:, 2] * 0.2
    o = target + np.array([0.01, -0.02, 3.0], np.float32)
    h, _, _ = oracle.rs_traverse(w, rays_of(o, target - o), 0)
    assert ((h["flags"] & 1) == 1).all()
    assert np.array_equal(h["tri"][:, 0], np.arange(len(target)))
    assert np.abs(h["u"] - 0.3).max() < 1e-4 and np.abs(h["v"] - 0.2).max() < 1e-4
    assert np.abs(h["t"] - 1.0).max() < 1e-4
    assert (h["slot"] == 3).all() and (h["inst"] == 0).all()
    again = rq.Scene.build([(pos, idx, 0)], [dict(model=np.eye(4), mesh=0, placement=3)], reflection=True)
    assert np.array_equal(again.words, w)


def rot_y(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def test_reflection_instance_records():
    """reflection_software.rs:516-545: a singular model is skipped; the record's rows take the
    transformed corner back to (1, 0, 0); the world box encloses it; words 18-21 are the
    node base, triangle base, root and slot."""
    tri = (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([0, 1, 2], np.uint32), 0)
    m = np.eye(4)
    m[:3, :3] = rot_y(0.5) * 2.0
    m[:3, 3] = [3.0, -1.0, 4.0]
    sc = rq.Scene.build([tri], [dict(model=m, mesh=0, placement=5), dict(model=np.zeros((4, 4)), mesh=0, placement=6)],
                        reflection=True)
    w = sc.words
    assert w[0] == 1
    r = w[RS_HEADER:RS_HEADER + RS_REC]
    f = r.view(np.float32)
    world = (m @ np.array([1.0, 0, 0, 1]))[:3]
    obj = [f[k * 4] * world[0] + f[k * 4 + 1] * world[1] + f[k * 4 + 2] * world[2] + f[k * 4 + 3] for k in range(3)]
    assert np.abs(np.array(obj) - [1, 0, 0]).max() < 1e-5
    for k in range(3):
        assert f[12 + k] <= world[k] <= f[15 + k]
    nodes, tris, root, slot = (int(x) for x in r[18:22])
    assert slot == 5 and r[22] == 0 and r[23] == 0
    assert tris >= nodes and (tris - nodes) % 16 == 0
    if root & rq.LEAF_BIT:
        assert root == rq.LEAF_BIT | (0 << 4) | 1  # the first triangle, one of them
    t = w[tris:tris + RS_TRI]
    assert np.array_equal(t[:9].view(np.float32), [0, 0, 0, 1, 0, 0, 0, 1, 0]) and t[9] == 0
    # the scene bounds follow the one packed record and the top-level reserve for both
    # slots: word 4, lo then hi
    b = int(w[4])
    assert b == RS_HEADER + 1 * RS_REC + rq.instance_bvh.reserve(2)
    assert np.array_equal(w[b:b + 3], r[12:15]) and np.array_equal(w[b + 4:b + 7], r[15:18])


def random_reflection_scene(seed, instances=50, per=40):
    g = np.random.default_rng(seed)
    meshes = []
    for k in range(4):
        pts = g.normal(size=(3 * per, 3)).astype(np.float32)
        idx = np.arange(3 * per, dtype=np.uint32)
        if k == 0:  # duplicated triangles: equal t, the lower primitive wins
            idx = np.concatenate([idx, idx[:30]])
        meshes.append((pts, idx, 0))
    inst = []
    for i in range(instances):
        q = g.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        rot = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        m = np.eye(4)
        m[:3, :3] = rot * g.uniform(0.5, 2, 3)
        m[:3, 3] = g.uniform(-8, 8, 3)
        inst.append(dict(model=m, mesh=int(g.integers(4)), flags=int(g.random() < 0.2), placement=100 + i))
    return meshes, inst, g


def random_rays(g, n, spread=12.0):
    return rays_of(g.uniform(-spread, spread, (n, 3)), g.normal(size=(n, 3)))


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("top", [-1, 0])
def test_rs_trace_equals_f64_brute_force(seed, top):
    meshes, inst, g = random_reflection_scene(seed)
    sc = rq.Scene.build(meshes, inst, top_level=top, reflection=True)
    w = sc.words
    assert (w[1] != 0) == (top != 0)
    rays = random_rays(g, 3000)
    h, _, _ = oracle.rs_traverse(w, rays, 0)
    b = oracle.rs_brute_f64(w, rays, 0)
    hit = (h["flags"] & 1) == 1
    assert hit.sum() > 100
    assert np.array_equal(hit, b["hit"] == 1)
    dn = np.linalg.norm(rays["direction"].astype(np.float64), axis=1)
    on = np.linalg.norm(rays["origin"].astype(np.float64), axis=1)
    close = np.abs(h["t"].astype(np.float64) - b["t"]) * dn <= 1e-5 * (on + np.abs(b["t"]) * dn + 1.0)
    assert close[hit].all()
    agree = (h["inst"] == b["inst"]) & (h["tri"][:, 0] == b["slot"])
    assert (agree | close)[hit].all()
    assert agree[hit].mean() > 0.995
    # the record's slot is the instance's placement
    assert np.array_equal(h["slot"][hit], 100 + h["inst"][hit])
    # any hit: blocked exactly when a shadow-casting instance is hit
    a, vis, _ = oracle.rs_traverse(w, rays, 1, want_vis=True)
    ba = oracle.rs_brute_f64(w, rays, 1)
    assert np.array_equal((a["flags"] & 1) == 1, ba["hit"] == 1)
    bits = ((vis[np.arange(len(rays)) // 32] >> (np.arange(len(rays)) % 32)) & 1).astype(bool)
    assert np.array_equal(bits, ba["hit"] == 0)


def test_equal_distance_ties():
    """Duplicated triangles: the lower primitive wins; the same mesh placed twice at one
    transform: the lower instance wins (rs_limit and the tie rule)."""
    pos = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    idx = np.array([0, 1, 2, 0, 1, 2, 0, 1, 2], np.uint32)
    inst = [dict(model=np.eye(4), mesh=0, placement=9), dict(model=np.eye(4), mesh=0, placement=8)]
    for top in (-1, 0):
        sc = rq.Scene.build([(pos, idx, 0)], inst * 25, top_level=top, reflection=True)
        g = np.random.default_rng(4)
        o = np.c_[g.uniform(-0.3, 0.3, (200, 2)), np.full(200, 2.0)]
        h, _, _ = oracle.rs_traverse(sc.words, rays_of(o, [0, 0, -1]), 0)
        assert ((h["flags"] & 1) == 1).all()
        assert (h["inst"] == 0).all() and (h["tri"][:, 0] == 0).all() and (h["slot"] == 9).all()


def test_scene_bounds_pre_cull_and_no_shadow():
    pos = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    tri = (pos, np.array([0, 1, 2], np.uint32), 0)
    sc = rq.Scene.build([tri], [dict(model=np.eye(4), mesh=0, flags=1)], reflection=True)
    w = sc.words
    # a ray passing beside the scene box misses; one through it hits (closest) and is not
    # blocked (any hit: the instance casts no shadow)
    rays = rays_of([[5.0, 5.0, 2.0], [0.0, -0.2, 2.0]], [[0, 0, -1], [0, 0, -1]])
    h, _, _ = oracle.rs_traverse(w, rays, 0)
    assert h["flags"][0] == 0 and h["flags"][1] == 1
    assert abs(h["t"][1] - 2.0) < 1e-6
    a, _, _ = oracle.rs_traverse(w, rays, 1)
    assert (a["flags"] & 1).sum() == 0
    # a hit exactly on the box exit stays eligible: the triangle's plane is the box face
    h, _, _ = oracle.rs_traverse(w, rays_of([0.0, -0.2, 2.0], [0, 0, -1], 0.0, 2.0000002), 0)
    assert h["flags"][0] == 1


def test_invalid_rays_are_flagged_misses():
    pos, idx = grid(4)
    sc = rq.Scene.build([(pos, idx, 0)], [dict(model=np.eye(4), mesh=0)], reflection=True)
    r = rays_of([[1, 1, 2]] * 4, [[0, 0, -1]] * 4)
    r["origin"][0, 0] = np.nan
    r["tmin"][1] = -1.0
    r["direction"][2, 1] = np.inf
    h, _, _ = oracle.rs_traverse(sc.words, r, 0)
    assert list(h["flags"]) == [2, 2, 2, 1]


def test_rl_words_round_trip_and_checks():
    pos, idx = grid(6)
    sc = rq.Scene.build([(pos, idx, 0)], [dict(model=np.eye(4), mesh=0)] * 45, reflection=True)
    w = sc.words
    assert w[1] != 0  # 45 >= 40 instances: a top level
    again = rq.Scene.from_rl_words(w)
    assert np.array_equal(again.words, w)
    with pytest.raises(rq.InvalidInput):
        rq.Scene.from_rl_words(w[:7])
    bad = w.copy()
    bad[RS_HEADER + 19] = len(w) + 100
    with pytest.raises(rq.InvalidInput):
        rq.Scene.from_rl_words(bad)
    with pytest.raises(rq.InvalidInput):
        sc.layout_image("plain", 0)


def test_rl_words_must_hold_finite_trees_inside_the_buffer():
    """External rl_geometry words whose traversal would loop or read outside the buffer are
    refused: a node reached twice, a node or leaf past the end, a top-level leaf past the
    records."""
    pos, idx = grid(6)
    sc = rq.Scene.build([(pos, idx, 0)], [dict(model=np.eye(4), mesh=0)] * 45, reflection=True)
    w = sc.words
    rec = RS_HEADER
    nodes_at, root = int(w[rec + 18]), int(w[rec + 20])
    assert not root & rq.LEAF_BIT
    kids = w[nodes_at + 16 * root + 6:nodes_at + 16 * root + 10]
    inner = [s for s, c in enumerate(kids) if c != rq.EMPTY and not c & rq.LEAF_BIT]
    leaf = [s for s, c in enumerate(kids) if c != rq.EMPTY and c & rq.LEAF_BIT]

    def bad(edit, match):
        b = w.copy()
        edit(b)
        with pytest.raises(rq.InvalidInput, match=match):
            rq.Scene.from_rl_words(b)

    if inner:
        bad(lambda b: b.__setitem__(nodes_at + 16 * root + 6 + inner[0], root), "twice")
    slot = (inner + leaf)[0]
    bad(lambda b: b.__setitem__(nodes_at + 16 * root + 6 + slot, len(w)), "node past the end")
    bad(lambda b: b.__setitem__(nodes_at + 16 * root + 6 + slot, rq.LEAF_BIT | (len(w) << 4) | 2),
        "triangles past the end")
    top = int(w[1])
    bad(lambda b: b.__setitem__(top + 32 * int(w[2]) + 24, rq.LEAF_BIT | 45), "top-level leaf")


def rs_node_children(words, base, n):
    """Children of an rl_geometry BVH4 node: (tag, lo, hi) decoded as rs_trace does (o + s * q)."""
    w = words[base + 16 * n:base + 16 * n + 16]
    f = w.view(np.float32)
    out = []
    for s in range(4):
        c = int(w[6 + s])
        if c == rq.EMPTY:
            continue
        lo = np.array([f[a] + f[3 + a] * np.float32((int(w[10 + a]) >> (8 * s)) & 0xFF) for a in range(3)], np.float32)
        hi = np.array([f[a] + f[3 + a] * np.float32((int(w[13 + a]) >> (8 * s)) & 0xFF) for a in range(3)], np.float32)
        out.append((c, lo, hi))
    return out


def test_rs_cw8_relayout_covers_every_triangle_once():
    """The RS_CW8 image: the words are kept but for each record's root (word 20) and triangle base
    (word 19); each mesh's tree is built again from its triangle records, which follow the words in
    the new leaf order (the mesh's own rows, moved whole). Each record's CW8 tree reaches each of its
    mesh's triangles exactly once, every CW8 box on the way to a leaf encloses the leaf's triangles
    (a, a + e1, a + e2), and the CW8 top level reaches every instance once, its leaf boxes enclosing the
    instances' triangles in world space."""
    from test_layouts import w8_children
    meshes, inst, _ = random_reflection_scene(3, instances=60)
    sc = rq.Scene.build(meshes, inst, reflection=True)
    w = sc.words
    img = sc.layout_image("rs_cw8", 0)
    nodes = sc.layout_image("rs_cw8", 2)
    top = sc.layout_image("rs_cw8", 1)
    meta = sc.layout_image("rs_cw8", 5)
    count = int(w[0])
    keep = np.ones(len(w), bool)
    keep[RS_HEADER + RS_REC * np.arange(count) + 19] = False
    keep[RS_HEADER + RS_REC * np.arange(count) + 20] = False
    assert len(img) > len(w) and np.array_equal(img[:len(w)][keep], w[keep])

    def rs_leaves(base, tag):
        if tag & rq.LEAF_BIT:
            return {}
        out = {}
        for c, lo, hi in rs_node_children(w, base, tag):
            if c & rq.LEAF_BIT:
                out[c] = (lo, hi)
            else:
                out.update(rs_leaves(base, c))
        return out

    def rows(words, tb, first, n):
        return words[tb + 12 * first: tb + 12 * (first + n)].reshape(-1, 12)

    def cw8_leaves(tag):
        """(leaf tag, [CW8 boxes on the path to it]) under `tag`."""
        got = []
        for c, lo, hi in w8_children(nodes, tag):
            if c & rq.LEAF_BIT:
                got.append((c, [(lo, hi)]))
            else:
                got += [(t, [(lo, hi)] + path) for t, path in cw8_leaves(c)]
        return got

    tris_of = {}
    for i in range(count):
        r = RS_HEADER + RS_REC * i
        old_tb, new_tb = int(w[r + 19]), int(img[r + 19])
        want = rs_leaves(int(w[r + 18]), int(w[r + 20]))
        n = max(((t & ~rq.LEAF_BIT) >> 4) + (t & 15) for t in want)
        tris_of[i] = rows(img, new_tb, 0, n).view(np.float32).astype(np.float64)
        got = cw8_leaves(int(img[r + 20]))
        covered = np.zeros(n, int)
        for t, path in got:
            first, cnt = (t & ~rq.LEAF_BIT & 0xFFFFFFFF) >> 4, t & 15
            covered[first:first + cnt] += 1
            tri = rows(img, new_tb, first, cnt).view(np.float32).astype(np.float64)
            pts = np.concatenate([tri[:, 0:3], tri[:, 0:3] + tri[:, 3:6], tri[:, 0:3] + tri[:, 6:9]])
            for lo, hi in path:  # every CW8 box on the way down encloses the leaf's triangles
                assert (lo <= pts.min(axis=0)).all() and (hi >= pts.max(axis=0)).all()
        assert (covered == 1).all()
        assert np.array_equal(np.unique(rows(img, new_tb, 0, n), axis=0), np.unique(rows(w, old_tb, 0, n), axis=0))

    assert meta[0] == 1  # 60 instances: a top level
    def top_leaves(tag):
        if tag & rq.LEAF_BIT:
            return [tag & ~rq.LEAF_BIT]
        out = []
        for c, _, _ in w8_children(top, tag):
            out += top_leaves(c)
        return out
    assert sorted(top_leaves(int(meta[1]))) == list(range(count))

    # Each top-level leaf box encloses its instance's triangle vertices taken to world space in f64
    # (the inverse of the record's world-to-object rows), with a margin far above the f32 roundings of
    # an object-space ray; most are tighter than the record's world box.
    def leaf_boxes(tag, out):
        for c, lo, hi in w8_children(top, tag):
            if c & rq.LEAF_BIT:
                out[c & ~rq.LEAF_BIT] = (lo, hi)
            else:
                leaf_boxes(c, out)
        return out
    boxes = leaf_boxes(int(meta[1]), {})
    tighter = 0
    for i in range(count):
        r = RS_HEADER + RS_REC * i
        m = w[r:r + 12].view(np.float32).astype(np.float64).reshape(3, 4)
        t = tris_of[i]
        pts = np.concatenate([t[:, 0:3], t[:, 0:3] + t[:, 3:6], t[:, 0:3] + t[:, 6:9]])
        wv = (np.linalg.inv(m[:, :3]) @ (pts - m[:, 3]).T).T
        lo, hi = boxes[i]
        margin = 5e-6 * (1.0 + np.abs(wv))
        assert (lo <= wv - margin).all() and (hi >= wv + margin).all()
        blo, bhi = w[r + 12:r + 15].view(np.float32), w[r + 15:r + 18].view(np.float32)
        tighter += bool(np.prod(hi - lo, dtype=np.float64) < 0.9 * np.prod(bhi - blo, dtype=np.float64))
    assert tighter > count // 2

    # The kernel loads an instance's mesh-tree root from a copy beside the records (part 7), one
    # memory round trip with the record: each copy is that root node word for word, zero for a leaf.
    roots = sc.layout_image("rs_cw8", 7).reshape(-1, 24)
    nodes24 = nodes.reshape(-1, 24)
    assert len(roots) == count
    for i in range(count):
        r = int(img[RS_HEADER + RS_REC * i + 20])
        want = np.zeros(24, np.uint32) if r & rq.LEAF_BIT else nodes24[r]
        assert np.array_equal(roots[i], want)
