"""Relayouts, checked on the host: PLAIN boxes are the reference decode bit
for bit; CW8 leaves are parts of PLAIN leaves (split to fill free slots) whose
boxes, decoded exactly as the kernel decodes them, enclose their triangles;
both keep every triangle exactly once."""
import numpy as np
import pytest

from helpers import chain_words
from renderlab_amd import rq

LEAF, EMPTY = rq.LEAF_BIT, rq.EMPTY


def plain_children(nodes, i):
    w = nodes[i * 32:(i + 1) * 32]
    f = w.view(np.float32)
    out = []
    for s in range(4):
        if w[24 + s] == EMPTY:
            continue
        lo = np.array([f[a * 4 + s] for a in range(3)], np.float32)
        hi = np.array([f[12 + a * 4 + s] for a in range(3)], np.float32)
        out.append((int(w[24 + s]), lo, hi))
    return out


def w8_children(nodes, i):
    w = nodes[i * 24:(i + 1) * 24]
    p = w[:3].view(np.float32)
    sc = [np.float32(2.0) ** int(np.uint8((w[3] >> (8 * a)) & 0xFF).view(np.int8)) for a in range(3)]
    valid = int(w[3] >> 24)
    first = int(w[16 + ((valid & -valid).bit_length() - 1)])
    out = []
    for s in range(8):
        def q(plane):
            return (int(w[4 + plane * 2 + s // 4]) >> (8 * (s % 4))) & 0xFF
        if not (valid >> s) & 1:
            # an empty slot: an inverted box (low codes 255, high codes 0) and a copy of the
            # first child's tag, so visiting it by a slack-widened test changes no result
            assert w[16 + s] == first
            assert all(q(a) == 255 and q(3 + a) == 0 for a in range(3))
            continue
        lo = np.array([p[a] + sc[a] * np.float32(q(a)) for a in range(3)], np.float32)
        hi = np.array([p[a] + sc[a] * np.float32(q(3 + a)) for a in range(3)], np.float32)
        out.append((int(w[16 + s]), lo, hi))
    return out


def leaf_ranges(tag):
    return (tag & ((1 << 27) - 1), ((tag >> 27) & 0xF) + 1)


@pytest.mark.parametrize("kind,seed,meshes,tris", [(0, 1, 0, 0), (1, 5, 60, 20000)])
def test_plain_boxes_are_the_reference_decode(kind, seed, meshes, tris):
    sc = rq.Scene.generate(kind, seed, meshes, tris)
    w = sc.words
    inst = sc.layout_image("plain", 0).reshape(-1, 32)
    nodes = sc.layout_image("plain", 2)
    r = 16
    nodes_at = int(w[r + 18])
    node_count = (int(w[r + 19]) - nodes_at) // 16
    root = int(inst[0, 18])
    if root & LEAF:
        return
    base = root - int(w[r + 22])  # the mesh's first node in the PLAIN node array
    for n in range(node_count):
        ref = w[nodes_at + 16 * n: nodes_at + 16 * (n + 1)]
        o, s_ = ref[:3].view(np.float32), ref[3:6].view(np.float32)
        kids = plain_children(nodes, base + n)
        for k, (tag, lo, hi) in enumerate(kids):
            ql = np.array([(int(ref[10 + a]) >> (8 * k)) & 0xFF for a in range(3)], np.float32)
            qh = np.array([(int(ref[13 + a]) >> (8 * k)) & 0xFF for a in range(3)], np.float32)
            assert np.array_equal(lo.view(np.uint32), (o + s_ * ql).astype(np.float32).view(np.uint32))
            assert np.array_equal(hi.view(np.uint32), (o + s_ * qh).astype(np.float32).view(np.uint32))


@pytest.mark.parametrize("kind,seed,meshes,tris", [(0, 1, 0, 0), (1, 5, 60, 20000)])
def test_cw8_leaves_cover_their_mesh_and_their_boxes_enclose_their_triangles(kind, seed, meshes, tris):
    """CW8 keeps PLAIN's triangles, each moved at most within its mesh's range (its triangle
    record, side record and vertices together); the CW8 leaves under a mesh root, at most 8
    triangles each, cover the mesh's triangles exactly once; every CW8 child box encloses its
    leaf's vertices."""
    sc = rq.Scene.generate(kind, seed, meshes, tris)
    pn, cn = sc.layout_image("plain", 2), sc.layout_image("cw8", 2)
    pi, ci = sc.layout_image("plain", 0).reshape(-1, 32), sc.layout_image("cw8", 0).reshape(-1, 32)
    ptris, ctris = sc.layout_image("plain", 3).reshape(-1, 9), sc.layout_image("cw8", 3).reshape(-1, 9)
    pids, cids = sc.layout_image("plain", 4).reshape(-1, 6), sc.layout_image("cw8", 4).reshape(-1, 6)
    pv, cv = sc.layout_image("plain", 6).reshape(-1, 9), sc.layout_image("cw8", 6).reshape(-1, 9)
    assert ptris.shape == ctris.shape
    # one row per triangle: its record, side record (indices, slot, primitive) and vertices
    prow = np.concatenate([ptris, pids, pv], axis=1)
    crow = np.concatenate([ctris, cids, cv], axis=1)

    def span(tag):
        first = tag & ((1 << 27) - 1)
        return first, first + ((tag >> 27) & 0xF) + 1

    checked = 0
    for i in range(len(pi)):
        proot, croot = int(pi[i, 18]), int(ci[i, 18])
        if proot & LEAF:
            assert croot == proot
            continue
        plain_spans = []
        stack = [proot]
        while stack:
            t = stack.pop()
            for c, _, _ in plain_children(pn, t):
                (plain_spans.append(span(c)) if c & LEAF else stack.append(c))
        got = []
        stack = [croot]
        while stack:
            t = stack.pop()
            for c, lo, hi in w8_children(cn, t):
                if not c & LEAF:
                    stack.append(c)
                    continue
                a, b = span(c)
                assert b - a <= 8
                got.append((a, b))
                v = cv[a:b].reshape(-1, 3).view(np.float32)
                assert (lo <= v.min(axis=0)).all() and (hi >= v.max(axis=0)).all()
                checked += 1
        m0, m1 = min(a for a, _ in plain_spans), max(b for _, b in plain_spans)
        covered = np.zeros(m1, int)
        for a, b in got:
            covered[a:b] += 1
        assert (covered[:m0] == 0).all() and (covered[m0:m1] == 1).all()
        # the mesh's triangles, moved only among themselves (rows moved whole)
        assert np.array_equal(np.unique(crow[m0:m1], axis=0), np.unique(prow[m0:m1], axis=0))
        if i > 40:
            break
    assert checked > 0


def instance_world_vertices(sc, i):
    """The vertices of instance i's triangles (through its CW8 tree's leaves) taken to world space in f64
    by inverting the record's world-to-object rows."""
    inst, nodes = sc.layout_image("cw8", 0).reshape(-1, 32), sc.layout_image("cw8", 2)
    verts = sc.layout_image("cw8", 6).reshape(-1, 9).view(np.float32)
    spans, stack = [], [int(inst[i, 18])]
    while stack:
        t = stack.pop()
        if t & LEAF:
            spans.append(leaf_ranges(t))
            continue
        stack.extend(c for c, _, _ in w8_children(nodes, t))
    v = np.concatenate([verts[a:a + n] for a, n in spans]).reshape(-1, 3).astype(np.float64)
    rows = sc.words[16 + 32 * i:16 + 32 * i + 12].view(np.float32).astype(np.float64).reshape(3, 4)
    return (np.linalg.inv(rows[:, :3]) @ (v - rows[:, 3]).T).T


def check_cw8_top_level(sc):
    """Every instance is one leaf of the CW8 top level; each leaf box encloses the instance's vertices in
    world space with a margin far above the f32 roundings of an object-space ray, and, for instances whose
    transform rotates the mesh, is tighter than the record's world box (the box of the transformed mesh-box
    corners)."""
    meta = sc.layout_image("cw8", 5)
    top = sc.layout_image("cw8", 1)
    assert meta[0] == 1
    seen = np.zeros(int(meta[2]), int)
    stack = [int(meta[1])]
    w = sc.words
    tighter = 0
    while stack:
        t = stack.pop()
        if t & LEAF:
            i = t & ~LEAF & 0xFFFFFFFF
            seen[i] += 1
            continue
        for c, lo, hi in w8_children(top, t):
            if c & LEAF:
                i = c & ~LEAF & 0xFFFFFFFF
                wv = instance_world_vertices(sc, i)
                margin = 5e-6 * (1.0 + np.abs(wv))
                assert (lo <= wv - margin).all() and (hi >= wv + margin).all()
                r = 16 + 32 * i
                blo, bhi = w[r + 12:r + 15].view(np.float32), w[r + 15:r + 18].view(np.float32)
                tighter += bool(np.prod(hi - lo, dtype=np.float64) < 0.9 * np.prod(bhi - blo, dtype=np.float64))
            stack.append(c)
    assert (seen == 1).all()
    return tighter


def test_cw8_root_copies_equal_each_instance_root_node():
    """The kernel loads an instance's BLAS root from a copy beside the records (with the record, one
    memory round trip): each copy is the instance's root node word for word, zero for a leaf root."""
    for sc in (rq.Scene.generate(0, 2), rq.Scene.generate(1, 3, 300, 30_000)):
        inst = sc.layout_image("cw8", 0).reshape(-1, 32)
        nodes = sc.layout_image("cw8", 2).reshape(-1, 24)
        roots = sc.layout_image("cw8", 7).reshape(-1, 24)
        assert len(roots) == len(inst)
        for i, r in enumerate(inst[:, 18]):
            want = np.zeros(24, np.uint32) if r & LEAF else nodes[r]
            assert np.array_equal(roots[i], want)


def test_cw8_has_a_top_level_whatever_the_instance_count():
    """The CW8 kernel tests no instance's world box at entry: the top-level leaf box it came through
    holds the box of the instance's triangles. So CW8 builds a top level even where the words have
    none (fewer than 40 instances), reaching every instance once with leaf boxes that hold the
    instances' world vertices; its records are PLAIN's but for the root (word 18)."""
    g = np.random.default_rng(7)
    pos = g.uniform(-1, 1, (30, 3)).astype(np.float32)
    mesh = (pos, np.arange(30, dtype=np.uint32))
    def shifted(k):  # row-major model matrix: a translation by 3k along x
        m = np.eye(4)
        m[0, 3] = 3.0 * k
        return m

    one = rq.Scene.build([mesh], [dict(model=shifted(0), mesh=0)])
    sixty = rq.Scene.build([mesh], [dict(model=shifted(k), mesh=0) for k in range(60)], top_level=0)
    for sc in (rq.Scene.generate(1, 3, 10, 2000), one, sixty):
        assert sc.layout_image("plain", 5)[0] == 0 and sc.layout_image("cw8", 5)[0] == 1
        pi, ci = sc.layout_image("plain", 0).reshape(-1, 32), sc.layout_image("cw8", 0).reshape(-1, 32)
        other = np.ones(32, bool)
        other[18] = False
        assert np.array_equal(pi[:, other], ci[:, other])
        check_cw8_top_level(sc)


def test_cw8_top_level_reaches_every_instance_once():
    # rotated boxes: the transformed corners are the vertices, so only the 8 spheres' boxes tighten
    assert check_cw8_top_level(rq.Scene.generate(0, 2)) >= 1
    # rotated lognormal clusters: nearly every leaf box is tighter than the record box
    assert check_cw8_top_level(rq.Scene.generate(1, 3, 300, 30_000)) > 150


def check_cw8_mesh_leaves(sc):
    """Every CW8 mesh tree's leaves hold 1 to 8 triangles inside the triangle records, and each
    leaf's box encloses its triangles' vertices."""
    inst, nodes = sc.layout_image("cw8", 0).reshape(-1, 32), sc.layout_image("cw8", 2)
    verts = sc.layout_image("cw8", 6).reshape(-1, 9)
    for root in {int(r) for r in inst[:, 18]}:
        stack = [root]
        while stack:
            t = stack.pop()
            if t & LEAF:
                continue
            for c, lo, hi in w8_children(nodes, t):
                if not c & LEAF:
                    stack.append(c)
                    continue
                first, n = leaf_ranges(c)
                assert 1 <= n <= 8 and first + n <= len(verts)
                v = verts[first:first + n].reshape(-1, 3).view(np.float32)
                assert (lo <= v.min(axis=0)).all() and (hi >= v.max(axis=0)).all()


def test_cw8_trees_over_boxes_of_any_finite_size():
    """Meshes spanning 1 to 1e36 (a strip of triangles at x = 1.5^k) and far from the origin: the
    trees built from their triangles and the instances' boxes price areas relative to their root, so
    their costs stay finite -- the top level still reaches every instance once and every leaf keeps
    at most 8 triangles."""
    g = np.random.default_rng(3)
    meshes = []
    for base, n in ((1.5, 210), (1.2, 450), (2.0, 120)):
        xs = base ** np.arange(n)
        pos = np.array([[x, 0, 0] for x in xs] + [[x + 0.01, 0, 0] for x in xs] + [[x, 0.01, 0] for x in xs],
                       np.float32)
        idx = np.stack([np.arange(n), np.arange(n) + n, np.arange(n) + 2 * n], 1).reshape(-1).astype(np.uint32)
        meshes.append((pos, idx, 0))
    far = (g.normal(size=(3000, 3)) * 1e3 + [3e6, -2e6, 5e5]).astype(np.float32)
    meshes.append((far, np.arange(3000, dtype=np.uint32), 0))
    soup = g.normal(size=(3 * 2000, 3)).astype(np.float32)
    meshes.append((soup, np.arange(3 * 2000, dtype=np.uint32), 1))
    inst = []
    for k in range(48):
        m = np.eye(4, dtype=np.float32)
        m[:3, :3] *= g.uniform(0.5, 2.0)
        m[:3, 3] = g.uniform(-20, 20, 3)
        inst.append(dict(model=m, mesh=k % len(meshes)))
    sc = rq.Scene.build(meshes, inst)
    check_cw8_top_level(sc)
    check_cw8_mesh_leaves(sc)


def far_tiny_scene():
    """Millimetre triangles 5 km from the origin (object space), and
    instances of them placed across binade boundaries in world space."""
    g = np.random.default_rng(11)
    centres = 5000.0 + g.uniform(-0.5, 0.5, (400, 1, 3))
    pts = (centres + g.normal(scale=1e-3, size=(400, 3, 3))).astype(np.float32).reshape(-1, 3)
    meshes = [(pts, np.arange(len(pts), dtype=np.uint32), 0)]
    inst = []
    for k in range(48):
        m = np.eye(4)
        m[:3, 3] = [-5000.0 + 2.0 ** (k % 12) - 1.0, -5000.0 - 0.3 * k, -5000.0 + 1e-3 * k]
        inst.append(dict(model=m, mesh=0))
    return rq.Scene.build(meshes, inst)


@pytest.mark.parametrize("which", ["lognormal", "far_tiny"])
def test_cw8_codes_decode_exactly(which):
    """Every code of every CW8 node (BLAS and top level) decodes to an f32
    with no rounding, so the kernel's ray-relative evaluation of origin +
    2^e * q and the builder's enclosure check see the same plane."""
    sc = rq.Scene.generate(1, 5, 60, 20000) if which == "lognormal" else far_tiny_scene()
    checked = 0
    for part in (2, 1):
        nodes = sc.layout_image("cw8", part)
        w = nodes.reshape(-1, 24)
        if len(w) == 0:
            continue
        origin = w[:, :3].view(np.float32).astype(np.float64)
        q = np.arange(256, dtype=np.float64)
        for a in range(3):
            e = ((w[:, 3] >> (8 * a)) & 0xFF).astype(np.uint8).view(np.int8).astype(np.int64)
            assert (e >= -126).all()
            dec = origin[:, a:a + 1] + np.ldexp(1.0, e)[:, None] * q[None, :]  # exact in f64
            assert np.array_equal(dec.astype(np.float32).astype(np.float64), dec)
            checked += dec.size
    assert checked > 0


def test_plain_triangle_records_hold_vertex_and_edges():
    """Triangle records are vertex a and the f32 edges b - a, c - a (9 floats) of the
    positions the side record (ia, ib, ic, slot, primitive, 0) names; the slot is the
    triangle's index in the mesh's BVH order."""
    sc = rq.Scene.generate(0, 1)
    words = sc.words
    tris = sc.layout_image("plain", 3).view(np.float32).reshape(-1, 9)
    ids = sc.layout_image("plain", 4).reshape(-1, 6)
    inst = sc.layout_image("plain", 0).reshape(-1, 32)
    bases = sorted(set(int(x) for x in inst[:, 20]) | {len(tris)})
    checked = 0
    for i in range(0, len(inst), 97):
        rec = words[16 + 32 * i:16 + 32 * (i + 1)]
        pos = words[int(rec[20]):].view(np.float32)
        base = int(inst[i, 20])
        end = bases[bases.index(base) + 1]  # the mesh's triangles end where the next mesh's begin
        for k in range(base, min(base + 64, end)):
            ia, ib, ic = (int(x) for x in ids[k, :3])
            a = pos[3 * ia:3 * ia + 3]
            b = pos[3 * ib:3 * ib + 3]
            c = pos[3 * ic:3 * ic + 3]
            assert np.array_equal(tris[k, 0:3].view(np.uint32), a.view(np.uint32))
            assert np.array_equal(tris[k, 3:6].view(np.uint32), (b - a).view(np.uint32))
            assert np.array_equal(tris[k, 6:9].view(np.uint32), (c - a).view(np.uint32))
            assert ids[k, 3] == k - base and ids[k, 5] == 0
            idx = int(rec[21])
            assert [int(x) for x in words[idx + 3 * (k - base):idx + 3 * (k - base) + 3]] == [ia, ib, ic]
            checked += 1
    assert checked > 0


def test_cw8_rejects_trees_deeper_than_16_levels():
    """The CW8 kernel pushes without a stack bound check: 7 entries a level for at most 16 levels
    fit its stack. A mesh whose tree does not hold each of its triangles once keeps that tree's
    shape: a chain collapses 7 levels into one 8-wide node, so 100 levels give 15 and 120 give 18,
    and the relayout refuses the second (PLAIN still takes it). A chain over all of its mesh's
    triangles is built again from them, a few levels deep."""
    ok = rq.Scene.from_gi_words(chain_words(100, hole=True))
    nodes = ok.layout_image("cw8", 2).reshape(-1, 24)
    assert len(nodes) == 15
    deep = rq.Scene.from_gi_words(chain_words(120, hole=True))
    assert deep.layout_image("plain", 2).size == 120 * 32
    with pytest.raises(rq.InvalidInput, match="deeper than 16"):
        deep.layout_image("cw8", 2)
    rebuilt = rq.Scene.from_gi_words(chain_words(120))
    nodes = rebuilt.layout_image("cw8", 2).reshape(-1, 24)
    root = int(rebuilt.layout_image("cw8", 0)[18])
    levels, todo = 0, [(root, 1)]
    while todo:
        n, lv = todo.pop()
        levels = max(levels, lv)
        todo += [(int(c), lv + 1) for c in nodes[n, 16:24] if not int(c) & LEAF]
    assert 2 <= levels <= 4


def test_external_words_must_hold_finite_trees_inside_their_arrays():
    """rq_scene_from_gi_words refuses words whose traversal would read outside the buffer or never
    end: a node reached twice (here a cycle), ids past their arrays, a leaf past the indices, an
    index past the vertices, a top level pointing past the instance records."""
    base = chain_words(8)
    rq.Scene.from_gi_words(base)
    r = 16
    nodes_at, leaves_at, pos_at, idx_at = (int(v) for v in base[r + 18:r + 22])

    def bad(edit, match):
        w = base.copy()
        edit(w)
        with pytest.raises(rq.InvalidInput, match=match):
            rq.Scene.from_gi_words(w)

    def cycle(w):
        w[nodes_at + 16 * 3 + 7] = 5  # node 3's inner child becomes its ancestor 5
    bad(cycle, "twice")
    bad(lambda w: w.__setitem__(nodes_at + 16 * 2 + 7, 8), "past the mesh's nodes")
    bad(lambda w: w.__setitem__(r + 22, 99), "past the mesh's nodes")
    bad(lambda w: w.__setitem__(nodes_at + 6, LEAF | 9), "past the mesh's leaves")
    bad(lambda w: w.__setitem__(leaves_at + 2 * 4, 3 * 400), "past the end of the buffer")
    bad(lambda w: w.__setitem__(idx_at + 5, 27), "past the mesh's vertices")
    # a leaf whose index offset or count is not whole triangles (the reference reads indices
    # first + k, the device layouts step by triangles: such words would trace differently)
    bad(lambda w: w.__setitem__(leaves_at + 2 * 4, 3 * 4 + 1), "multiple of 3")
    bad(lambda w: w.__setitem__(leaves_at + 2 * 4 + 1, 2), "multiple of 3")
    # a leaf running past its mesh's indices (into words after them, still valid vertex indices): the
    # relayouts address a leaf's triangles inside the mesh's own range
    w = np.concatenate([base, np.zeros(60, np.uint32)])
    w[leaves_at + 2 * 2 + 1] = 3 * 15
    with pytest.raises(rq.InvalidInput, match="mesh's index count"):
        rq.Scene.from_gi_words(w)
    # the primitive table (word 28) and a merged world's identity table (words 24-26) are read
    # after hits; both must lie inside the buffer, the identity table in position order
    bad(lambda w: w.__setitem__(r + 28, len(base) - 3), "primitive table")
    bad(lambda w: w.__setitem__(r + 28, 0xFFFFFFF0), "primitive table")
    ok = base.copy()
    ok[r + 28] = len(base) - 9  # a table of the mesh's 9 triangles ending at the last word
    rq.Scene.from_gi_words(ok)

    def table(first, n):
        def edit(w):
            w[r + 24], w[r + 25], w[r + 26] = 0, first, n
        return edit
    bad(table(len(base) - 4, 3), "identity table")

    def descending(w):
        table(idx_at, 2)(w)
        w[idx_at], w[idx_at + 2] = 5, 1  # pairs (5, .) then (1, .); still valid vertex indices
    bad(descending, "position order")
    ok = base.copy()
    table(len(base) - 6, 3)(ok)  # the last index words as pairs (21, 22), (23, 24), (25, 26)
    rq.Scene.from_gi_words(ok)
    two = rq.Scene.build([(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.arange(3, dtype=np.uint32), 0)],
                         [dict(model=np.eye(4, dtype=np.float32), mesh=0)] * 50).words
    rq.Scene.from_gi_words(two)
    assert two[1] != 0
    w = two.copy()
    w[int(w[1]) + 24] = LEAF | 50
    with pytest.raises(rq.InvalidInput, match="top-level leaf"):
        rq.Scene.from_gi_words(w)
    w = two.copy()
    w[int(w[1]) + 24] = 0  # the root lists itself
    with pytest.raises(rq.InvalidInput, match="twice"):
        rq.Scene.from_gi_words(w)
