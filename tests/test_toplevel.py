"""Top level over instance boxes: the reference's own tests
(src/instance_bvh.rs:481-722), restated over the C ABI."""
import numpy as np

from renderlab_amd import rq

IB = rq.instance_bvh
LEAF, EMPTY = IB.LEAF, IB.EMPTY


class XorShift:
    def __init__(self, seed):
        self.s = seed

    def next(self, mod, div):
        s = self.s
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        self.s = s
        return np.float32(s % mod) / np.float32(div)


def children(nodes, nid):
    at = nid * 32
    f = nodes.view(np.float32)
    out = []
    for c in range(4):
        k = int(nodes[at + 24 + c])
        if k == EMPTY:
            continue
        lo = [f[at + a * 4 + c] for a in range(3)]
        hi = [f[at + 12 + a * 4 + c] for a in range(3)]
        out.append((k, (lo, hi)))
    return out


def wide_depth(nodes, nid):
    kids = [k for k, _ in children(nodes, nid) if not k & LEAF]
    return 1 + max((wide_depth(nodes, k) for k in kids), default=0)


def boxes_array(boxes):
    return np.array([list(lo) + list(hi) for lo, hi in boxes], np.float32)


def test_top_level_covers_every_instance_once():
    g = XorShift(0x12345678)
    boxes = []
    for _ in range(257):
        x, y, z = g.next(1000, 10), g.next(1000, 10), g.next(1000, 10)
        boxes.append(((x, y, z), (x + np.float32(1), y + np.float32(2), z + np.float32(0.5))))
    nodes = IB.build(boxes_array(boxes))
    assert nodes.size <= IB.reserve(len(boxes))
    seen = np.zeros(len(boxes), int)
    stack = [(0, None)]
    f32 = lambda v: np.array(v, np.float32)
    while stack:
        nid, parent = stack.pop()
        if nid & LEAF:
            i = nid & ~LEAF & 0xFFFFFFFF
            seen[i] += 1
            assert np.array_equal(f32(parent[0]), f32(boxes[i][0])) and np.array_equal(f32(parent[1]), f32(boxes[i][1]))
            continue
        for k, b in children(nodes, nid):
            if parent is not None:
                assert all(parent[0][a] <= b[0][a] and b[1][a] <= parent[1][a] for a in range(3))
            stack.append((k, b))
    assert (seen == 1).all()
    assert 1 + 3 * wide_depth(nodes, 0) <= 48


def test_top_level_walk_reaches_every_crossed_instance():
    g = XorShift(0x0BAD5EED)
    nx = lambda: g.next(10000, 100)
    boxes = []
    for i in range(3000):
        x, y, z = nx(), nx() * np.float32(0.1), nx()
        s = np.float32(30.0) if i % 50 == 0 else np.float32(0.5) + np.float32(i % 7) * np.float32(0.3)
        boxes.append(((x, y, z), (x + s, y + s * np.float32(0.5), z + s)))
    nodes = IB.build(boxes_array(boxes))

    def crosses(b, o, d):
        near, far = np.float32(0), np.float32(np.inf)
        with np.errstate(divide="ignore", invalid="ignore"):
            for a in range(3):
                p = (np.float32(b[0][a]) - o[a]) / d[a]
                q = (np.float32(b[1][a]) - o[a]) / d[a]
                near = max(near, min(p, q))
                far = min(far, max(p, q))
        return near <= far

    for r in range(200):
        o = np.array([nx(), nx() * np.float32(0.1) + np.float32(1), nx()], np.float32)
        d = np.array([nx() - np.float32(50) + np.float32(0.01), np.float32(r % 5) * np.float32(0.1) - np.float32(0.2) +
                      np.float32(0.003), nx() - np.float32(50) + np.float32(0.02)], np.float32)
        want = sorted(i for i, b in enumerate(boxes) if crosses(b, o, d))
        got, stack = [], [0]
        while stack:
            nid = stack.pop()
            for k, b in children(nodes, nid):
                if crosses(b, o, d):
                    if k & LEAF:
                        got.append(k & ~LEAF & 0xFFFFFFFF)
                    else:
                        stack.append(k)
        assert sorted(got) == want, f"ray {r}"


def test_refit_follows_moved_boxes():
    boxes = [(((i * 7) % 50, 0.0, 0.0), ((i * 7) % 50 + 1.0, 1.0, 1.0)) for i in range(50)]
    b = boxes_array(boxes)
    nodes = IB.build(b)
    moved = b.copy()
    moved[:, 1] += 3.0
    moved[:, 4] += 3.0
    IB.refit(nodes, moved)
    assert np.array_equal(nodes, IB.build(moved))


def test_place_sets_the_header_and_reuses_unchanged_records():
    header, stride = 8, 24
    words = np.zeros(header + 3 * stride, np.uint32)
    words[0] = 3
    for i in range(3):
        r = header + i * stride
        words[r + 12:r + 15] = np.float32(i * 4.0).view(np.uint32)
        words[r + 15:r + 18] = np.float32(i * 4.0 + 1.0).view(np.uint32)
    placed = IB.place(words, header, stride, 3)
    assert placed.size == words.size + IB.reserve(3)
    # Three records are under MIN_INSTANCES: no top level unless forced.
    assert placed[1] == 0
    forced = IB.place(words, header, stride, 3, force=1)
    assert (forced[1], forced[2]) == (words.size, 0)
    again = IB.place(words, header, stride, 3, previous=forced, force=1)
    assert np.array_equal(again, forced)


def test_place_refits_moved_records_then_rebuilds_after_refits():
    header, stride, count = 16, 32, 60
    words = np.zeros(header + count * stride, np.uint32)
    words[0] = count
    for i in range(count):
        r = header + i * stride
        words[r + 12:r + 15] = np.array([i, 0, 0], np.float32).view(np.uint32)
        words[r + 15:r + 18] = np.array([i + 1, 1, 1], np.float32).view(np.uint32)
    first = IB.place(words, header, stride, count)
    assert first[1] == words.size and first[3] == 0
    moved = words.copy()
    for i in range(count):
        r = header + i * stride
        moved[r + 13] = np.float32(2.0).view(np.uint32)
        moved[r + 16] = np.float32(3.0).view(np.uint32)
    second = IB.place(moved, header, stride, count, previous=first)
    assert second[3] == 1
    # The same records again: reused as they are.
    assert np.array_equal(IB.place(moved, header, stride, count, previous=second), second)
    prev = second
    for k in range(2, IB.REFITS + 1):
        prev = IB.place(words if k % 2 == 0 else moved, header, stride, count, previous=prev)
        assert prev[3] == k
    prev = IB.place(words if (IB.REFITS + 1) % 2 == 0 else moved, header, stride, count, previous=prev)
    assert prev[3] == 0  # rebuilt after REFITS refits
