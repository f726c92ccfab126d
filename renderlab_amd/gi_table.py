"""GI radiance-cache tables in the reference's word layout (shaders/gi_cells.wgsl:18-185, decoded
the same way by tests/ara_backend/gi_cache_reference.rs:50-118): a 64-word header, 24-word hot
records, cold records of max(24, header[23]) words. Only the words the ray stages read are set:
the seeded tables the tests and the benchmark feed to rq_gi_rays / rq_gi_gather_rays."""
import numpy as np

HEADER, HOT = 64, 24


def pack_normal(n):
    """gi_pack_normal (shaders/gi_cache.wgsl:1055-1059): octahedral map, then pack2x16snorm."""
    n = np.asarray(n, np.float32)
    p = n[:2] / np.float32(np.abs(n[0]) + np.abs(n[1]) + np.abs(n[2]))
    if n[2] < 0:
        p = (np.float32(1) - np.abs(p[::-1])) * np.where(p >= 0, np.float32(1), np.float32(-1))
    q = np.rint(np.clip(p, -1, 1) * np.float32(32767)).astype(np.int32)
    return int(q[0] & 0xffff) | (int(q[1] & 0xffff) << 16)


def make_table(cells, slots=None, min_edge=0.25, rotate=True, stride=1, cold_words=24):
    """cells: {slot: dict(normal, point, level=0, key_x=1, key_y=0, bits=0)}."""
    slots = slots or 1 << max(1, int(np.ceil(np.log2(max(cells) + 1))))
    cw = max(24, cold_words)
    w = np.zeros(HEADER + slots * (HOT + cw), np.uint32)
    w[1] = slots - 1
    w[2] = 4 if rotate else 0
    w[5] = np.float32(min_edge).view(np.uint32)
    w[21] = stride
    w[23] = cold_words
    for slot, c in cells.items():
        hot, cold = HEADER + slot * HOT, HEADER + slots * HOT + slot * cw
        w[hot] = (c.get("key_x", 1) & 0x7fffffff) | 0x80000000
        w[cold] = c.get("key_y", 0)
        w[cold + 3] = c.get("level", 0) & 0xff
        w[cold + 7] = pack_normal(c["normal"])
        w[cold + 8:cold + 11] = np.asarray(c["point"], np.float32).view(np.uint32)
        bits = int(c.get("bits", 0))
        w[cold + 12], w[cold + 13] = bits & 0xffffffff, bits >> 32
    return w


def random_cells(n, g, box=40.0, levels=4):
    cells = {}
    for slot in range(n):
        nrm = g.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        cells[slot] = dict(normal=nrm, point=g.uniform(-box, box, 3), level=int(g.integers(levels)),
                           key_x=int(g.integers(1, 1 << 31)), key_y=int(g.integers(0, 1 << 32)),
                           bits=int(g.integers(0, 1 << 63)) if slot % 3 == 0 else 0)
    return cells
