"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, by __graft_entry__.smoke() and by
the CPU-baseline leg of bench.py, always as the checker. The product library
never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("direction", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("inst", "<u4"), ("tri", "<u4", 3), ("u", "<f4"), ("v", "<f4"),
                      ("normal", "<f4", 3), ("flags", "<u4"), ("slot", "<u4")])
HIT64_DTYPE = np.dtype([("t", "<f8"), ("u", "<f8"), ("v", "<f8"), ("inst", "<u4"), ("slot", "<u4"),
                        ("hit", "<u4"), ("pad", "<u4")])
COUNTS_FIELDS = ["words", "top_nodes", "inst_tests", "inst_entries", "blas_nodes", "leaves", "tri_steps",
                 "tri_tests"]

if not os.path.exists(_LIB):
    raise ImportError(f"{_LIB} is not built: run `make -C oracle`")
lib = ctypes.CDLL(_LIB)
_P = ctypes.c_void_p
_u32 = ctypes.c_uint32
_sz = ctypes.c_size_t


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


_sig("orc_gi_traverse_batch", ctypes.c_int, _P, _sz, _P, _u32, _u32, _P, _P, _P, _u32)
_sig("orc_flat_f32", ctypes.c_int, _P, _sz, _P, _u32, _u32, _P)
_sig("orc_brute_f64", ctypes.c_int, _P, _sz, _P, _u32, _u32, _P, _u32)
_sig("orc_rs_trace_batch", ctypes.c_int, _P, _sz, _P, _u32, _u32, _P, _P, _P, _u32)
_sig("orc_rs_brute_f64", ctypes.c_int, _P, _sz, _P, _u32, _u32, ctypes.c_double, _P, _u32)
_sig("orc_reintersect_f64", ctypes.c_int, _P, _sz, _P, _P, _u32, _P)
_sig("orc_validate_blas", ctypes.c_int, _P, _sz, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _P)
_sig("orc_primary_ray", None, _P, _u32, _u32, _u32, _u32, _P)
_sig("orc_secondary_rays", _u32, _P, _P, _P, _u32, _u32, _u32, _P)
_sig("orc_pcg", _u32, _u32)
_sig("orc_pcg_batch", None, _P, _u32, _P)
_sig("orc_rand_sequence", None, _u32, _u32, _P)
_sig("orc_basis", None, _P, _P)
_sig("orc_cosine_dir", None, _P, ctypes.c_float, ctypes.c_float, _P)
_sig("orc_offset", None, _P, _P, _P)
_sig("orc_hit_records", None, _P, _sz, ctypes.c_int, _P, _u32, _P, _P)
_sig("orc_gi_item_rays", None, _P, _sz, _P, _u32, _P)
_sig("orc_near_bin", _u32, _P, _u32)
_sig("orc_near_bin_direction", None, _u32, _u32, _P)
_sig("orc_near_blocked", ctypes.c_int, _P, _sz, _u32, _u32, _P, _P, ctypes.c_int, _P)
_sig("orc_gi_gather_rays", _u32, _P, _sz, _P, _u32, _P, _P, _u32)


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def traverse(words, rays, mode, threads=0, want_hits=True, want_vis=False):
    """gi_traverse over a batch; returns (hits, vis_bits, counts dict)."""
    w = np.ascontiguousarray(words, np.uint32)
    r = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(len(r), HIT_DTYPE) if want_hits else None
    vis = np.zeros((len(r) + 31) // 32, np.uint32) if want_vis else None
    counts = np.zeros(8, np.uint64)
    fails = lib.orc_gi_traverse_batch(_p(w), w.size, _p(r), len(r), int(mode), _p(hits), _p(vis), _p(counts),
                                      int(threads))
    if fails:
        raise RuntimeError(f"oracle traversal read outside the buffer on {fails} rays")
    return hits, vis, dict(zip(COUNTS_FIELDS, (int(x) for x in counts)))


def rs_traverse(words, rays, mode, threads=0, want_hits=True, want_vis=False):
    """rs_trace (the reflection rays) over rl_geometry words; returns (hits, vis_bits, counts dict)."""
    w = np.ascontiguousarray(words, np.uint32)
    r = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(len(r), HIT_DTYPE) if want_hits else None
    vis = np.zeros((len(r) + 31) // 32, np.uint32) if want_vis else None
    counts = np.zeros(8, np.uint64)
    fails = lib.orc_rs_trace_batch(_p(w), w.size, _p(r), len(r), int(mode), _p(hits), _p(vis), _p(counts),
                                   int(threads))
    if fails:
        raise RuntimeError(f"oracle rs_trace read outside the buffer on {fails} rays")
    return hits, vis, dict(zip(COUNTS_FIELDS, (int(x) for x in counts)))


def rs_brute_f64(words, rays, mode, slack=4e-6, threads=0):
    """f64 brute force over rl_geometry words (hit64 slot = primitive)."""
    w = np.ascontiguousarray(words, np.uint32)
    r = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(len(r), HIT64_DTYPE)
    lib.orc_rs_brute_f64(_p(w), w.size, _p(r), len(r), int(mode), float(slack), _p(hits), int(threads))
    return hits


def flat_f32(words, rays, mode):
    w = np.ascontiguousarray(words, np.uint32)
    r = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(len(r), HIT_DTYPE)
    lib.orc_flat_f32(_p(w), w.size, _p(r), len(r), int(mode), _p(hits))
    return hits


def brute_f64(words, rays, mode, threads=0):
    w = np.ascontiguousarray(words, np.uint32)
    r = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(len(r), HIT64_DTYPE)
    lib.orc_brute_f64(_p(w), w.size, _p(r), len(r), int(mode), _p(hits), int(threads))
    return hits


def reintersect_f64(words, rays, hits):
    w = np.ascontiguousarray(words, np.uint32)
    r = np.ascontiguousarray(rays, RAY_DTYPE)
    h = np.ascontiguousarray(hits, HIT_DTYPE)
    out = np.zeros(len(r), HIT64_DTYPE)
    lib.orc_reintersect_f64(_p(w), w.size, _p(r), _p(h), len(r), _p(out))
    return out


def validate_blas(words, record):
    """Bvh::validate on the mesh of instance `record`; returns (code, depth)."""
    w = np.ascontiguousarray(words, np.uint32)
    r = 16 + 32 * record
    nodes, leaves, positions, indices = (int(x) for x in w[r + 18:r + 22])
    root = int(w[r + 22])
    node_count = (leaves - nodes) // 16
    leaf_count = (positions - leaves) // 2
    vertex_count = (indices - positions) // 3
    last = leaves + 2 * (leaf_count - 1)
    index_count = int(w[last] + w[last + 1]) if leaf_count else 0
    depth = np.zeros(1, np.uint32)
    code = lib.orc_validate_blas(_p(w), w.size, nodes, node_count, leaves, leaf_count, positions, vertex_count,
                                 indices, index_count, root, _p(depth))
    return code, int(depth[0])


def primary_rays(inverse, width, height):
    inv = np.ascontiguousarray(inverse, np.float32)
    out = np.zeros(width * height, RAY_DTYPE)
    one = np.zeros(1, RAY_DTYPE)
    for py in range(height):
        for px in range(width):
            lib.orc_primary_ray(_p(inv), px, py, width, height, _p(one))
            out[py * width + px] = one[0]
    return out


def secondary_rays(primary, hits, sun, seed, n_gi):
    """Per pixel (valid mask, rays[1 + n_gi])."""
    s = np.ascontiguousarray(sun, np.float32)
    masks = np.zeros(len(primary), np.uint32)
    rays = np.zeros((len(primary), 1 + n_gi), RAY_DTYPE)
    p = np.ascontiguousarray(primary, RAY_DTYPE)
    h = np.ascontiguousarray(hits, HIT_DTYPE)
    for i in range(len(p)):
        out = np.zeros(1 + n_gi, RAY_DTYPE)
        masks[i] = lib.orc_secondary_rays(_p(p[i:i + 1]), _p(h[i:i + 1]), _p(s), i, int(seed), int(n_gi), _p(out))
        rays[i] = out
    return masks, rays


def pcg(v):
    return int(lib.orc_pcg(int(v)))


def pcg_batch(values):
    v = np.ascontiguousarray(values, np.uint32)
    out = np.zeros(v.size, np.uint32)
    lib.orc_pcg_batch(_p(v), v.size, _p(out))
    return out


def rand_sequence(state, count):
    out = np.zeros(count, np.float32)
    lib.orc_rand_sequence(int(state) & 0xFFFFFFFF, int(count), _p(out))
    return out


def basis(n):
    """basis(n) as a 3x3 array whose columns are c0, c1, c2."""
    v = np.ascontiguousarray(n, np.float32)
    out = np.zeros(9, np.float32)
    lib.orc_basis(_p(v), _p(out))
    return out.reshape(3, 3).T


def cosine_dir(n, u1, u2):
    v = np.ascontiguousarray(n, np.float32)
    out = np.zeros(3, np.float32)
    lib.orc_cosine_dir(_p(v), float(u1), float(u2), _p(out))
    return out


def offset(p, ng):
    a = np.ascontiguousarray(p, np.float32)
    b = np.ascontiguousarray(ng, np.float32)
    out = np.zeros(3, np.float32)
    lib.orc_offset(_p(a), _p(b), _p(out))
    return out


def hit_records(words, hits, reflection=False):
    """(records (n, 4) u32: placement + 1, primitive, bits(u), bits(v); identities (n,) u32)."""
    w = np.ascontiguousarray(words, np.uint32)
    h = np.ascontiguousarray(hits, HIT_DTYPE)
    rec = np.zeros((len(h), 4), np.uint32)
    ids = np.zeros(len(h), np.uint32)
    lib.orc_hit_records(_p(w), w.size, int(bool(reflection)), _p(h), len(h), _p(rec), _p(ids))
    return rec, ids


def gi_item_rays(table, items):
    """gi_ray's rays for ray items (slot << 7 | k) of a GI cache table."""
    t = np.ascontiguousarray(table, np.uint32)
    it = np.ascontiguousarray(items, np.uint32)
    out = np.zeros(len(it), RAY_DTYPE)
    lib.orc_gi_item_rays(_p(t), t.size, _p(it), len(it), _p(out))
    return out


def gi_gather_rays(table, slots):
    """gi_gather's fixed-direction rays of the listed cells: (rays, items)."""
    t = np.ascontiguousarray(table, np.uint32)
    sl = np.ascontiguousarray(slots, np.uint32)
    n = lib.orc_gi_gather_rays(_p(t), t.size, _p(sl), len(sl), None, None, 0)
    rays = np.zeros(max(n, 1), RAY_DTYPE)
    items = np.zeros(max(n, 1), np.uint32)
    lib.orc_gi_gather_rays(_p(t), t.size, _p(sl), len(sl), _p(rays), _p(items), n)
    return rays[:n], items[:n]


def near_bin(d, bins):
    v = np.ascontiguousarray(d, np.float32)
    return int(lib.orc_near_bin(_p(v), int(bins)))


def near_bin_direction(b, bins):
    out = np.zeros(3, np.float32)
    lib.orc_near_bin_direction(int(b), int(bins), _p(out))
    return out


def near_blocked(words, bins, queries):
    """gi_near_blocked per query (records with inst, tri, bary, back, d)."""
    w = np.ascontiguousarray(words, np.uint32)
    out = np.zeros(len(queries), bool)
    for k, q in enumerate(queries):
        tri = np.ascontiguousarray(q["tri"], np.uint32)
        bary = np.ascontiguousarray(q["bary"], np.float32)
        d = np.ascontiguousarray(q["d"], np.float32)
        out[k] = bool(lib.orc_near_blocked(_p(w), w.size, int(bins), int(q["inst"]), _p(tri), _p(bary),
                                           int(q["back"]), _p(d)))
    return out
