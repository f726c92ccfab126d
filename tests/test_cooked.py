"""Cooked geometry: the cluster BVH sidecar ("RLB\\0") and near field ("RNF\\0") codecs, and scenes
packed from them. Pinned by the byte layout the reference documents and writes
(crates/asset/src/cluster_bvh/codec.rs, crates/asset/src/near_field.rs:540-583) -- restated here
as an independent writer -- and by the reference's own tests restated:
near_field.rs:784-828 (encodes_validates_and_rejects_corruption), :762-782
(bins_round_trip_through_their_centres) and src/gi_bvh.rs:344-410
(packs_instances_meshes_and_near_masks)."""
import struct

import numpy as np
import pytest

import oracle
from renderlab_amd import rq

HEADER, INST = 16, 32


def grid(n):
    """The reference tests' grid (near_field.rs:723-739): y = 0 plane, n x n quads."""
    p = np.array([[x / n, 0.0, y / n] for y in range(n + 1) for x in range(n + 1)], np.float32)
    i = []
    for y in range(n):
        for x in range(n):
            a = y * (n + 1) + x
            b, c = a + 1, a + n + 1
            i += [a, c, b, b, c, c + 1]
    return p, np.array(i, np.uint32)


def rlb_bytes(records, first_tris, n_pos, positions, indices):
    """The cluster BVH file of a runtime BVH (rq_bvh records): the layout of codec.rs:4-75 with the
    cook's leaf records (near_field.rs:380-401) and one zero-error level over every leaf."""
    nodes, leaves = records["nodes"], records["leaves"]
    out = b"RLB\0" + struct.pack("<7I", 1, len(nodes), len(leaves), 1, len(leaves), n_pos, len(indices))
    for nd in nodes:
        out += nd.astype("<u4").tobytes() + b"\0" * 16  # 16 record words, 4 zero child errors
    for k, (off, cnt) in enumerate(leaves):
        out += struct.pack("<6If I", 0, k, 0, 3 * first_tris[k], off, cnt, 0.0, 0)
    out += struct.pack("<2I2f2I", 0, records["root"], 0.0, 0.0, 0, len(leaves))
    out += np.arange(len(leaves), dtype="<u4").tobytes()
    pos4 = np.zeros((n_pos, 4), "<f4")
    pos4[:, :3] = positions
    return out + pos4.tobytes() + np.asarray(indices, "<u4").tobytes()


def leaf_first_triangles(records, src_idx):
    tri = {tuple(src_idx[3 * t:3 * t + 3]): t for t in range(len(src_idx) // 3)}
    return [tri[tuple(records["indices"][off:off + 3])] for off, _ in records["leaves"]]


def rnf_bytes(settings, source, blobs, placements):
    """near_field.rs:546-583: header, digest, length-prefixed RLB blobs, placements."""
    lod, reach, skip, bins = settings
    out = b"RNF\0" + struct.pack("<4I3f", 1, len(blobs), len(placements), bins, lod, reach, skip) + bytes(source)
    for b in blobs:
        out += struct.pack("<Q", len(b)) + b
    for mesh, transform, masks in placements:
        out += struct.pack("<I", mesh) + np.asarray(transform, "<f4").reshape(-1).tobytes()
        out += struct.pack("<Q", len(masks)) + np.asarray(masks, "<u4").tobytes()
    return out


def reference_field(bins=16):
    """The field of near_field.rs:784-807: grid(6) median BVH, one placement at x = 2."""
    p, i = grid(6)
    bvh = rq.ClusterBvh.from_mesh(p, i, sah=False)
    words = bins * bins // 32
    masks = (np.arange(len(p) * 2 * words, dtype=np.uint64) * 2654435761 % (1 << 32)).astype(np.uint32)
    transform = np.eye(4, dtype=np.float32)
    transform[3, :3] = [2.0, 0.0, 0.0]  # column 3 (translation), columns stored one after another
    return p, i, bvh, masks, transform


def test_rlb_layout_matches_an_independent_writer():
    p, i = grid(6)
    for sah in (False, True):
        rec = (rq.build_bvh_sah if sah else rq.build_bvh)(p, i).records()
        want = rlb_bytes(rec, leaf_first_triangles(rec, i), len(p), p, rec["indices"])
        bvh = rq.ClusterBvh.from_mesh(p, i, sah=sah)
        assert bvh.encode() == want
        again = rq.ClusterBvh.decode(want)
        assert again.encode() == want
        assert again.counts() == dict(nodes=len(rec["nodes"]), leaves=len(rec["leaves"]), levels=1,
                                      level_leaves=len(rec["leaves"]), positions=len(p), indices=len(i))


def test_rlb_rejects_corruption():
    p, i = grid(6)
    good = rq.ClusterBvh.from_mesh(p, i).encode()
    n_nodes = struct.unpack_from("<I", good, 8)[0]
    cases = {
        "truncated": good[:-1],
        "extra": good + b"\0",
        "magic": b"RLX\0" + good[4:],
        "version": good[:4] + struct.pack("<I", 2) + good[8:],
    }
    for name, data in cases.items():
        with pytest.raises(rq.InvalidInput):
            rq.ClusterBvh.decode(data)
    pos_at = 32 + n_nodes * 80 + struct.unpack_from("<I", good, 12)[0] * 32 + 24 + \
        struct.unpack_from("<I", good, 20)[0] * 4
    b = bytearray(good)
    b[pos_at + 12:pos_at + 16] = struct.pack("<f", 1.0)  # position w must be 0
    with pytest.raises(rq.InvalidInput, match="invalid cluster BVH position"):
        rq.ClusterBvh.decode(bytes(b))
    b = bytearray(good)
    b[pos_at:pos_at + 4] = struct.pack("<f", float("nan"))
    with pytest.raises(rq.InvalidInput, match="invalid cluster BVH position"):
        rq.ClusterBvh.decode(bytes(b))
    b = bytearray(good)  # shrink the root's first child box: lower code of slot 0 on x to 255
    root = n_nodes - 1
    b[32 + root * 80 + 40] = 255
    with pytest.raises(rq.InvalidInput, match="do not enclose"):
        rq.ClusterBvh.decode(bytes(b))
    b = bytearray(good)  # a nonzero error on a child
    b[32 + root * 80 + 64:32 + root * 80 + 68] = struct.pack("<f", 0.5)
    with pytest.raises(rq.InvalidInput):
        rq.ClusterBvh.decode(bytes(b))
    with pytest.raises(rq.InvalidInput, match="byte limit"):
        rq.ClusterBvh.decode(good, max_bytes=64)


def test_near_field_reference_case():
    p, i, bvh, masks, transform = reference_field()
    data = rnf_bytes((1.0, 0.5, 0.01, 16), [7] * 32, [bvh.encode()], [(0, transform, masks)])
    field = rq.NearField.decode(data)
    assert field.encode() == data
    info = field.info()
    assert info["meshes"] == 1 and info["placements"] == 1 and info["bins"] == 16
    assert info["source"] == bytes([7] * 32) and abs(info["reach"] - 0.5) < 1e-7
    for bad in (data[:-1], data + b"\0", data[:4] + b"\x09" + data[5:]):
        with pytest.raises(rq.InvalidInput):
            rq.NearField.decode(bad)
    short = rnf_bytes((1.0, 0.5, 0.01, 16), [7] * 32, [bvh.encode()], [(0, transform, masks[:-1])])
    with pytest.raises(rq.InvalidInput, match="masks or transform"):
        rq.NearField.decode(short)
    skip = rnf_bytes((1.0, 0.5, 1.0, 16), [7] * 32, [bvh.encode()], [(0, transform, masks)])
    with pytest.raises(rq.InvalidInput, match="settings"):
        rq.NearField.decode(skip)
    with pytest.raises(rq.InvalidInput):
        field.encode(max_bytes=64)
    two = rnf_bytes((1.0, 0.5, 0.01, 16), [7] * 32, [bvh.encode(), bvh.encode()], [(1, transform, masks)])
    assert rq.NearField.decode(two).info()["meshes"] == 2
    bad_mesh = rnf_bytes((1.0, 0.5, 0.01, 16), [7] * 32, [bvh.encode()], [(3, transform, masks)])
    with pytest.raises(rq.InvalidInput, match="placement mesh"):
        rq.NearField.decode(bad_mesh)


def test_bins_round_trip_through_their_centres():
    for bins in (8, 16, 32):
        for b in range(bins * bins):
            d = oracle.near_bin_direction(b, bins)
            assert abs(float(d @ d) - 1.0) < 1e-5
            assert oracle.near_bin(d, bins) == b
    axes = [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
    assert len({oracle.near_bin(np.array(a, np.float32), 16) for a in axes}) == 6


def test_cooked_scene_packs_records_and_masks():
    """src/gi_bvh.rs:344-410 restated: every field software traversal reads survives bit for bit;
    only the instance with the cooked transform's exact bits gets the masks."""
    p, i = grid(6)
    bvh = rq.ClusterBvh.from_mesh(p, i, sah=True)
    rec = rq.build_bvh_sah(p, i).records()
    masks = np.arange(len(p) * 16, dtype=np.uint32)
    model = np.eye(4, dtype=np.float32)
    model[3, :3] = [5.0, 0.0, 0.0]  # column-major: translation in column 3
    data = rnf_bytes((0.0, 1.0, 0.0, 16), [0] * 32, [bvh.encode()], [(0, model, masks)])
    field = rq.NearField.decode(data)
    row_major = model.T
    inst = [dict(model=row_major, mesh=0), dict(model=np.eye(4, dtype=np.float32), mesh=0)]
    scene = rq.Scene.build_cooked([("field", field, 0)], inst)
    w = scene.words
    nodes, leaves, positions, indices = (int(v) for v in w[HEADER + 18:HEADER + 22])
    assert nodes == HEADER + 2 * INST + rq.instance_bvh.reserve(2)
    assert np.array_equal(w[nodes:leaves], rec["nodes"].reshape(-1))
    assert np.array_equal(w[leaves:positions], rec["leaves"].reshape(-1))
    assert np.array_equal(w[positions:indices], p.reshape(-1).view(np.uint32))
    assert np.array_equal(w[indices:indices + len(i)], rec["indices"])
    base = int(w[HEADER + 23])
    assert base and np.array_equal(w[base:base + len(masks)], masks) and base + len(masks) == len(w)
    assert w[HEADER + INST + 23] == 0
    assert abs(w[HEADER + 12:HEADER + 13].view(np.float32)[0] - 5.0) < 1e-3
    # a near-field mesh, a cooked cluster BVH and a triangle list of the same grid pack the same
    # mesh words as rq_scene_build (SAH trees on both sides)
    for src in (("bvh", bvh, 0), ("mesh", (p, i, 0))):
        s2 = rq.Scene.build_cooked([src], inst[1:])
        ref = rq.Scene.build([(p, i, 0)], inst[1:])
        assert np.array_equal(s2.words, ref.words)


def test_cooked_scene_errors():
    p, i, bvh, masks, transform = reference_field(16)
    f16 = rq.NearField.decode(rnf_bytes((1.0, 0.5, 0.01, 16), [7] * 32, [bvh.encode()], [(0, transform, masks)]))
    m8 = masks[:len(p) * 2 * 2]
    f8 = rq.NearField.decode(rnf_bytes((1.0, 0.5, 0.01, 8), [7] * 32, [bvh.encode()], [(0, transform, m8)]))
    one = [dict(model=np.eye(4, dtype=np.float32), mesh=0)]
    with pytest.raises(rq.InvalidInput, match="share their bins"):
        rq.Scene.build_cooked([("field", f16, 0), ("field", f8, 0)], one)
    with pytest.raises(rq.InvalidInput, match="level"):
        rq.Scene.build_cooked([("bvh", bvh, 1)], one)
    with pytest.raises(rq.InvalidInput, match="mesh index"):
        rq.Scene.build_cooked([("field", f16, 2)], one)
