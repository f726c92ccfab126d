"""The reference's buffer form of a hit, vec4<u32>(placement + 1, primitive,
bits(u), bits(v)) with zero for a miss (shaders/primary_rays.wgsl:19,
shaders/reflections_software_bvh4.wgsl:42, 90), and the hit decoders
gi_hit_identity / gi_hit_primitive / gi_hit_placement
(shaders/gi_rays_software.wgsl:297-316).

The oracle's records are pinned by closed forms: rays aimed at known triangles
of instances with known placement and identity words give exactly those words,
the source primitive (the closest-hit remap) or GI_NO_PRIMITIVE, and the
barycentrics' bits; a merged world's position table gives the identity of the
last pair at or below the hit's position. The host decoder (rq_gi_hit_decode)
must equal the oracle on every hit."""
import numpy as np
import pytest

import oracle
from renderlab_amd import rq
from test_oracle_raygen import grid_scene

R = oracle.RAY_DTYPE
NO_PRIMITIVE = 0xFFFFFFFF


def aimed_rays(pos, idx, model, lift=3.0):
    """One ray per triangle toward its point with weights (0.5, 0.3, 0.2), from `lift` above in world space."""
    p = pos[idx.reshape(-1, 3)].astype(np.float64)
    target = p[:, 0] * 0.5 + p[:, 1] * 0.3 + p[:, 2] * 0.2
    world = (np.c_[target, np.ones(len(target))] @ model.astype(np.float64).T)[:, :3]
    o = world + np.array([0.01, -0.02, lift])
    rays = np.zeros(len(o), R)
    rays["origin"], rays["direction"], rays["tmax"] = o, world - o, 1e20
    return rays


def two_instance_scene():
    pos, idx = grid_scene(9)
    far = np.eye(4, dtype=np.float32)
    far[0, 3] = 50.0
    scene = rq.Scene.build([(pos, idx, 1), (pos, idx, 0)],
                           [dict(model=np.eye(4, dtype=np.float32), mesh=0, identity=42, placement=7),
                            dict(model=far, mesh=1, identity=0, placement=3)])
    return scene, pos, idx, far


def test_records_of_aimed_rays_are_the_closed_form():
    scene, pos, idx, far = two_instance_scene()
    n = len(idx) // 3
    rays = np.concatenate([aimed_rays(pos, idx, np.eye(4)), aimed_rays(pos, idx, far)])
    miss = np.zeros(2, R)
    miss["origin"], miss["direction"], miss["tmax"] = [0, 0, 10], [0, 0, 1], 1e20  # away from everything
    rays = np.concatenate([rays, miss])
    hits, _, _ = oracle.traverse(scene.words, rays, 0)
    rec, ids = oracle.hit_records(scene.words, hits)
    assert (hits["flags"][:2 * n] & 1).all() and not (hits["flags"][2 * n:] & 1).any()
    first, second = slice(0, n), slice(n, 2 * n)
    assert (rec[first, 0] == 8).all() and (rec[second, 0] == 4).all()  # placement + 1 (word 29)
    assert np.array_equal(rec[first, 1], np.arange(n))  # the source primitive: the mesh keeps primitives
    assert (rec[second, 1] == NO_PRIMITIVE).all()  # no primitive table (word 28 zero)
    assert np.array_equal(rec[:2 * n, 2], hits["u"][:2 * n].view(np.uint32))
    assert np.array_equal(rec[:2 * n, 3], hits["v"][:2 * n].view(np.uint32))
    assert np.abs(rec[:2 * n, 2:].view(np.float32) - [0.3, 0.2]).max() < 1e-4
    assert (ids[first] == 42).all() and (ids[second] == 0).all()  # word 24; 0 without identity or table
    assert (rec[2 * n:] == 0).all() and (ids[2 * n:] == 0).all()  # a miss is all zero


def merged_world(firsts, identities):
    """The one-instance scene of grid_scene(4) with word 24 zero and a position table after the words."""
    pos, idx = grid_scene(4)
    w = rq.Scene.build([(pos, idx, 0)], [dict(model=np.eye(4, dtype=np.float32), mesh=0, identity=9)]).words.copy()
    table = np.array([v for pair in zip(firsts, identities) for v in pair], np.uint32)
    r = 16
    w[r + 24], w[r + 25], w[r + 26] = 0, len(w), len(firsts)
    return np.concatenate([w, table]), pos, idx


def test_merged_world_identity_is_the_last_pair_at_or_below_the_position():
    words, pos, idx = merged_world([0, 10, 10, 17], [5, 6, 7, 8])
    scene = rq.Scene.from_gi_words(words)
    rays = aimed_rays(pos, idx, np.eye(4))
    hits, _, _ = oracle.traverse(words, rays, 0)
    assert (hits["flags"] & 1).all()
    _, ids = oracle.hit_records(words, hits)
    a = hits["tri"][:, 0]  # no primitive table: tri holds position indices
    want = np.where(a >= 17, 8, np.where(a >= 10, 7, 5))  # equal firsts: the later pair wins
    assert np.array_equal(ids, want)
    assert {5, 7, 8} <= set(ids.tolist())
    # host decoder: the same words
    for k in range(len(hits)):
        rec, ident = scene.hit_record(hits[k])
        assert ident == ids[k]
    # no pair at or below the position: identity 0
    words, _, _ = merged_world([100, 200], [1, 2])
    _, ids = oracle.hit_records(words, hits)
    assert (ids == 0).all()


def test_host_decoder_equals_the_oracle():
    scene, pos, idx, far = two_instance_scene()
    g = np.random.default_rng(3)
    rays = np.concatenate([aimed_rays(pos, idx, np.eye(4)), aimed_rays(pos, idx, far)])
    rays = np.concatenate([rays, np.zeros(50, R)])
    rays["origin"][-50:] = g.uniform(-5, 60, (50, 3))
    rays["direction"][-50:] = g.normal(size=(50, 3))
    rays["tmax"][-50:] = 1e20
    hits, _, _ = oracle.traverse(scene.words, rays, 0)
    rec, ids = oracle.hit_records(scene.words, hits)
    for k in range(len(hits)):
        r, ident = scene.hit_record(hits[k])
        assert np.array_equal(r, rec[k]) and ident == ids[k], k


def test_reflection_records_carry_slot_and_primitive():
    pos, idx = grid_scene(6)
    far = np.eye(4, dtype=np.float32)
    far[1, 3] = 40.0
    scene = rq.Scene.build([(pos, idx, 0)], [dict(model=np.eye(4, dtype=np.float32), mesh=0, placement=11),
                                             dict(model=far, mesh=0, placement=2)], reflection=True)
    rays = np.concatenate([aimed_rays(pos, idx, np.eye(4)), aimed_rays(pos, idx, far)])
    hits, _, _ = oracle.rs_traverse(scene.words, rays, 0)
    n = len(idx) // 3
    assert (hits["flags"] & 1).all()
    rec, ids = oracle.hit_records(scene.words, hits, reflection=True)
    assert (rec[:n, 0] == 12).all() and (rec[n:, 0] == 3).all()
    assert np.array_equal(rec[:n, 1], np.arange(n)) and np.array_equal(rec[n:, 1], np.arange(n))
    assert (ids == 0).all()
    for k in range(len(hits)):
        r, ident = scene.hit_record(hits[k])
        assert np.array_equal(r, rec[k]) and ident == 0
