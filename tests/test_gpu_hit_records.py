"""rq_hit_records on the GPU: the reference's 16-byte hit records and
gi_hit_identity for traced batches, equal word for word to the oracle's decode
of the same hits (pinned by closed forms in test_hit_records.py)."""
import numpy as np
import pytest

import oracle
from helpers import gpu_workload, to_device
from renderlab_amd import rq
from test_hit_records import aimed_rays, merged_world, two_instance_scene

pytestmark = pytest.mark.gpu


def device_records(scene, hits_dev, n):
    rec = rq.DeviceBuffer(max(n, 1) * 16)
    ids = rq.DeviceBuffer(max(n, 1) * 4)
    scene.hit_records(hits_dev, n, rec, ids)
    rq.synchronize()
    return rec.download(np.uint32, 4 * n).reshape(n, 4), ids.download(np.uint32, n)


@pytest.mark.parametrize("layout", ["ref", "plain", "cw8"])
def test_records_of_a_frame_equal_the_oracle(layout):
    scene = rq.Scene.generate(0, 1)
    scene.upload(("ref", layout))
    _, prim, hits, _, _ = gpu_workload(scene, 0, 1, 240, 160, 0, layout)
    n = len(prim)
    dh = to_device(hits)
    rec, ids = device_records(scene, dh, n)
    want, want_ids = oracle.hit_records(scene.words, hits)
    assert np.array_equal(rec, want) and np.array_equal(ids, want_ids)
    assert (rec[(hits["flags"] & 1) == 0] == 0).all()
    if layout != "cw8":  # the hits themselves are the oracle's bit for bit, so the records are too
        o, _, _ = oracle.traverse(scene.words, prim, 0)
        assert np.array_equal(rec, oracle.hit_records(scene.words, o)[0])


def test_records_and_identities_of_built_and_merged_scenes():
    scene, pos, idx, far = two_instance_scene()
    scene.upload(("ref",))
    rays = np.concatenate([aimed_rays(pos, idx, np.eye(4)), aimed_rays(pos, idx, far)])
    dr = to_device(rays)
    dh = rq.DeviceBuffer(len(rays) * 48)
    scene.trace("ref", dr, len(rays), rq.CLOSEST, hits=dh)
    hits = dh.download(rq.HIT_DTYPE, len(rays))
    rec, ids = device_records(scene, dh, len(rays))
    want, want_ids = oracle.hit_records(scene.words, hits)
    assert np.array_equal(rec, want) and np.array_equal(ids, want_ids)
    assert set(ids.tolist()) == {42, 0}
    words, pos, idx = merged_world([0, 10, 10, 17], [5, 6, 7, 8])
    merged = rq.Scene.from_gi_words(words)
    merged.upload(("ref",))
    rays = aimed_rays(pos, idx, np.eye(4))
    dr = to_device(rays)
    dh = rq.DeviceBuffer(len(rays) * 48)
    merged.trace("ref", dr, len(rays), rq.CLOSEST, hits=dh)
    hits = dh.download(rq.HIT_DTYPE, len(rays))
    rec, ids = device_records(merged, dh, len(rays))
    want, want_ids = oracle.hit_records(words, hits)
    assert np.array_equal(rec, want) and np.array_equal(ids, want_ids)
    assert {5, 7, 8} <= set(ids.tolist())
    # identities alone, and an empty batch
    only = rq.DeviceBuffer(len(rays) * 4)
    merged.hit_records(dh, len(rays), None, only)
    merged.hit_records(dh, 0, None, only)
    rq.synchronize()
    assert np.array_equal(only.download(np.uint32, len(rays)), want_ids)


@pytest.mark.parametrize("layout", ["rs", "rs_cw8"])
def test_reflection_records_equal_rs_closest(layout):
    scene = rq.Scene.generate(0, 1, reflection=True)
    scene.upload((layout,))
    w, h = 160, 96
    inv = rq.camera(0, 1, w, h)
    prim = rq.DeviceBuffer(w * h * 32)
    rq.gen_primary(inv, w, h, prim)
    dh = rq.DeviceBuffer(w * h * 48)
    scene.trace(layout, prim, w * h, rq.CLOSEST, hits=dh)
    rq.synchronize()
    rays = prim.download(rq.RAY_DTYPE, w * h)
    rec, ids = device_records(scene, dh, w * h)
    o, _, _ = oracle.rs_traverse(scene.words, rays, 0)
    want, _ = oracle.hit_records(scene.words, o, reflection=True)
    assert np.array_equal(rec, want) and (ids == 0).all()
    assert (rec[:, 0] != 0).sum() > 1000
