"""Ray reordering changes which lane traces which ray and nothing else: with
the sort on, every layout must return the same 48-byte hit records and the
same visibility bits as in submission order, also when the ray count is read
on the device. The order-off results are pinned against the oracle in the
parity tests; here the oracle also checks a sample of the reordered records.
"""
import numpy as np
import pytest

import oracle
from helpers import compare_exact, gpu_workload, to_device
from renderlab_amd import rq

pytestmark = pytest.mark.gpu

W, H = 384, 256  # 98,304 rays: above the 65,536-ray reorder threshold


@pytest.fixture(scope="module")
def scene_c1():
    s = rq.Scene.generate(1, 7, 64, 20000)
    s.upload(("ref", "plain", "cw8"))
    return s


@pytest.fixture(scope="module")
def gi_rays(scene_c1):
    print(f"workload: lognormal seed 7 {W}x{H} P+S+GI", flush=True)
    _, p, _, s, gis = gpu_workload(scene_c1, 1, 7, W, H, 1, layout="plain")
    rays = np.concatenate([p, s, gis[0]])
    assert len(rays) >= 65536
    return rays


def run(scene, layout, rays, mode, reorder, indirect=None):
    n = len(rays)
    mode = mode | (rq.REORDER if reorder else 0)
    dr = to_device(rays)
    dh = rq.DeviceBuffer(n * rq.HIT_DTYPE.itemsize)
    dv = rq.DeviceBuffer(((n + 31) // 32) * 4)
    if indirect is None:
        scene.trace(layout, dr, n, mode, hits=dh, vis=dv)
        m = n
    else:
        cnt = to_device(np.array([indirect], np.uint32))
        scene.trace_indirect(layout, dr, cnt, n, mode, hits=dh, vis=dv)
        m = indirect
    rq.synchronize()
    return dh.download(rq.HIT_DTYPE, n)[:m], dv.download(np.uint32, (n + 31) // 32)


@pytest.mark.parametrize("layout", ["ref", "plain", "cw8"])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_reorder_same_records(scene_c1, gi_rays, layout, mode):
    print(f"case: reorder {layout} mode {mode} n={len(gi_rays)}", flush=True)
    h0, v0 = run(scene_c1, layout, gi_rays, mode, 0)
    h1, v1 = run(scene_c1, layout, gi_rays, mode, 1)
    assert len(compare_exact(h1, h0)) == 0
    assert np.array_equal(v0, v1)


def test_reorder_indirect_count(scene_c1, gi_rays):
    m = len(gi_rays) - 1000
    print(f"case: reorder indirect plain count {m} of {len(gi_rays)}", flush=True)
    h0, _ = run(scene_c1, "plain", gi_rays, 0, 0, indirect=m)
    h1, _ = run(scene_c1, "plain", gi_rays, 0, 1, indirect=m)
    assert len(compare_exact(h1, h0)) == 0


def test_reorder_matches_oracle_sample(scene_c1, gi_rays):
    h1, _ = run(scene_c1, "plain", gi_rays, 0, 1)
    idx = np.random.default_rng(3).choice(len(gi_rays), 4000, replace=False)
    o, _, _ = oracle.traverse(scene_c1.words, gi_rays[idx], 0)
    assert len(compare_exact(h1[idx], o)) == 0
