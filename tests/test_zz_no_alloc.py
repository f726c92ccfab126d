"""rq_trace_noalloc: rq_trace with its ray cursor and stack spill area taken from
a static device array instead of allocated memory. Called through the C ABI
twice in a row (the second launch smaller, so it runs over the first one's
leftover scratch), its results must equal rq_trace's on the same rays: every
hit record bit for bit for the REF and PLAIN layouts, and for CW8 the same
triangle except on equal-distance ties (whose visit order depends on how rays
share waves); the visibility bits equal in every layout."""
import numpy as np
import pytest

from helpers import classify_mismatch, compare_exact, gpu_workload, to_device
from renderlab_amd import rq

pytestmark = pytest.mark.gpu


def both(scene, layout, rays, n, mode):
    """(rq_trace, rq_trace_noalloc) records and bits of the first n rays."""
    out = []
    for fn in (scene.trace, scene.trace_noalloc):
        hits = rq.DeviceBuffer(max(n, 1) * 48)
        vis = rq.DeviceBuffer(max((n + 31) // 32, 1) * 4)
        fn(layout, rays, n, mode, hits=hits, vis=vis)
        rq.synchronize()
        out.append((hits.download(rq.HIT_DTYPE, n), vis.download(np.uint32, (n + 31) // 32)))
    return out


@pytest.mark.parametrize("layout", ["ref", "plain", "cw8"])
def test_noalloc_trace_equals_rq_trace_on_two_calls_in_a_row(layout):
    assert rq.lib.rq_trace_noalloc_scratch_bytes() == 32 * 256 + 131072 * 224 * 4
    scene = rq.Scene.generate(0, 1)
    scene.upload((layout,))
    _, prim, _, shadow, gis = gpu_workload(scene, 0, 1, 640, 480, 1, layout)
    for rays in (prim, gis[0]):  # 307,200 rays, then a smaller launch of GI rays
        dr = to_device(rays)
        for mode in (rq.CLOSEST, rq.ANY):
            (h0, v0), (h1, v1) = both(scene, layout, dr, len(rays), mode)
            assert np.array_equal(v0, v1)
            if mode != rq.CLOSEST:
                continue
            if layout == "cw8":
                bad, _ = classify_mismatch(h1, h0)
                assert len(bad) == 0
            else:
                assert len(compare_exact(h1, h0)) == 0
    # a launch of fewer rays than one wave, then of none
    dr = to_device(prim[:37])
    (h0, v0), (h1, v1) = both(scene, layout, dr, 37, rq.CLOSEST)
    assert np.array_equal(v0, v1) and np.array_equal(h0["inst"], h1["inst"])
    scene.trace_noalloc(layout, dr, 0, rq.CLOSEST)


def test_noalloc_trace_refuses_reordering():
    scene = rq.Scene.generate(0, 1)
    scene.upload(("plain",))
    rays = to_device(np.zeros(70000, rq.RAY_DTYPE))
    with pytest.raises(rq.RqError, match="RQ_REORDER"):
        scene.trace_noalloc("plain", rays, 70000, rq.CLOSEST | rq.REORDER)
