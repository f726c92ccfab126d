"""One persistent launch over several scenes and many uneven batches gives
each batch exactly what its own rq_trace call gives (alone vs in the mix)."""
import numpy as np
import pytest

from helpers import gpu_trace
from renderlab_amd import rq

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout", ["plain", "cw8"])
def test_batched_equals_separate_launches(layout):
    scenes = [rq.Scene.generate(0, 1), rq.Scene.generate(1, 2, 200, 40000), rq.Scene.generate(0, 3)]
    for s in scenes:
        s.upload((layout,))
    g = np.random.default_rng(4)
    batches, host = [], []
    for b in range(24):
        si = int(g.integers(len(scenes)))
        n = int(np.exp(g.uniform(np.log(50), np.log(6000))))
        mode = int(g.choice([0, 1, 2]))
        r = np.zeros(n, rq.RAY_DTYPE)
        r["origin"] = g.uniform(-60, 60, (n, 3))
        r["direction"] = g.normal(size=(n, 3))
        r["tmax"] = 1e20
        dr = rq.DeviceBuffer.from_array(r)
        dh = rq.DeviceBuffer(n * 48) if mode == 0 else None
        dv = rq.DeviceBuffer(((n + 31) // 32) * 4) if mode != 0 else None
        batches.append((si, mode, n, dr, dh, dv))
        host.append(r)
    print(f"batched {layout}: {len(batches)} batches, {sum(b[2] for b in batches)} rays", flush=True)
    rq.trace_batches(scenes, layout, batches)
    rq.synchronize()
    for (si, mode, n, dr, dh, dv), r in zip(batches, host):
        want_h, want_v = gpu_trace(scenes[si], layout, r, mode)
        if mode == 0:
            got = dh.download(rq.HIT_DTYPE, n)
            assert np.array_equal(got.view(np.uint32), want_h.view(np.uint32))
        else:
            assert np.array_equal(dv.download(np.uint32, (n + 31) // 32), want_v)
