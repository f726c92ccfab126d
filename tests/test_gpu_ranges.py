"""Several passes over one scene in one persistent launch (rq_trace_ranges):
each range's hit records and visibility bits must equal a separate rq_trace
of that range, bit for bit, for every layout, with modes mixed, an empty
range in the middle, and the four-range maximum."""
import numpy as np
import pytest

from helpers import classify_mismatch, compare_exact, gpu_trace, gpu_workload, to_device
from renderlab_amd import rq

pytestmark = pytest.mark.gpu

W, H = 240, 160


@pytest.fixture(scope="module")
def scene():
    s = rq.Scene.generate(1, 3, 80, 30000)
    s.upload(("ref", "plain", "cw8"))
    return s


@pytest.fixture(scope="module")
def passes(scene):
    print(f"workload: lognormal seed 3 {W}x{H} P+S+2GI", flush=True)
    _, p, _, s, gis = gpu_workload(scene, 1, 3, W, H, 2, layout="plain")
    return p, s, gis[0], gis[1]


def run_ranges(scene, layout, specs):
    """specs: [(rays, mode)]; returns [(hits, vis)] per range."""
    bufs, ranges = [], []
    for rays, mode in specs:
        n = len(rays)
        dr = to_device(rays) if n else rq.DeviceBuffer(32)
        dh = rq.DeviceBuffer(max(n, 1) * rq.HIT_DTYPE.itemsize)
        dv = rq.DeviceBuffer(max((n + 31) // 32, 1) * 4)
        dh.zero()
        bufs.append((dr, dh, dv, n))
        ranges.append((dr, n, mode, dh, dv))
    scene.trace_ranges(layout, ranges)
    rq.synchronize()
    return [(dh.download(rq.HIT_DTYPE, n), dv.download(np.uint32, (n + 31) // 32)) for _, dh, dv, n in bufs]


@pytest.mark.parametrize("layout", ["ref", "plain", "cw8"])
def test_ranges_equal_separate_launches(scene, passes, layout):
    p, s, g0, g1 = passes
    specs = [(s, rq.ANY), (np.zeros(0, rq.RAY_DTYPE), rq.ANY), (g0, rq.CLOSEST), (g1, rq.ANY_OPAQUE)]
    print(f"case: ranges {layout} sizes {[len(x) for x, _ in specs]}", flush=True)
    got = run_ranges(scene, layout, specs)
    for (rays, mode), (h, v) in zip(specs, got):
        if len(rays) == 0:
            continue
        h0, v0 = gpu_trace(scene, layout, rays, mode)
        if mode == rq.CLOSEST:
            assert len(compare_exact(h, h0)) == 0
        assert np.array_equal(v, v0)


@pytest.mark.parametrize("layout", ["plain", "cw8"])
def test_ranges_of_odd_sizes_equal_separate_launches(scene, passes, layout):
    """Ranges of 1, 63, 65 and 130 rays: the cursor's 64-ray chunks stop at each range's end. CW8 closest
    hits may differ only on equal-distance ties (its speculation follows the wave's other lanes)."""
    p, s, g0, g1 = passes
    specs = [(p[:1], rq.CLOSEST), (s[:63], rq.ANY), (g0[:65], rq.CLOSEST), (g1[:130], rq.ANY_OPAQUE)]
    print(f"case: ranges {layout} sizes {[len(x) for x, _ in specs]}", flush=True)
    got = run_ranges(scene, layout, specs)
    for (rays, mode), (h, v) in zip(specs, got):
        h0, v0 = gpu_trace(scene, layout, rays, mode)
        if mode == rq.CLOSEST and layout == "cw8":
            assert len(classify_mismatch(h, h0)[0]) == 0
        elif mode == rq.CLOSEST:
            assert len(compare_exact(h, h0)) == 0
        assert np.array_equal(v, v0)


def test_ranges_single_and_primary(scene, passes):
    p, s, g0, _ = passes
    print("case: ranges cw8 primary + shadow, and one range", flush=True)
    got = run_ranges(scene, "cw8", [(p, rq.CLOSEST), (s, rq.ANY)])
    h0, _ = gpu_trace(scene, "cw8", p, rq.CLOSEST)
    _, v0 = gpu_trace(scene, "cw8", s, rq.ANY)
    assert len(compare_exact(got[0][0], h0)) == 0
    assert np.array_equal(got[1][1], v0)
    one = run_ranges(scene, "plain", [(g0, rq.CLOSEST)])
    h1, _ = gpu_trace(scene, "plain", g0, rq.CLOSEST)
    assert len(compare_exact(one[0][0], h1)) == 0


def test_ranges_reject_bad_input(scene, passes):
    p = passes[0]
    dr = to_device(p[:64])
    with pytest.raises(RuntimeError):
        scene.trace_ranges("plain", [(dr, 64, 0, None, None)] * 5)
    with pytest.raises(RuntimeError):
        scene.trace_ranges("plain", [(dr, 64, 7, None, None)])


def test_tuner_times_the_launch_on_its_stream(scene, passes):
    """The autotuner's event timing must bracket launches on the caller's stream: the winner's
    time is within a factor of a separately timed launch with the installed tuning."""
    from renderlab_amd import autotune
    p, s, g0, _ = passes
    stream = rq.Stream()
    bufs = [to_device(x) for x in (p, s, g0)]
    hits = [rq.DeviceBuffer(len(x) * 48) for x in (p, g0)]
    vis = rq.DeviceBuffer(((len(s) + 31) // 32) * 4)
    ranges = [(bufs[0], len(p), rq.CLOSEST, hits[0], None), (bufs[1], len(s), rq.ANY, None, vis),
              (bufs[2], len(g0), rq.CLOSEST, hits[1], None)]
    base = rq.get_tuning("cw8", 0)
    try:
        t = autotune.tune_ranges(scene, "cw8", ranges, quick=True, stream=stream)
        start, end = rq.Event(), rq.Event()
        scene.trace_ranges("cw8", ranges, stream=stream)
        stream.synchronize()
        start.record(stream)
        for _ in range(3):
            scene.trace_ranges("cw8", ranges, stream=stream)
        end.record(stream)
        end.synchronize()
        ms = start.elapsed_ms(end) / 3
        print(f"tuned {t}, timed {ms:.4f} ms", flush=True)
        assert 0.33 * ms <= t["ms"] <= 3.0 * ms
    finally:
        rq.set_tuning("cw8", 0, base.block, base.stack, base.refill, base.grid_per_cu, speculate=base.speculate,
                      drain=base.drain)


def test_stream_wait_orders_two_streams(scene, passes):
    """rq_stream_wait_event: a copy queued on stream B after waiting on an event recorded on stream
    A behind a trace copies the finished records (equal to a plain trace's)."""
    p, _, _, _ = passes
    want, _ = gpu_trace(scene, "cw8", p, rq.CLOSEST)
    a, b = rq.Stream(), rq.Stream()
    dr = to_device(p)
    dh = rq.DeviceBuffer(len(p) * 48)
    dh.zero()
    out = rq.DeviceBuffer(len(p) * 48)
    done = rq.Event()
    print("case: trace on A, copy on B after A's event", flush=True)
    scene.trace("cw8", dr, len(p), rq.CLOSEST, hits=dh, stream=a)
    done.record(a)
    b.wait(done)
    rq.device_copy(out, dh, len(p) * 48, stream=b)
    b.synchronize()
    assert len(compare_exact(out.download(rq.HIT_DTYPE, len(p)), want)) == 0
    a.synchronize()
