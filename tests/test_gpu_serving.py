"""Serving loop: continuous batching (one persistent launch at a time carrying
several jobs' primary, shadow and GI rays) under an HBM budget and a latency
budget. Each job's result (a device digest of its hit records and visibility
bits) is the same when the job runs alone and when it runs in the mix with
others in flight; jobs that cannot meet the latency budget are refused."""
import numpy as np
import pytest

from renderlab_amd import rq

pytestmark = pytest.mark.gpu

RESOLUTIONS = [(320, 240), (960, 640), (1920, 1080), (3840, 2160)]


def make_jobs(n, rate, seed, resolutions, scenes, mixes=(0, 1)):
    g = np.random.default_rng(seed)
    t = np.cumsum(g.exponential(1.0 / rate, n))
    jobs = []
    for k in range(n):
        w, h = resolutions[int(g.integers(len(resolutions)))]
        jobs.append(dict(scene=int(g.integers(scenes)), kind=1, width=w, height=h,
                         n_gi=int(mixes[int(g.integers(len(mixes)))]), seed=int(g.integers(1 << 30)),
                         arrival_s=float(t[k])))
    return jobs


def alone_vs_mix(scenes, jobs, layout="cw8", budget=1 << 30, slots=8):
    mixed = rq.serve(scenes, jobs, layout=layout, slots=slots, hbm_budget=budget)
    alone = [rq.serve(scenes, [dict(j, arrival_s=0.0)], layout=layout, slots=1, hbm_budget=budget)[0] for j in jobs]
    assert all(r["status"] == 0 for r in mixed)
    assert [r["digest"] for r in mixed] == [r["digest"] for r in alone]
    assert [r["secondary"] for r in mixed] == [r["secondary"] for r in alone]
    assert all(r["digest"] != 0 for r in mixed)
    return mixed


@pytest.mark.parametrize("layout", ["cw8", "plain"])
def test_each_job_is_the_same_alone_and_in_the_mix(layout):
    scenes = [rq.Scene.generate(1, 11, 300, 60000), rq.Scene.generate(1, 12, 500, 150000)]
    for s in scenes:
        s.upload((layout,))
    jobs = make_jobs(24, 4000.0, 5, [(320, 240), (640, 360), (960, 640)], len(scenes))
    print(f"serving {layout}: {len(jobs)} jobs, mixed", flush=True)
    mixed = alone_vs_mix(scenes, jobs, layout)
    # jobs overlapped: some job started before an earlier one ended
    starts = sorted((r["start_s"], r["end_s"]) for r in mixed)
    assert any(b[0] < a[1] for a, b in zip(starts, starts[1:]))


def test_budget_refuses_a_job_larger_than_the_arena_and_serves_the_rest():
    scene = rq.Scene.generate(1, 13, 100, 20000)
    scene.upload(("cw8",))
    jobs = [dict(scene=0, kind=1, width=320, height=240, n_gi=1, seed=1, arrival_s=0.0),
            dict(scene=0, kind=1, width=3840, height=2160, n_gi=1, seed=2, arrival_s=0.0),
            dict(scene=0, kind=1, width=320, height=240, n_gi=0, seed=3, arrival_s=0.0)]
    out = rq.serve([scene], jobs, slots=2, hbm_budget=64 << 20)
    assert [r["status"] for r in out] == [0, rq.RQ_ERR_BUDGET, 0]
    assert out[0]["digest"] != 0 and out[2]["digest"] != 0


def test_latency_budget_refuses_jobs_it_cannot_finish_in_time():
    scene = rq.Scene.generate(1, 14, 500, 200000)
    scene.upload(("cw8",))
    # a burst at one instant: 2160p jobs cannot finish within 4 ms, 320x240 ones can while the queue is short
    jobs = [dict(scene=0, kind=1, width=3840, height=2160, n_gi=1, seed=k, arrival_s=0.0) for k in range(4)]
    jobs += [dict(scene=0, kind=1, width=320, height=240, n_gi=1, seed=10 + k, arrival_s=0.05 + 0.01 * k)
             for k in range(4)]
    # warm the loop's throughput estimate on the same scene first
    rq.serve([scene], [dict(j, arrival_s=0.0) for j in jobs[4:]], slots=4, hbm_budget=4 << 30)
    out = rq.serve([scene], jobs, slots=8, hbm_budget=4 << 30, latency_budget_ms=4.0)
    status = [r["status"] for r in out]
    print(f"latency budget 4 ms: statuses {status}", flush=True)
    assert all(s in (0, rq.RQ_ERR_DEADLINE) for s in status)
    assert status[:4] == [rq.RQ_ERR_DEADLINE] * 4  # a 2160p job alone needs well over 4 ms
    assert status[4:] == [0, 0, 0, 0]
    # refusals are immediate; admitted jobs equal their runs alone
    for r, j in zip(out, jobs):
        if r["status"] == 0:
            alone = rq.serve([scene], [dict(j, arrival_s=0.0)], slots=1, hbm_budget=4 << 30)[0]
            assert alone["digest"] == r["digest"]
        else:
            assert r["end_s"] == r["start_s"]


def test_serving_form_over_eight_resident_scenes():
    """The serving configuration: 8 resident scenes of 100K-4M triangles, Poisson arrivals, resolutions
    from 320x240 to 3840x2160, primary + shadow or primary + shadow + GI, a fixed HBM budget; every job
    equal to its run alone."""
    from bench import C3_TRIS
    scenes = []
    for k, t in enumerate(C3_TRIS):
        sc = rq.Scene.generate(1, 100 + k, int(min(5000, max(100, t // 800))), t, gpu=True)
        sc.upload(("cw8",))
        scenes.append(sc)
    jobs = make_jobs(32, 500.0, 9, RESOLUTIONS, len(scenes))
    print(f"serving form: {len(jobs)} jobs over {len(scenes)} scenes", flush=True)
    mixed = alone_vs_mix(scenes, jobs, budget=8 << 30, slots=16)
    assert len({j["scene"] for j in jobs}) >= 6
