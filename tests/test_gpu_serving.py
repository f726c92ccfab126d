"""Serving loop: continuous batching under an HBM budget. Each job's result
(a device digest of its hit records and visibility bits) is the same when
the job runs alone and when it runs in the mix with others in flight."""
import numpy as np
import pytest

from renderlab_amd import rq

pytestmark = pytest.mark.gpu


def make_jobs(n, rate, seed, resolutions, scenes):
    g = np.random.default_rng(seed)
    t = np.cumsum(g.exponential(1.0 / rate, n))
    jobs = []
    for k in range(n):
        w, h = resolutions[int(g.integers(len(resolutions)))]
        jobs.append(dict(scene=int(g.integers(scenes)), kind=1, width=w, height=h, n_gi=int(g.integers(2)),
                         seed=int(g.integers(1 << 30)), arrival_s=float(t[k])))
    return jobs


def test_each_job_is_the_same_alone_and_in_the_mix():
    scenes = [rq.Scene.generate(1, 11, 300, 60000), rq.Scene.generate(1, 12, 500, 150000)]
    for s in scenes:
        s.upload(("cw8",))
    jobs = make_jobs(24, 4000.0, 5, [(320, 240), (640, 360), (960, 640)], len(scenes))
    print(f"serving: {len(jobs)} jobs, mixed", flush=True)
    mixed = rq.serve(scenes, jobs, slots=4, hbm_budget=1 << 30)
    alone = []
    for j in jobs:
        one = dict(j, arrival_s=0.0)
        alone.append(rq.serve(scenes, [one], slots=1, hbm_budget=1 << 30)[0])
    assert all(r["status"] == 0 for r in mixed)
    assert [r["digest"] for r in mixed] == [r["digest"] for r in alone]
    assert [r["secondary"] for r in mixed] == [r["secondary"] for r in alone]
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
