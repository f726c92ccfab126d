"""Benchmark: Mrays/s of the software BVH ray query on one MI355X.

A step is one pass of the hot path over one frame of synthetic input, all on
the device: primary rays -> closest hit -> sun-shadow and cosine GI rays from
every hit -> any-hit shadow visibility and closest-hit GI. `value` is the
rays traced per second over the K timed steps (inputs resident in HBM).

Workload (default, the metric's precision form): 4M-triangle seeded scene
(5,000 lognormal-edge clusters in a 200 m cube plus ground), 3840x2160 primary
+ 1 sun shadow + 2 cosine GI bounces per primary hit, traced through the 8-wide
quantized layout; the plain FP32 layout is timed on the same rays for the
quantized-vs-plain ratio, and both trace one frame's identical rays for the
hit-ID agreement. The 1M-triangle 1080p frame (c1) is reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1|c0]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    # name: (scene kind, meshes, triangles, width, height, gi bounces, description)
    "c0": (0, 0, 0, 960, 640, 0, "2,000 boxes + 8 UV spheres + ground (31,682 tris), 960x640 primary + sun shadow"),
    "c1": (1, 5000, 1_000_000, 1920, 1080, 1,
           "1M-tri seeded scene (5,000 lognormal clusters, 200 m cube, ground), 1920x1080 primary + shadow + 1 GI"),
    "c2": (1, 5000, 4_000_000, 3840, 2160, 2,
           "4M-tri seeded scene (5,000 lognormal clusters, 200 m cube, ground), 3840x2160 primary + shadow + 2 GI"),
}
SUN = np.array([0.3, 0.8, 0.2], np.float32) / np.float32(np.linalg.norm([0.3, 0.8, 0.2]))
HBM_PEAK = 8.0e12


def log(msg):
    print(msg, file=sys.stderr, flush=True)


PROFILES = {"c1": "trace_kernel_c1_1080p_psgi.md", "c2": "trace_kernel_c2_2160p_ps2gi.md"}


def profiled_traffic(config, layout, launch):
    """HBM bytes per launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration) for this workload, layout and launch, from the committed profile summary; None if absent."""
    import re
    path = os.path.join(ROOT, "profiles", PROFILES.get(config, ""))
    if not os.path.isfile(path):
        return None
    pat = re.compile(r"^traffic (\w+) (\w+): read (\d+) B, written (\d+) B per launch")
    with open(path) as f:
        for line in f:
            m = pat.match(line.strip().lstrip("- "))
            if m and m.group(1) == layout and m.group(2) == launch:
                return int(m.group(3)) + int(m.group(4)), os.path.relpath(path, ROOT)
    return None


def profiled_issue(config, layout, launch):
    """The launch's VALU issue rate as a fraction of the chip's (SQ_INSTS_VALU over 1024 SIMDs x 2.4 GHz /
    4 cycles for its duration), measured by rocprofv3, from the committed profile summary; None if absent."""
    import re
    path = os.path.join(ROOT, "profiles", PROFILES.get(config, ""))
    if not os.path.isfile(path):
        return None
    pat = re.compile(r"^VALU issue (\w+) (\w+): .* = ([0-9.]+) of ")
    with open(path) as f:
        for line in f:
            m = pat.match(line.strip().lstrip("- "))
            if m and m.group(1) == layout and m.group(2) == launch:
                return float(m.group(3))
    return None


class Frame:
    """Device buffers of one frame and the launches of one step."""

    def __init__(self, rq, scene, kind, seed, w, h, n_gi, stream, reorder=False, schedule="fused", tiled=False):
        self.rq, self.scene, self.w, self.h, self.n_gi, self.seed = rq, scene, w, h, n_gi, seed
        # primary rays in 8x8 screen-tile order (a wave's 64 rays through one tile) or row-major
        self.tiled = tiled and w % 8 == 0 and h % 8 == 0
        # secondary launches sorted by (octant, origin Morton code) before tracing
        self.sec = rq.REORDER if reorder else 0
        # fused: shadow and GI rays as ranges of one persistent launch (they do not depend on each
        # other). pipelined: one launch per step carries the next frame's primary rays and this
        # frame's shadow and GI rays (software pipelining over a double-buffered primary stage).
        self.schedule = "separate" if reorder else schedule
        self.fused = self.schedule == "fused"
        self.parity = 0
        n = w * h
        self.n = n
        self.inv = rq.camera(kind, seed, w, h)
        self.prim = rq.DeviceBuffer(n * 32)
        self.hits = rq.DeviceBuffer(n * 48)
        # the second primary stage of the pipelined schedule
        self.prim2 = rq.DeviceBuffer(n * 32) if self.schedule == "pipelined" else None
        self.hits2 = rq.DeviceBuffer(n * 48) if self.schedule == "pipelined" else None
        self.shadow = rq.DeviceBuffer(n * 32)
        self.gi = rq.DeviceBuffer(max(n_gi, 1) * n * 32)
        self.gi_hits = rq.DeviceBuffer(max(n_gi, 1) * n * 48)
        self.vis = rq.DeviceBuffer(((n + 31) // 32) * 4)
        self.counts = rq.DeviceBuffer(4 * (n_gi + 1))
        self.stream = stream
        self.counts_host = None

    def stage(self, parity):
        return (self.prim, self.hits) if parity == 0 else (self.prim2, self.hits2)

    def step(self, layout, events=None):
        rq, s = self.rq, self.stream
        rec = (lambda name: events.mark(name)) if events else (lambda name: None)
        if self.schedule == "pipelined":
            # this frame's primary hits are in stage `parity` (traced by the previous step)
            cur_p, cur_h = self.stage(self.parity)
            nxt_p, nxt_h = self.stage(1 - self.parity)
            rq.gen_primary(self.inv, self.w, self.h, nxt_p, stream=s, tiled=self.tiled)
            rec("raygen")
            rq.gen_secondary(cur_p, cur_h, self.n, SUN, self.seed, self.n_gi, self.shadow,
                             self.gi if self.n_gi else None, self.counts, stream=s)
            rec("secondary_gen")
            ranges = [(nxt_p, self.n, rq.CLOSEST, nxt_h, None)] + self.secondary_ranges()
            self.scene.trace_ranges(layout, ranges, stream=s)
            rec("frame")
            self.parity = 1 - self.parity
            return
        rq.gen_primary(self.inv, self.w, self.h, self.prim, stream=s, tiled=self.tiled)
        rec("raygen")
        self.scene.trace(layout, self.prim, self.n, rq.CLOSEST, hits=self.hits, stream=s)
        rec("primary")
        rq.gen_secondary(self.prim, self.hits, self.n, SUN, self.seed, self.n_gi, self.shadow,
                         self.gi if self.n_gi else None, self.counts, stream=s)
        rec("secondary_gen")
        c = self.counts_host
        if self.fused:
            self.scene.trace_ranges(layout, self.secondary_ranges(), stream=s)
            rec("secondary")
            return
        self.scene.trace(layout, self.shadow, int(c[0]), rq.ANY | self.sec, vis=self.vis, stream=s)
        rec("shadow")
        for b in range(self.n_gi):
            ptr = self.gi.ptr.value + b * self.n * 32
            hp = self.gi_hits.ptr.value + b * self.n * 48
            self.scene.trace(layout, ptr, int(c[1 + b]), rq.CLOSEST | self.sec, hits=hp, stream=s)
            rec(f"gi{b}")

    def secondary_ranges(self):
        """GI ranges first, shadow last: the launch's pool then drains on the cheapest (any-hit) rays."""
        rq, c = self.rq, self.counts_host
        out = []
        for b in range(self.n_gi):
            out.append((self.gi.ptr.value + b * self.n * 32, int(c[1 + b]), rq.CLOSEST,
                        self.gi_hits.ptr.value + b * self.n * 48, None))
        out.append((self.shadow, int(c[0]), rq.ANY, None, self.vis))
        return out

    def tuning_ranges(self):
        """The ranges of this schedule's largest launch (for the autotuner), or None."""
        rq = self.rq
        if self.schedule == "pipelined":
            return [(self.prim, self.n, rq.CLOSEST, self.hits2, None)] + self.secondary_ranges()
        if self.schedule == "fused":
            return self.secondary_ranges()
        return None

    def launch_parts(self, name):
        """(rays pointer, count, mode) of each pass a timed launch traces."""
        rq, c = self.rq, self.counts_host
        shadow = [(self.shadow.ptr.value, int(c[0]), rq.ANY | self.sec)]
        gi = [(self.gi.ptr.value + b * self.n * 32, int(c[1 + b]), rq.CLOSEST | self.sec) for b in range(self.n_gi)]
        if name == "primary":
            return [(self.prim.ptr.value, self.n, rq.CLOSEST)]
        if name == "shadow":
            return shadow
        if name == "secondary":
            return shadow + gi
        if name == "frame":
            return [(self.prim.ptr.value, self.n, rq.CLOSEST)] + shadow + gi
        return [gi[int(name[2:])]]

    def prime(self, layout):
        """One untimed pass that fixes the secondary ray counts (deterministic) and, for the pipelined
        schedule, leaves the first frame's primary hits in stage 0."""
        rq = self.rq
        self.parity = 0
        rq.gen_primary(self.inv, self.w, self.h, self.prim, stream=self.stream, tiled=self.tiled)
        self.scene.trace(layout, self.prim, self.n, rq.CLOSEST, hits=self.hits, stream=self.stream)
        rq.gen_secondary(self.prim, self.hits, self.n, SUN, self.seed, self.n_gi, self.shadow,
                         self.gi if self.n_gi else None, self.counts, stream=self.stream)
        self.stream.synchronize()
        self.counts_host = self.counts.download(np.uint32)
        return self.rays_per_step()

    def rays_per_step(self):
        return int(self.n + sum(int(x) for x in self.counts_host))


class Marks:
    """HIP events between launches on the benchmark stream."""

    def __init__(self, rq, stream):
        self.rq, self.stream = rq, stream
        self.events = []

    def start(self):
        e = self.rq.Event()
        e.record(self.stream)
        self.events.append(("start", e))

    def mark(self, name):
        e = self.rq.Event()
        e.record(self.stream)
        self.events.append((name, e))

    def durations(self):
        out = {}
        for (_, a), (name, b) in zip(self.events, self.events[1:]):
            if name == "start":
                continue
            out.setdefault(name, []).append(a.elapsed_ms(b))
        return out


def time_layout(rq, frame, layout, steps, warmup):
    for _ in range(warmup):
        frame.step(layout)
    frame.stream.synchronize()
    marks = Marks(rq, frame.stream)
    t0 = time.perf_counter()
    marks.start()
    for _ in range(steps):
        frame.step(layout, marks)
        marks.mark("start")
    frame.stream.synchronize()
    wall = time.perf_counter() - t0
    per = marks.durations()
    total_ms = sum(sum(v) for v in per.values())
    return total_ms / steps, wall, {k: float(np.mean(v)) for k, v in per.items()}


C3_TRIS = [100_000, 250_000, 500_000, 1_000_000, 1_000_000, 2_000_000, 3_000_000, 4_000_000]


def c3_scenes(rq, layout, tris=C3_TRIS):
    scenes = []
    for k, t in enumerate(tris):
        meshes = int(min(5000, max(100, t // 800)))
        sc = rq.Scene.generate(1, 100 + k, meshes, t)
        sc.upload((layout, "plain") if layout != "plain" else ("plain",))
        scenes.append(sc)
    return scenes


def c3_batches(rq, scenes, layout, n_batches=256, seed=7):
    """256 batches, sizes lognormal in [4K, 2M], 40% primary / 30% shadow / 30% GI, drawn from
    per-scene pools of 1920x1080 primary rays and the shadow and GI rays of their hits."""
    g = np.random.default_rng(seed)
    pools = []
    for k, sc in enumerate(scenes):
        w, h = 1920, 1080
        n = w * h
        inv = rq.camera(1, 100 + k, w, h)
        prim = rq.DeviceBuffer(n * 32)
        rq.gen_primary(inv, w, h, prim)
        hits = rq.DeviceBuffer(n * 48)
        sc.trace(layout, prim, n, rq.CLOSEST, hits=hits)
        shadow = rq.DeviceBuffer(n * 32)
        gi = rq.DeviceBuffer(n * 32)
        counts = rq.DeviceBuffer(8)
        rq.gen_secondary(prim, hits, n, SUN, 100 + k, 1, shadow, gi, counts)
        rq.synchronize()
        c = counts.download(np.uint32)
        pools.append({0: (prim, n), 1: (shadow, int(c[0])), 2: (gi, int(c[1]))})
    sizes = np.clip(np.exp(g.normal(np.log(64_000), 1.6, n_batches)), 4096, 2_000_000).astype(int)
    kinds = g.choice([0, 1, 2], n_batches, p=[0.4, 0.3, 0.3])
    batches, outs = [], []
    for b in range(n_batches):
        si = int(g.integers(len(scenes)))
        buf, avail = pools[si][int(kinds[b])]
        n = int(min(sizes[b], avail))
        start = int(g.integers(0, max(avail - n, 0) + 1))
        mode = rq.ANY if kinds[b] == 1 else rq.CLOSEST
        hits = rq.DeviceBuffer(n * 48) if mode == rq.CLOSEST else None
        vis = rq.DeviceBuffer(((n + 31) // 32) * 4) if mode != rq.CLOSEST else None
        outs.append((hits, vis))
        batches.append((si, mode, n, buf.view(32 * start, 32 * n), hits, vis))
    return batches, pools, outs


def run_batching(rq, layout, quick_tune=True):
    from renderlab_amd import autotune
    t0 = time.perf_counter()
    scenes = c3_scenes(rq, layout)
    log(f"c3: 8 scenes built and uploaded in {time.perf_counter() - t0:.1f}s")
    batches, pools, outs = c3_batches(rq, scenes, layout)
    total = sum(b[2] for b in batches)
    tuned = autotune.tune_batches(scenes, layout, batches, quick=quick_tune)
    log(f"c3: {len(batches)} batches, {total} rays; tuned {tuned}")
    start, end = rq.Event(), rq.Event()
    reps = 3
    start.record()
    for _ in range(reps):
        rq.trace_batches(scenes, layout, batches)
    end.record()
    end.synchronize()
    one = start.elapsed_ms(end) / reps
    start.record()
    for _ in range(reps):
        for si, mode, n, rays, hits, vis in batches:
            scenes[si].trace(layout, rays, n, mode, hits=hits, vis=vis)
    end.record()
    end.synchronize()
    sep = start.elapsed_ms(end) / reps
    return {"rays": total, "batches": len(batches), "one_launch_ms": round(one, 3),
            "one_launch_mrays_s": round(total / one / 1e3, 1), "per_batch_launches_ms": round(sep, 3),
            "per_batch_mrays_s": round(total / sep / 1e3, 1),
            "tuning": {k: tuned[k] for k in ("block", "stack", "refill", "grid_per_cu", "speculate")}}


def run_serving(rq, layout, scenes, rate, n_jobs, resolutions, mixes, seed=3, slots=16, budget=16 << 30,
                latency_budget_ms=0.0):
    """A Poisson stream of render-query jobs through rq_serve: p50/p99 latency (arrival to completion) of the
    served jobs, refusals by the HBM and latency budgets, and Mrays/s over the run."""
    g = np.random.default_rng(seed)
    t = np.cumsum(g.exponential(1.0 / rate, n_jobs))
    jobs = []
    for k in range(n_jobs):
        w, h = resolutions[int(g.integers(len(resolutions)))]
        jobs.append(dict(scene=int(g.integers(len(scenes))), kind=1, width=w, height=h,
                         n_gi=int(mixes[int(g.integers(len(mixes)))]), seed=int(g.integers(1 << 30)),
                         arrival_s=float(t[k])))
    out = rq.serve(scenes, jobs, layout=layout, slots=slots, hbm_budget=budget, latency_budget_ms=latency_budget_ms)
    ok = [r for r in out if r["status"] == 0]
    lat = np.array([r["latency_ms"] for r in ok]) if ok else np.zeros(1)
    rays = sum(r["primary"] + r["secondary"] for r in ok)
    span = max(r["end_s"] for r in out) - min(r["arrival_s"] for r in out)
    return {"rate_jobs_s": rate, "jobs": n_jobs, "served": len(ok),
            "refused_deadline": sum(r["status"] == rq.RQ_ERR_DEADLINE for r in out),
            "refused_hbm": sum(r["status"] == rq.RQ_ERR_BUDGET for r in out),
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mrays_s": round(rays / span / 1e6, 1), "latency_budget_ms": latency_budget_ms}


def run_serving_forms(rq, layout):
    """The serving configuration over the 8 resident scenes: mixed resolutions and ray mixes at rates from
    50 to 2,000 jobs/s under a 16 ms latency budget; then 1080p primary + shadow + GI jobs without the latency
    budget at rising rates, whose p99 crossing 16 ms is the saturation point, and the top rate again with the
    budget."""
    t0 = time.perf_counter()
    scenes = []
    for k, tris in enumerate(C3_TRIS):
        sc = rq.Scene.generate(1, 100 + k, int(min(5000, max(100, tris // 800))), tris, gpu=True)
        sc.upload((layout,))
        scenes.append(sc)
    log(f"serving: 8 scenes built and uploaded in {time.perf_counter() - t0:.1f}s")
    res = [(320, 240), (960, 640), (1920, 1080), (3840, 2160)]
    mixed = [run_serving(rq, layout, scenes, rate=r, n_jobs=int(min(600, max(60, 0.4 * r))), resolutions=res,
                         mixes=[0, 1], latency_budget_ms=16.0) for r in (50, 200, 500, 1000, 2000)]
    log(f"serving, mixed resolutions, 16 ms budget: {mixed}")
    rates = (50, 100, 200, 300, 400, 500, 600, 700)
    sweep = [run_serving(rq, layout, scenes, rate=r, n_jobs=int(max(200, 0.5 * r)), resolutions=[(1920, 1080)],
                         mixes=[1]) for r in rates]
    over = [s["rate_jobs_s"] for s in sweep if s["p99_ms"] >= 16.0]
    held = [r for r in rates if not over or r < min(over)]
    slo = run_serving(rq, layout, scenes, rate=rates[-1], n_jobs=int(0.4 * rates[-1]), resolutions=[(1920, 1080)],
                      mixes=[1], latency_budget_ms=16.0)
    log(f"serving 1080p P+S+GI sweep: {sweep}; with the 16 ms budget at {rates[-1]}/s: {slo}")
    return ({"scenes": "8 resident (100K-4M triangles)", "resolutions": "320x240..3840x2160", "mixes": "P+S, P+S+GI",
             "rates": mixed},
            {"scenes": "8 resident (100K-4M triangles)", "sweep": sweep,
             "p99_under_16ms_at_every_rate_through_jobs_s": max(held) if held else None,
             "saturation_jobs_s": min(over) if over else None,
             "note": "saturation: the lowest swept rate whose p99 reaches 16 ms (no latency budget)",
             "with_16ms_budget_at_top_rate": slo})


def run_reflection(rq, kind, meshes, tris, frame, reps=3):
    """The reflection rays' traversal (rs_trace over rl_geometry) on the same seeded scene: the frame's
    primary and GI rays as closest-hit queries (rs_closest), the shadow rays as any-hit (rs_blocked)."""
    t0 = time.perf_counter()
    host = rq.Scene.generate(kind, 1, meshes, tris, reflection=True)
    rq.Scene.generate(kind, 1, meshes, tris, reflection=True, gpu=True)  # first GPU build: pins staging
    sc = rq.Scene.generate(kind, 1, meshes, tris, reflection=True, gpu=True)
    build = {"host_build_ms": round(host.stats()["build_us"] / 1e3, 2),
             "gpu_build_ms": round(sc.stats()["build_us"] / 1e3, 2),
             "words_equal": bool(np.array_equal(host.words, sc.words))}
    del host
    sc.upload(("rs", "rs_cw8"))
    build_s = time.perf_counter() - t0
    c = frame.counts_host
    n_gi = int(c[1]) if frame.n_gi else 0
    hits = rq.DeviceBuffer(max(frame.n, n_gi, 1) * 48)
    vis = rq.DeviceBuffer(((int(c[0]) + 31) // 32 + 1) * 4)
    ranges = [(frame.prim, frame.n, rq.CLOSEST, hits, None), (frame.shadow, int(c[0]), rq.ANY, None, vis)]
    if n_gi:
        ranges.insert(1, (frame.gi, n_gi, rq.CLOSEST, frame.gi_hits, None))
    rays = sum(r[1] for r in ranges)
    out = {"rays": int(rays), "build_and_upload_s": round(build_s, 2), "words": int(len(sc.words)),
           "scene_build": build}
    for layout in ("rs", "rs_cw8"):
        start, end = rq.Event(), rq.Event()
        sc.trace_ranges(layout, ranges)
        rq.synchronize()
        start.record()
        for _ in range(reps):
            sc.trace_ranges(layout, ranges)
        end.record()
        end.synchronize()
        ms = start.elapsed_ms(end) / reps
        nbytes = sum(sc.count_stats(layout, r[0], r[1], r[2])["bytes"] for r in ranges)
        out[layout] = {"ms": round(ms, 4), "mrays_s": round(rays / ms / 1e3, 1),
                       "bytes_per_ray": round(nbytes / max(rays, 1), 1),
                       "achieved_gb_s": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    # the RS_CW8 records equal the RS ones (rs_trace's closest hit does not depend on the visit order)
    check = rq.DeviceBuffer(max(frame.n, 1) * 48)
    sc.trace("rs", frame.prim, frame.n, rq.CLOSEST, hits=hits)
    sc.trace("rs_cw8", frame.prim, frame.n, rq.CLOSEST, hits=check)
    rq.synchronize()
    out["rs_cw8_records_equal"] = bool(np.array_equal(hits.download(np.uint32, 12 * frame.n),
                                                      check.download(np.uint32, 12 * frame.n)))
    out["mrays_s"] = out["rs_cw8"]["mrays_s"]
    return out


def run_refit(rq, n=724, instances=4, reps=20):
    """GPU refit of one skinned-style mesh (an n x n grid, 2n^2 triangles, drawn by `instances`
    instances) to new poses: BVH refit, instance boxes, top-level place and the PLAIN refresh,
    against near_field::refit_bvh (the CPU restatement) on the same mesh and poses."""
    pos = np.stack(np.meshgrid(np.arange(n + 1), np.arange(n + 1)), -1).reshape(-1, 2) / n
    pos = np.concatenate([pos, np.zeros((len(pos), 1))], 1).astype(np.float32)
    q = (np.arange(n)[None, :] + (n + 1) * np.arange(n)[:, None]).reshape(-1)
    idx = np.stack([q, q + 1, q + n + 2, q, q + n + 2, q + n + 1], 1).reshape(-1).astype(np.uint32)

    def pose(k):
        p = pos.astype(np.float64).copy()
        p[:, 2] += 0.15 * np.sin(6.0 * p[:, 0] + k) * np.cos(4.0 * p[:, 1])
        p[:, :3] += [0.01 * k, 0.0, 0.0]
        return p.astype(np.float32)

    inst = []
    for k in range(instances):
        m = np.eye(4, dtype=np.float32)
        m[0, 3] = 1.5 * k
        inst.append(dict(model=m, mesh=0))
    sc = rq.Scene.build([(pos, idx, 1)], inst)
    sc.upload(("ref", "plain"))
    poses = [pose(k) for k in range(3)]
    dev = [rq.DeviceBuffer.from_array(p) for p in poses]
    for k in range(3):
        sc.refit_mesh(0, dev[k], len(pos))
    t0 = time.perf_counter()
    for r in range(reps):
        sc.refit_mesh(0, dev[r % 3], len(pos))
    gpu_ms = (time.perf_counter() - t0) / reps * 1e3
    bvh = rq.build_bvh_sah(pos, idx)
    t0 = time.perf_counter()
    for r in range(3):
        rq.refit_bvh(bvh, poses[r % 3])
    cpu_ms = (time.perf_counter() - t0) / 3 * 1e3
    tris = len(idx) // 3
    return {"triangles": tris, "vertices": len(pos), "instances": instances, "gpu_ms": round(gpu_ms, 3),
            "gpu_mtris_s": round(tris / gpu_ms / 1e3, 1), "cpu_refit_bvh_ms": round(cpu_ms, 2), "cpu_threads": 1,
            "speedup": round(cpu_ms / gpu_ms, 1),
            "note": "gpu_ms: host clock around the synchronous call (bounds, refit, records, top level, PLAIN refresh)"}


def run_gi_cache(rq, scene, layout, cells=32768, reps=5):
    """The GI cache's ray stages on the benchmark scene: every direction of `cells` seeded cells (the
    reference's per-frame item capacity at 32,768 cells) as gi_rays items (ray generation + gi_visible
    into the staged bits), then gi_gather's fixed directions (those left blocked) as closest-hit rays."""
    from renderlab_amd.gi_table import make_table, random_cells
    g = np.random.default_rng(7)
    cell = random_cells(cells, g, box=95.0)
    for c in cell.values():
        c["bits"] = 0
    table = make_table(cell, stride=1)
    items = np.array([(s << 7) | k for s in range(cells) for k in range(64)], np.uint32)
    n = len(items)
    dt, di = rq.DeviceBuffer.from_array(table), rq.DeviceBuffer.from_array(items)
    work, bits = rq.DeviceBuffer(n * 32), rq.DeviceBuffer(((n + 31) // 32) * 4)
    start, end = rq.Event(), rq.Event()
    rq.gi_rays(scene, layout, dt, table.size, di, n, work, bits)
    rq.synchronize()
    start.record()
    for _ in range(reps):
        rq.gi_rays(scene, layout, dt, table.size, di, n, work, bits)
    end.record()
    end.synchronize()
    vis_ms = start.elapsed_ms(end) / reps
    staged = bits.download(np.uint32, (n + 31) // 32)
    # the staged bits become each cell's known-visible directions; the gather traces the rest
    cold = 64 + (int(table[1]) + 1) * 24
    cw = max(24, int(table[23]))
    for s in range(cells):
        table[cold + s * cw + 12] = staged[2 * s]
        table[cold + s * cw + 13] = staged[2 * s + 1]
    dt = rq.DeviceBuffer.from_array(table)
    slots = rq.DeviceBuffer.from_array(np.arange(cells, dtype=np.uint32))
    rays, gitems, count = rq.DeviceBuffer(n * 32), rq.DeviceBuffer(n * 4), rq.DeviceBuffer(4)
    hits, ws = rq.DeviceBuffer(n * 48), rq.DeviceBuffer(rq.gi_gather_workspace(cells))

    def gather():
        rq.gi_gather_rays(dt, table.size, slots, cells, rays, gitems, n, count, ws)
        scene.trace_indirect(layout, rays, count.ptr, n, rq.CLOSEST, hits=hits)

    gather()
    rq.synchronize()
    start.record()
    for _ in range(reps):
        gather()
    end.record()
    end.synchronize()
    gather_ms = start.elapsed_ms(end) / reps
    m = int(count.download(np.uint32, 1)[0])
    visible = int(np.unpackbits(staged.view(np.uint8)).sum())
    return {"cells": cells, "items": n, "visible": visible, "gi_rays_ms": round(vis_ms, 4),
            "gi_rays_mrays_s": round(n / vis_ms / 1e3, 1), "gather_rays": m, "gather_ms": round(gather_ms, 4),
            "gather_mrays_s": round(m / gather_ms / 1e3, 1), "layout": layout,
            "note": "gi_rays: item rays + gi_visible; gather: fixed-direction rays + closest hits (device count)"}


def cpu_baseline(scene_words, rays_sets, budget_s=12.0):
    """The oracle traversal (scalar C port of gi_traverse) on this host's cores, over a bounded sample."""
    import oracle
    threads = os.cpu_count() or 1
    sample = []
    for mode, rays in rays_sets:
        sample.append((mode, rays[: max(1, min(len(rays), 20000))]))
    # calibrate, then scale the sample to the budget
    t0 = time.perf_counter()
    n0 = 0
    for mode, r in sample:
        oracle.traverse(scene_words, r, mode, threads=threads, want_hits=False, want_vis=True)
        n0 += len(r)
    dt = time.perf_counter() - t0
    scale = max(1.0, budget_s / max(dt, 1e-6))
    total, secs = 0, 0.0
    for mode, rays in rays_sets:
        k = int(min(len(rays), max(1, int(20000 * scale))))
        t0 = time.perf_counter()
        oracle.traverse(scene_words, rays[:k], mode, threads=threads, want_hits=False, want_vis=True)
        secs += time.perf_counter() - t0
        total += k
    return total / secs / 1e6, threads, total


def build_scene(rq, cfg, compare_builds):
    """The configuration's seeded scene built on the GPU; with compare_builds, also on the host (timed,
    words checked equal) and a first GPU build (which pins the host staging buffer)."""
    kind, meshes, tris = CONFIGS[cfg][:3]
    build = None
    if compare_builds:
        t0 = time.perf_counter()
        host_scene = rq.Scene.generate(kind, 1, meshes, tris)
        log(f"scene {cfg}: {host_scene.stats()} synthesized and built on the host in {time.perf_counter() - t0:.2f}s")
        first = rq.Scene.generate(kind, 1, meshes, tris, gpu=True)
        gb_first = first.stats()["build_us"]
        del first
    scene = rq.Scene.generate(kind, 1, meshes, tris, gpu=True)
    if compare_builds:
        same = np.array_equal(scene.words, host_scene.words)
        if not same:
            raise SystemExit("GPU-built scene words differ from the host build")
        hb, gb = host_scene.stats()["build_us"], scene.stats()["build_us"]
        build = {"host_build_ms": round(hb / 1e3, 2), "gpu_build_ms": round(gb / 1e3, 2),
                 "speedup": round(hb / max(gb, 1), 2), "gpu_build_first_ms": round(gb_first / 1e3, 2),
                 "words_equal": bool(same), "host_threads": os.cpu_count(),
                 "note": "rq_scene_build (host BVH builder, threads over meshes) vs rq_scene_build_gpu, each timed "
                         "inside the call (inputs in host memory, packed words out); mesh synthesis excluded; "
                         "gpu_build_ms is a process's second build, gpu_build_first_ms its first (which also pins "
                         "the host staging buffer)"}
        del host_scene
        log(f"GPU build of the same scene: {build}")
    return scene, build


def time_config(rq, scene, cfg, layouts, args, stream):
    """Autotune and time each layout on the configuration's frame; returns (frame, rays per step,
    {layout: (ms/step, wall s, per-launch ms)}, {layout: tuning})."""
    kind, _, _, w, h, n_gi, _ = CONFIGS[cfg]
    frame = Frame(rq, scene, kind, 1, w, h, n_gi, stream, reorder=args.reorder == "secondary",
                  schedule=args.secondary, tiled=args.primary_order == "tiles")
    rays = frame.prime(layouts[0])
    log(f"{cfg}: rays per step {rays} (primary {frame.n}, secondary {[int(x) for x in frame.counts_host]})")
    results, tuned = {}, {}
    for layout in layouts:
        frame.prime(layout)
        tr = frame.tuning_ranges()
        fixed = dict(t.split("=", 1) for t in args.tuning).get(layout)
        if fixed:
            block, stack, refill, grid, spec = (int(x) for x in fixed.split(","))
            rq.set_tuning(layout, rq.CLOSEST, block, stack, refill, grid_per_cu=grid, speculate=spec)
            rq.set_tuning(layout, rq.ANY, block, stack, refill, grid_per_cu=grid, speculate=spec)
            tuned[layout] = {"block": block, "stack": stack, "refill": refill, "grid_per_cu": grid, "speculate": spec}
            log(f"{cfg} {layout}: fixed tuning {tuned[layout]}")
        elif tr is not None and not args.no_tune:
            from renderlab_amd import autotune
            t = autotune.tune_ranges(scene, layout, tr, stream=stream)
            tuned[layout] = {k: t[k] for k in ("block", "stack", "refill", "grid_per_cu", "speculate")}
            log(f"{cfg} {layout}: autotuned {t}")
            frame.prime(layout)
        ms, wall, per = time_layout(rq, frame, layout, args.steps, args.warmup)
        results[layout] = (ms, wall, per)
        log(f"{cfg} {layout}: {ms:.3f} ms/step, {rays / ms / 1e3:.1f} Mrays/s, per launch {per}")
    return frame, rays, results, tuned


def dominant_launch(scene, frame, layout, per):
    """The longest trace launch of a step, its rays and the kernel's step statistics on them (counted
    pass by pass: a ray's traversal and the bytes it moves do not depend on which launch carries it)."""
    names = [k for k in per if k in ("primary", "shadow", "secondary", "frame") or k.startswith("gi")]
    dom = max(names, key=lambda k: per[k])
    parts = frame.launch_parts(dom)
    tot = None
    for ptr, n, mode in parts:
        st1 = scene.count_stats(layout, ptr, n, mode)
        tot = st1 if tot is None else {k: tot[k] + st1[k] for k in tot}
    return dom, sum(n for _, n, _ in parts), tot


def compare_hits(q, p, rel=1e-5):
    """Quantized-layout hits q against plain-layout hits p of the same rays: hit-ID agreement (instance and
    BVH-order triangle), ties (|dt|/t <= rel), genuine mismatches, and t / barycentric differences where both
    found the same triangle."""
    hq, hp = (q["flags"] & 1) == 1, (p["flags"] & 1) == 1
    same_tri = hq & hp & (q["inst"] == p["inst"]) & (q["slot"] == p["slot"])
    diff = (hq != hp) | (hp & ~same_tri)
    dt = np.abs(q["t"].astype(np.float64) - p["t"]) / np.maximum(np.abs(p["t"].astype(np.float64)), 1e-30)
    tie = diff & hq & hp & (dt <= rel)
    out = {"rays": int(len(p)), "hits": int(hp.sum()), "differing_ids": int(diff.sum()), "ties": int(tie.sum()),
           "mismatches": int((diff & ~tie).sum())}
    if same_tri.any():
        out["max_t_rel_same_triangle"] = float(dt[same_tri].max())
        out["max_bary_same_triangle"] = float(max(np.abs(q["u"][same_tri] - p["u"][same_tri]).max(),
                                                  np.abs(q["v"][same_tri] - p["v"][same_tri]).max()))
    return out


def hit_id_agreement(rq, scene, frame, layout, base="plain"):
    """One frame's rays through both layouts: the primary rays, then shadow and GI rays generated from the
    base layout's primary hits, so both layouts trace identical rays. Closest hits are compared by hit ID,
    shadow rays by visibility bit."""
    n, n_gi = frame.n, frame.n_gi
    rq.gen_primary(frame.inv, frame.w, frame.h, frame.prim, tiled=frame.tiled)
    hb, hq = rq.DeviceBuffer(n * 48), rq.DeviceBuffer(n * 48)
    scene.trace(base, frame.prim, n, rq.CLOSEST, hits=hb)
    scene.trace(layout, frame.prim, n, rq.CLOSEST, hits=hq)
    rq.gen_secondary(frame.prim, hb, n, SUN, frame.seed, n_gi, frame.shadow, frame.gi if n_gi else None,
                     frame.counts)
    rq.synchronize()
    counts = frame.counts.download(np.uint32)
    passes = {"primary": compare_hits(hq.download(rq.HIT_DTYPE, n), hb.download(rq.HIT_DTYPE, n))}
    for b in range(n_gi):
        m = int(counts[1 + b])
        rays = frame.gi.view(b * n * 32, m * 32)
        scene.trace(base, rays, m, rq.CLOSEST, hits=hb)
        scene.trace(layout, rays, m, rq.CLOSEST, hits=hq)
        rq.synchronize()
        passes[f"gi{b}"] = compare_hits(hq.download(rq.HIT_DTYPE, m), hb.download(rq.HIT_DTYPE, m))
    m = int(counts[0])
    vb, vq = rq.DeviceBuffer(((m + 31) // 32) * 4 + 4), rq.DeviceBuffer(((m + 31) // 32) * 4 + 4)
    scene.trace(base, frame.shadow, m, rq.ANY, vis=vb)
    scene.trace(layout, frame.shadow, m, rq.ANY, vis=vq)
    rq.synchronize()
    words = (m + 31) // 32
    xor = vb.download(np.uint32, words) ^ vq.download(np.uint32, words)
    shadow_diff = int(np.unpackbits(xor.view(np.uint8)).sum())
    closest = list(passes.values())
    out = {"layouts": f"{layout} vs {base}", "closest_rays": sum(p["rays"] for p in closest),
           "differing_ids": sum(p["differing_ids"] for p in closest), "ties": sum(p["ties"] for p in closest),
           "mismatches": sum(p["mismatches"] for p in closest), "shadow_rays": m,
           "shadow_visibility_mismatches": shadow_diff,
           "max_t_rel_same_triangle": max(p.get("max_t_rel_same_triangle", 0.0) for p in closest),
           "max_bary_same_triangle": max(p.get("max_bary_same_triangle", 0.0) for p in closest),
           "passes": passes,
           "note": "hit ID = (instance, triangle); ties: a different triangle at |dt|/t <= 1e-5 (an equal-distance "
                   "candidate reached in another visit order); mismatches: any other difference"}
    return out


def gi_only(rq, scene, frame, layouts, stream, reps=3):
    """The frame's incoherent GI bounce rays alone (its GI ranges in one launch, no primary or shadow rays),
    timed with HIP events in each layout at that layout's tuned launch."""
    from renderlab_amd import autotune
    ranges = [r for r in frame.secondary_ranges() if r[2] == rq.CLOSEST]
    n = sum(int(r[1]) for r in ranges)
    out = {"rays": n}
    for layout in layouts:
        ms = autotune._time(lambda: scene.trace_ranges(layout, ranges, stream=stream), reps=reps, stream=stream)
        out[layout] = {"ms": round(ms, 4), "mrays_s": round(n / ms / 1e3, 1)}
    if "plain" in out and len(layouts) > 1:
        out["quantized_speedup"] = round(out["plain"]["ms"] / out[layouts[0]]["ms"], 3)
    return out


def fp64_all(rq, scene, frame, layout):
    """rq_verify_fp64 over every closest hit of one frame (primary and each GI bounce) in `layout`."""
    n, n_gi = frame.n, frame.n_gi
    rq.gen_primary(frame.inv, frame.w, frame.h, frame.prim, tiled=frame.tiled)
    scene.trace(layout, frame.prim, n, rq.CLOSEST, hits=frame.hits)
    rq.gen_secondary(frame.prim, frame.hits, n, SUN, frame.seed, n_gi, frame.shadow, frame.gi if n_gi else None,
                     frame.counts)
    rq.synchronize()
    counts = frame.counts.download(np.uint32)
    frame.counts_host = counts  # the frame's secondary buffers now hold these rays
    sets = [(frame.prim, frame.hits, n)]
    for b in range(n_gi):
        m = int(counts[1 + b])
        rays, hits = frame.gi.view(b * n * 32, m * 32), frame.gi_hits.view(b * n * 48, m * 48)
        scene.trace(layout, rays, m, rq.CLOSEST, hits=hits)
        sets.append((rays, hits, m))
    out = None
    for rays, hits, m in sets:
        v = scene.verify_fp64(rays, hits, m)
        if out is None:
            out = dict(v)
        else:
            for k in ("max_t_rel", "max_bary_abs", "max_t_scaled", "max_ulps_conditioned"):
                out[k] = max(out[k], v[k])
            for k in ("checked", "over_1e5", "conditioned"):
                out[k] += v[k]
    out["rays"] = "primary + " + ", ".join(f"GI bounce {b}" for b in range(n_gi)) if n_gi else "primary"
    out["layout"] = layout
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--layout", default="cw8", choices=["ref", "plain", "cw8"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip timing the plain layout")
    ap.add_argument("--no-agreement", action="store_true", help="skip the hit-ID agreement check against plain")
    ap.add_argument("--no-extra-c1", action="store_true", help="skip the 1M-triangle 1080p line (c1)")
    ap.add_argument("--batching", action="store_true", help="also run the 8-scene, 256-batch one-launch form")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving runs (8 scenes, rates 50..2000 jobs/s)")
    ap.add_argument("--reorder", default="none", choices=["none", "secondary"],
                    help="sort shadow and GI launches by direction octant and origin Morton code")
    ap.add_argument("--secondary", default="pipelined", choices=["fused", "separate", "pipelined"],
                    help="shadow and GI rays as ranges of one launch (fused), one launch per pass (separate), or "
                         "one launch per step with the next frame's primary rays (pipelined)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch tuning (no autotuner sweep)")
    ap.add_argument("--tuning", action="append", default=[], metavar="LAYOUT=B,S,R,G,SPEC",
                    help="fixed launch tuning of a layout instead of the autotuner's (block, stack, refill, "
                         "workgroups per CU, speculation), e.g. cw8=128,16,32,0,1")
    ap.add_argument("--primary-order", default="tiles", choices=["tiles", "rows"],
                    help="primary rays in 8x8 screen-tile order (rq_gen_primary_tiled) or row-major")
    ap.add_argument("--no-reflection", action="store_true", help="skip the reflection-ray (rs_trace) line")
    ap.add_argument("--no-refit", action="store_true", help="skip the GPU mesh refit line")
    ap.add_argument("--no-gi-cache", action="store_true", help="skip the GI cache ray-stage line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run one plain device copy of BYTES (the known-size reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the ray query shards by ray "
              "batch with the scene replicated per GPU; no multi-GPU figure is produced)", file=sys.stderr)
        sys.exit(2)

    from renderlab_amd import rq
    if args.copy_calibration:
        nb = (args.copy_calibration + 15) // 16 * 16
        src, dst = rq.DeviceBuffer(nb), rq.DeviceBuffer(nb)
        src.zero()
        rq.device_copy(dst, src, nb)
        rq.synchronize()
        log(f"copy calibration: {nb} bytes read and written by copy_kernel")
        del src, dst
    kind, meshes, tris, w, h, n_gi, desc = CONFIGS[args.config]
    rq.Scene.generate(1, 2, 20, 20000, gpu=True)  # load the build kernels before anything is timed
    scene, build = build_scene(rq, args.config, compare_builds=True)
    st = scene.stats()
    layouts = [args.layout] + ([] if args.no_compare or args.layout == "plain" else ["plain"])
    t0 = time.perf_counter()
    scene.upload(tuple(set(layouts) | {"plain", "ref"}))  # REF: original positions for the FP64 check
    log(f"uploaded {layouts} in {time.perf_counter() - t0:.1f}s; bytes " +
        ", ".join(f"{l}={scene.device_bytes(l)}" for l in set(layouts) | {"plain"}))
    stream = rq.Stream()
    frame, rays, results, tuned = time_config(rq, scene, args.config, layouts, args, stream)
    ms, wall, per = results[args.layout]
    value = rays / (ms * 1e-3) / 1e6

    dom, dom_rays, stats = dominant_launch(scene, frame, args.layout, per)
    nbytes = stats["bytes"]
    lanes = 64 * stats["iterations"]
    log(f"dominant launch {dom}: {dom_rays} rays, steps/ray node {stats['node_steps'] / dom_rays:.2f} "
        f"tri-pair {stats['tri_steps'] / dom_rays:.2f} inst {stats['inst_steps'] / dom_rays:.2f}; lane use "
        f"{(stats['node_steps'] + stats['tri_steps'] + stats['inst_steps']) / max(lanes, 1):.3f} (pass by pass); "
        f"per iteration: live lanes {stats['live_lanes'] / max(stats['iterations'], 1):.1f}, wanting node "
        f"{stats['want_node'] / max(stats['iterations'], 1):.1f} tri {stats['want_tri'] / max(stats['iterations'], 1):.1f} "
        f"inst {stats['want_inst'] / max(stats['iterations'], 1):.1f}")
    plain_stats = None
    if "plain" in results and args.layout != "plain":
        _, _, plain_stats = dominant_launch(scene, frame, "plain", per)
    achieved = nbytes / (per[dom] * 1e-3)
    # profile labels: primary, shadow, gi, ranges (the multi-range launch of the fused and pipelined schedules)
    kind_of = "gi" if dom.startswith("gi") else ("ranges" if dom in ("secondary", "frame") else dom)
    measured = profiled_traffic(args.config, args.layout, kind_of)
    roofline = {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK, 4), "traffic": measured[0] if measured else None,
                "traffic_source": measured[1] if measured else None,
                "kernel": {"secondary": "trace shadow+GI (ranges of one launch)",
                           "frame": "trace next primary+shadow+GI (ranges of one launch)"}.get(dom, f"trace {dom}"),
                "bytes_per_ray": round(nbytes / max(dom_rays, 1), 1), "launch_ms": round(per[dom], 4),
                "lane_use": round((stats["node_steps"] + stats["tri_steps"] + stats["inst_steps"]) /
                                  max(64 * stats["iterations"], 1), 3),
                "live_lanes": round(stats["live_lanes"] / max(64 * stats["iterations"], 1), 3),
                "steps_per_ray": {k: round(stats[f"{k}_steps"] / max(dom_rays, 1), 2) for k in ("node", "tri", "inst")},
                "valu_issue_frac": profiled_issue(args.config, args.layout, kind_of)}

    cpu = None
    if not args.no_cpu_baseline:
        counts = frame.counts_host
        prim = frame.prim.download(rq.RAY_DTYPE, frame.n)
        sh = frame.shadow.download(rq.RAY_DTYPE, int(counts[0]))
        sets = [(0, prim), (1, sh)]
        if n_gi:
            sets.append((0, frame.gi.download(rq.RAY_DTYPE, int(counts[1]))))
        mrays, threads, sampled = cpu_baseline(scene.words, sets)
        whole = sampled == sum(len(r) for _, r in sets)
        what = "every ray of one frame" if whole else "the first rays of each pass of one frame"
        cpu = {"value": round(mrays, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"{what} (primary, shadow, GI: {sampled} rays), the oracle's gi_traverse (C restatement "
                         f"of the reference traversal) on {threads} threads"}

    line = {
        "metric": "Mrays/s (primary+shadow+GI) on 1 MI355X, quantized vs plain; p50/p99 job latency",
        "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded scene and camera)",
        "config": {"workload": desc, "layout": args.layout, "rays_per_step": rays, "triangles": st["triangles"],
                   "instances": st["instances"], "reorder": args.reorder, "secondary": args.secondary,
                   "primary_order": "tiles" if frame.tiled else "rows",
                   "tuning": tuned.get(args.layout),
                   "parallelism": "rays across one GPU"},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    if "plain" in results and args.layout != "plain":
        pm = results["plain"][0]
        line["plain_f32"] = {"value": round(rays / (pm * 1e-3) / 1e6, 2), "ms_per_step": round(pm, 4),
                             "quantized_speedup": round(pm / ms, 3), "tuning": tuned.get("plain"),
                             "per_launch_ms": {k: round(v, 4) for k, v in results["plain"][2].items()}}
        if plain_stats:
            line["plain_f32"]["dominant_steps_per_ray"] = {
                k: round(plain_stats[f"{k}_steps"] / max(dom_rays, 1), 2) for k in ("node", "tri", "inst")}
            line["plain_f32"]["dominant_lane_use"] = round(
                (plain_stats["node_steps"] + plain_stats["tri_steps"] + plain_stats["inst_steps"]) /
                max(64 * plain_stats["iterations"], 1), 3)
        line["gi_rays_only"] = gi_only(rq, scene, frame, layouts, stream)
        log(f"GI bounce rays alone: {line['gi_rays_only']}")
        if not args.no_agreement:
            line["hit_id_agreement"] = hit_id_agreement(rq, scene, frame, args.layout)
            log(f"hit-ID agreement: {line['hit_id_agreement']}")
    # FP64 re-intersection of every closest hit of one frame (outside the timed region).
    line["fp64_verify"] = fp64_all(rq, scene, frame, args.layout)
    log(f"FP64 verification: {line['fp64_verify']}")
    line["per_launch_ms"] = {k: round(v, 4) for k, v in per.items()}
    line["scene_build"] = build
    if args.config != "c1" and not args.no_extra_c1:
        line["c1"] = run_extra(rq, "c1", layouts, args)
        log(f"c1: {line['c1']}")
    if not args.no_reflection:
        line["reflection_rs"] = run_reflection(rq, kind, meshes, tris, frame)
        log(f"reflection rays (rs_trace, one launch of primary + GI + shadow): {line['reflection_rs']}")
    if not args.no_serving:
        line["serving_c4"], line["serving_1080p_sgi"] = run_serving_forms(rq, args.layout)
    if not args.no_gi_cache:
        line["gi_cache_rays"] = run_gi_cache(rq, scene, args.layout)
        log(f"GI cache ray stages: {line['gi_cache_rays']}")
    if not args.no_refit:
        line["refit_1m"] = run_refit(rq)
        log(f"mesh refit (1M triangles): {line['refit_1m']}")
    if args.batching:
        line["batching_c3"] = run_batching(rq, args.layout)
        log(f"batching: {line['batching_c3']}")
    print(json.dumps(line), flush=True)


def run_extra(rq, cfg, layouts, args):
    """Another configuration's frame, timed the same way (its own scene, tuning and rays)."""
    scene, _ = build_scene(rq, cfg, compare_builds=False)
    scene.upload(tuple(set(layouts)))
    stream = rq.Stream()
    frame, rays, results, tuned = time_config(rq, scene, cfg, layouts, args, stream)
    ms, _, per = results[layouts[0]]
    dom, dom_rays, stats = dominant_launch(scene, frame, layouts[0], per)
    out = {"workload": CONFIGS[cfg][6], "layout": layouts[0], "value": round(rays / (ms * 1e-3) / 1e6, 2),
           "ms_per_step": round(ms, 4), "rays_per_step": rays, "tuning": tuned.get(layouts[0]),
           "bytes_per_ray": round(stats["bytes"] / max(dom_rays, 1), 1),
           "roofline_frac": round(stats["bytes"] / (per[dom] * 1e-3) / HBM_PEAK, 4)}
    if "plain" in results and layouts[0] != "plain":
        pm = results["plain"][0]
        out["plain_f32"] = round(rays / (pm * 1e-3) / 1e6, 2)
        out["quantized_speedup"] = round(pm / ms, 3)
    del frame, scene
    return out


if __name__ == "__main__":
    main()
