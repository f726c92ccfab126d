// Serving loop: a stream of render-query jobs, admitted by an HBM budget and a
// latency budget, traced with continuous batching. No counterpart in the
// reference, whose renderer traces once per frame on one queue
// (src/gi_cache.rs:5150-5190); its own admission is a bounded wait
// (src/engine/mod.rs:1247-1262).
//
// A job (scene, camera seed, resolution, ray mix) is two traces: its primary
// rays (closest hit), then the sun-shadow rays (any hit) and GI rays (closest
// hit) generated from those hits. The loop runs one persistent launch at a
// time (rq_trace_batches, many scenes and batches in one grid): each launch
// carries the primary rays of the jobs admitted since the previous one and the
// secondary rays of the jobs whose primary rays that previous launch traced.
// Jobs arriving while a launch runs join the next one, so every launch mixes
// jobs and the GPU runs one full grid instead of competing ones. After a launch
// the device writes the new jobs' secondary rays and their counts (copied to
// pinned host memory, which decides the next launch's batches) and folds the
// finished jobs' outputs into 64-bit digests (position-keyed, order-free sums),
// so a job's result can be compared alone and in the mix. Launches trace without
// the CW8 node-step speculation, so no job's records depend on its neighbours.
//
// Admission, first come first served: a job is admitted once it has arrived, a
// slot (jobs in flight) is free and its buffers fit in the HBM arena (first
// fit). With a latency budget, each job is due at its arrival plus the budget.
// A job joins the next launch N with its primary rays and finishes with N + 1;
// the jobs whose secondary rays N carries finish with N. Launch times are
// predicted as a fixed drain cost plus rays over the measured throughput (a
// running average of the completed launches, secondary rays from a running hit
// fraction), with a margin, from the end of the running launch. A job that
// would finish after it is due is refused (RQ_ERR_DEADLINE), on arrival or
// while it waits; a job that would make an admitted job late waits (and the
// jobs behind it with it). Primary rays are generated in 8x8 screen-tile order
// when the job's size is a multiple of 8 (positions are ray indices).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "internal.h"

namespace rq {
namespace serve {

__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// Sum over i < min(*count, max_count) of mix(tag, i, words of record i).
__global__ void digest_kernel(const uint32_t* data, uint32_t words_per, const uint32_t* count_dev,
                              uint32_t max_count, unsigned long long tag, unsigned long long* out) {
    uint32_t n = count_dev ? min(*count_dev, max_count) : max_count;
    unsigned long long acc = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        unsigned long long h = mix64(tag ^ (i * 0x9E3779B97F4A7C15ull));
        for (uint32_t k = 0; k < words_per; ++k) h = mix64(h ^ data[i * words_per + k]);
        acc += h;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, acc);
}

// The job's secondary ray total, for its Mrays/s.
__global__ void count_kernel(const uint32_t* counts, uint32_t n, unsigned long long* out) {
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (uint32_t k = 0; k < n; ++k) s += counts[k];
        *out = s;
    }
}

struct Arena {
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    void init(size_t s) {
        size = s;
        free_.clear();
        free_[0] = s;
    }
    bool alloc(size_t n, size_t& at) {
        n = (n + 4095) & ~(size_t)4095;
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                at = it->first;
                size_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[at + n] = rest;
                return true;
            }
        return false;
    }
    void release(size_t at, size_t n) {
        n = (n + 4095) & ~(size_t)4095;
        auto it = free_.emplace(at, n).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
};

size_t job_bytes(const rq_job& j) {
    size_t n = (size_t)j.width * j.height;
    size_t b = n * (32 + 48 + 32) + ((n + 31) / 32) * 4 + (size_t)j.n_gi * n * (32 + 48) + 4096;
    return (b + 4095) & ~(size_t)4095;
}

// One job's progress through the loop.
enum Phase : int { WAITING = 0, PRIMARY = 1, SECONDARY_READY = 2, SECONDARY = 3, DONE = 4 };

struct JobState {
    Phase phase = WAITING;
    size_t at = 0, bytes = 0;         // arena range
    uint32_t counts[2] = {0, 0};      // shadow and GI rays, once known
    int primary_launch = -1;          // index of the launch that traced its primary rays
};

// One launch in flight: its jobs and its events.
struct Launch {
    hipEvent_t start = nullptr, end = nullptr;
    std::vector<uint32_t> primary, secondary;  // jobs whose primary / secondary rays it traces
    uint64_t rays = 0;
    double submit_s = 0.0, est_ms = 0.0;
    bool busy = false;
};

// Buffers of a job inside the arena.
struct JobBufs {
    rq_ray* prim;
    rq_hit* hits;
    rq_ray* shadow;
    uint32_t* vis;
    rq_ray* gi;
    rq_hit* gi_hits;
    uint32_t* counts;
};

JobBufs bufs(char* arena, const rq_job& j, const JobState& s) {
    const size_t n = (size_t)j.width * j.height;
    char* b = arena + s.at;
    JobBufs o;
    o.prim = reinterpret_cast<rq_ray*>(b);
    o.hits = reinterpret_cast<rq_hit*>(b + n * 32);
    o.shadow = reinterpret_cast<rq_ray*>(b + n * 80);
    o.vis = reinterpret_cast<uint32_t*>(b + n * 112);
    const size_t vis_bytes = ((n + 31) / 32) * 4;
    char* gi_base = b + n * 112 + ((vis_bytes + 255) & ~(size_t)255);
    o.gi = reinterpret_cast<rq_ray*>(gi_base);
    o.gi_hits = reinterpret_cast<rq_hit*>(gi_base + (size_t)j.n_gi * n * 32);
    o.counts = reinterpret_cast<uint32_t*>(b + s.bytes - 256);
    return o;
}

// The loop's estimate of a launch's duration: a fixed drain plus rays over the
// running throughput.
struct Model {
    double drain_ms = 0.5;       // measured fixed cost of a persistent launch (the drain)
    double rays_per_ms = 1.0e6;  // replaced by a running average of completed launches
    std::vector<double> hit_fraction;  // per scene: secondary rays per primary ray and ray kind, running average
    bool measured = false;
    // with a margin for the estimate's error and the host's turnaround between launches
    double ms(uint64_t rays) const { return 1.2 * (drain_ms + (double)rays / rays_per_ms) + 0.3; }
    // Launches of 1M rays or more update the throughput, their whole duration counted (the drain
    // included: an estimate that errs long); shorter ones are mostly drain and would not.
    void launch_done(uint64_t rays, double ms) {
        if (rays < (1u << 20) || ms <= 0.0) return;
        const double r = (double)rays / ms;
        rays_per_ms = measured ? 0.7 * rays_per_ms + 0.3 * r : r;
        measured = true;
    }
    double secondary(const rq_job& j) const {
        return (double)j.width * j.height * (1 + j.n_gi) * hit_fraction[j.scene];
    }
};

}  // namespace serve
}  // namespace rq

using namespace rq;

static double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

extern "C" rq_status rq_serve(rq_scene* const* scenes, uint32_t scene_count, const rq_serve_config* cfg,
                              const rq_job* jobs, uint32_t job_count, rq_job_result* results) {
    if (!scenes || !cfg || (!jobs && job_count) || (!results && job_count) || cfg->slots == 0 || cfg->slots > 256 ||
        (int)cfg->layout < 0 || (int)cfg->layout > 2 || scene_count == 0) {
        set_error("rq_serve: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    for (uint32_t i = 0; i < job_count; ++i)
        if (jobs[i].scene >= scene_count || !scenes[jobs[i].scene] || scenes[jobs[i].scene]->kind != RQ_SCENE_GI ||
            jobs[i].n_gi > 1 || jobs[i].width == 0 || jobs[i].height == 0 ||
            (uint64_t)jobs[i].width * jobs[i].height >= (1ull << 31) || !scenes[jobs[i].scene]->dev[cfg->layout].ready) {
            set_error("rq_serve: job " + std::to_string(i) + " is malformed or names a scene without the layout");
            return RQ_ERR_INVALID_INPUT;
        }
    char* arena_base = nullptr;
    if (hipMalloc(&arena_base, cfg->hbm_budget) != hipSuccess) {
        set_error("rq_serve: cannot allocate the HBM budget arena");
        return RQ_ERR_OUT_OF_MEMORY;
    }
    unsigned long long* digests = nullptr;  // per job
    const size_t dbytes = sizeof(unsigned long long) * std::max<uint32_t>(job_count, 1);
    uint32_t* counts_host = nullptr;  // pinned: 2 words a job (shadow, GI ray counts)
    hipStream_t stream = nullptr;
    rq_status status = RQ_OK;
    if (hipMalloc(&digests, dbytes) != hipSuccess || hipMemset(digests, 0, dbytes) != hipSuccess ||
        hipHostMalloc(&counts_host, 8ull * std::max<uint32_t>(job_count, 1), hipHostMallocDefault) != hipSuccess ||
        hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("rq_serve: cannot allocate digests, pinned counts or the stream");
        status = RQ_ERR_OUT_OF_MEMORY;
    }
    serve::Launch launch;
    if (status == RQ_OK && hipEventCreate(&launch.end) != hipSuccess) {
        set_error("rq_serve: event creation failed");
        status = RQ_ERR_DEVICE;
    }
    uint32_t max_rays = 1;
    for (uint32_t i = 0; i < job_count; ++i) max_rays = std::max(max_rays, jobs[i].width * jobs[i].height);
    if (status == RQ_OK) status = rq_stream_reserve(stream, max_rays);
    serve::Arena arena;
    arena.init(cfg->hbm_budget);
    serve::Model model;
    model.hit_fraction.assign(scene_count, 0.8);
    std::vector<serve::JobState> js(job_count);
    std::vector<double> launch_start_s;  // host clock at each launch's submission
    std::vector<hipEvent_t> starts;      // each launch's start event (kept for the jobs' GPU times)
    std::vector<uint32_t> order(job_count);
    for (uint32_t i = 0; i < job_count; ++i) {
        order[i] = i;
        std::memset(&results[i], 0, sizeof(rq_job_result));
        results[i].arrival_s = jobs[i].arrival_s;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return jobs[a].arrival_s < jobs[b].arrival_s; });
    std::deque<uint32_t> waiting;      // arrived, not admitted (arrival order)
    size_t next_arrival = 0;
    std::vector<uint32_t> new_jobs;    // admitted, primary rays not yet launched
    std::vector<uint32_t> ready_jobs;  // secondary rays known, not yet launched
    uint32_t in_flight = 0, done = 0;
    const double budget_s = cfg->latency_budget_us * 1e-6;
    auto t0 = std::chrono::steady_clock::now();

    auto finish = [&](uint32_t ji, double t) {
        results[ji].end_s = t;
        js[ji].phase = serve::DONE;
        arena.release(js[ji].at, js[ji].bytes);
        in_flight--;
        done++;
    };
    auto refuse = [&](uint32_t ji, double t, uint32_t why) {
        results[ji].status = why;
        results[ji].start_s = results[ji].end_s = t;
        js[ji].phase = serve::DONE;
        done++;
    };
    // Latency admission of job ji into the next launch N (its primary rays), its secondary rays then
    // in launch N + 1. N also holds the ready jobs' and the running launch's jobs' secondary rays
    // (those jobs finish with N) and the other new jobs' primary rays (they finish with N + 1).
    // 0: admit; 1: refuse (ji itself cannot finish in time); 2: wait (admitting it now would make a
    // job already admitted late).
    auto admission = [&](uint32_t ji, double t) -> int {
        const double start = launch.busy ? std::max(t, launch.submit_s + launch.est_ms * 1e-3) : t;
        uint64_t n0 = (uint64_t)jobs[ji].width * jobs[ji].height;
        double n1 = model.secondary(jobs[ji]);
        double due0 = 1e300, due1 = jobs[ji].arrival_s + budget_s;
        for (uint32_t k : new_jobs) {
            n0 += (uint64_t)jobs[k].width * jobs[k].height;
            n1 += model.secondary(jobs[k]);
            due1 = std::min(due1, jobs[k].arrival_s + budget_s);
        }
        for (uint32_t k : ready_jobs) {
            n0 += (uint64_t)js[k].counts[0] + js[k].counts[1];
            due0 = std::min(due0, jobs[k].arrival_s + budget_s);
        }
        if (launch.busy)
            for (uint32_t k : launch.primary) {
                n0 += (uint64_t)model.secondary(jobs[k]);
                due0 = std::min(due0, jobs[k].arrival_s + budget_s);
            }
        const double end0 = start + model.ms(n0) * 1e-3, end1 = end0 + model.ms((uint64_t)n1) * 1e-3;
        if (end1 > jobs[ji].arrival_s + budget_s) return 1;
        return end0 > due0 || end1 > due1 ? 2 : 0;
    };

    while (status == RQ_OK && done < job_count) {
        bool progressed = false;
        double t = now_s(t0);
        // 1. the running launch completed: its secondary jobs are done, its primary jobs' counts are in
        if (launch.busy) {
            const hipError_t q = hipEventQuery(launch.end);
            if (q != hipErrorNotReady) {
                if (q != hipSuccess) {
                    set_error(std::string("rq_serve: launch failed: ") + hipGetErrorString(q));
                    status = RQ_ERR_DEVICE;
                    break;
                }
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, launch.start, launch.end);
                model.launch_done(launch.rays, ms);
                t = now_s(t0);
                for (uint32_t ji : launch.secondary) {
                    float job_ms = 0.0f;
                    (void)hipEventElapsedTime(&job_ms, starts[js[ji].primary_launch], launch.end);
                    results[ji].gpu_ms = job_ms;
                    finish(ji, t);
                }
                for (uint32_t ji : launch.primary) {
                    js[ji].counts[0] = counts_host[2 * ji];
                    js[ji].counts[1] = jobs[ji].n_gi ? counts_host[2 * ji + 1] : 0u;
                    const uint64_t sec = (uint64_t)js[ji].counts[0] + js[ji].counts[1];
                    results[ji].secondary = sec;
                    const double n = (double)jobs[ji].width * jobs[ji].height * (1 + jobs[ji].n_gi);
                    double& hf = model.hit_fraction[jobs[ji].scene];
                    hf = 0.7 * hf + 0.3 * (double)sec / n;
                    if (sec == 0) {  // nothing hit: the primary records are the whole job
                        results[ji].gpu_ms = ms;
                        finish(ji, t);
                    } else {
                        js[ji].phase = serve::SECONDARY_READY;
                        ready_jobs.push_back(ji);
                    }
                }
                launch.busy = false;
                progressed = true;
            }
        }
        // 2. arrivals, then admission in arrival order
        while (next_arrival < order.size() && jobs[order[next_arrival]].arrival_s <= t) {
            waiting.push_back(order[next_arrival++]);
            progressed = true;
        }
        while (!waiting.empty()) {
            const uint32_t ji = waiting.front();
            const size_t need = serve::job_bytes(jobs[ji]);
            if (need > cfg->hbm_budget) {
                refuse(ji, t, RQ_ERR_BUDGET);
                waiting.pop_front();
                progressed = true;
                continue;
            }
            if (budget_s > 0.0) {
                const int a = admission(ji, t);
                if (a == 1) {
                    refuse(ji, t, RQ_ERR_DEADLINE);
                    waiting.pop_front();
                    progressed = true;
                    continue;
                }
                if (a == 2) break;  // first come, first served: the jobs behind it wait too
            }
            size_t at = 0;
            if (in_flight >= cfg->slots || !arena.alloc(need, at)) break;
            js[ji].at = at;
            js[ji].bytes = (need + 4095) & ~(size_t)4095;
            js[ji].phase = serve::PRIMARY;
            results[ji].start_s = t;
            results[ji].primary = (uint64_t)jobs[ji].width * jobs[ji].height;
            in_flight++;
            new_jobs.push_back(ji);
            waiting.pop_front();
            progressed = true;
        }
        // 3. the next launch: the new jobs' primary rays and the ready jobs' secondary rays
        if (!launch.busy && (!new_jobs.empty() || !ready_jobs.empty())) {
            std::vector<rq_batch> batches;
            std::vector<rq_scene*> used;
            auto scene_index = [&](uint32_t s) -> uint32_t {
                for (uint32_t k = 0; k < used.size(); ++k)
                    if (used[k] == scenes[s]) return k;
                used.push_back(scenes[s]);
                return (uint32_t)used.size() - 1;
            };
            launch.rays = 0;
            for (uint32_t ji : new_jobs) {
                const rq_job& j = jobs[ji];
                const serve::JobBufs b = serve::bufs(arena_base, j, js[ji]);
                float inv[16];
                if ((status = rq_gen_camera(j.scene_kind, j.seed, j.width, j.height, inv)) != RQ_OK) break;
                const bool tiled = j.width % 8 == 0 && j.height % 8 == 0;
                if ((status = tiled ? rq_gen_primary_tiled(inv, j.width, j.height, b.prim, stream)
                                    : rq_gen_primary(inv, j.width, j.height, b.prim, stream)) != RQ_OK)
                    break;
                rq_batch q{};
                q.scene = scene_index(j.scene);
                q.mode = RQ_MODE_CLOSEST;
                q.count = j.width * j.height;
                q.rays = b.prim;
                q.hits = b.hits;
                batches.push_back(q);
                launch.rays += q.count;
            }
            for (uint32_t ji : ready_jobs) {
                const rq_job& j = jobs[ji];
                const serve::JobBufs b = serve::bufs(arena_base, j, js[ji]);
                rq_batch q{};
                q.scene = scene_index(j.scene);
                q.mode = RQ_MODE_ANY;
                q.count = js[ji].counts[0];
                q.rays = b.shadow;
                q.vis_bits = b.vis;
                if (q.count) batches.push_back(q);
                if (js[ji].counts[1]) {
                    rq_batch g{};
                    g.scene = q.scene;
                    g.mode = RQ_MODE_CLOSEST;
                    g.count = js[ji].counts[1];
                    g.rays = b.gi;
                    g.hits = b.gi_hits;
                    batches.push_back(g);
                }
                launch.rays += (uint64_t)js[ji].counts[0] + js[ji].counts[1];
            }
            if (status != RQ_OK) break;
            // a start event per launch: each job's GPU time runs from its primary launch's start
            if (hipEventCreate(&launch.start) != hipSuccess) {
                set_error("rq_serve: event creation failed");
                status = RQ_ERR_DEVICE;
                break;
            }
            starts.push_back(launch.start);
            const int index = (int)starts.size() - 1;
            (void)hipEventRecord(launch.start, stream);
            status = trace_batches_impl(used.data(), (uint32_t)used.size(), cfg->layout, batches.data(),
                                        (uint32_t)batches.size(), stream, 0);
            if (status != RQ_OK) break;
            for (uint32_t ji : new_jobs) {
                const rq_job& j = jobs[ji];
                const serve::JobBufs b = serve::bufs(arena_base, j, js[ji]);
                const uint32_t n = j.width * j.height;
                if ((status = rq_gen_secondary(b.prim, b.hits, n, cfg->sun, (uint32_t)j.seed, j.n_gi, b.shadow, b.gi,
                                               b.counts, stream)) != RQ_OK)
                    break;
                (void)hipMemcpyAsync(counts_host + 2ull * ji, b.counts, 4ull * (1 + j.n_gi), hipMemcpyDeviceToHost,
                                     stream);
                hipLaunchKernelGGL(serve::digest_kernel, dim3(1024), dim3(256), 0, stream,
                                   reinterpret_cast<const uint32_t*>(b.hits), 12u, nullptr, n, 1ull, digests + ji);
                js[ji].primary_launch = index;
            }
            if (status != RQ_OK) break;
            for (uint32_t ji : ready_jobs) {
                const rq_job& j = jobs[ji];
                const serve::JobBufs b = serve::bufs(arena_base, j, js[ji]);
                const uint32_t n = js[ji].counts[0];
                hipLaunchKernelGGL(serve::digest_kernel, dim3(256), dim3(256), 0, stream, b.vis, 1u, nullptr,
                                   (n + 31) / 32, 2ull, digests + ji);
                if (js[ji].counts[1])
                    hipLaunchKernelGGL(serve::digest_kernel, dim3(1024), dim3(256), 0, stream,
                                       reinterpret_cast<const uint32_t*>(b.gi_hits), 12u, nullptr, js[ji].counts[1],
                                       3ull, digests + ji);
                js[ji].phase = serve::SECONDARY;
            }
            if (hipEventRecord(launch.end, stream) != hipSuccess || hipGetLastError() != hipSuccess) {
                set_error("rq_serve: launch failed");
                status = RQ_ERR_DEVICE;
                break;
            }
            launch.primary = new_jobs;
            launch.secondary = ready_jobs;
            launch.submit_s = now_s(t0);
            launch.est_ms = model.ms(launch.rays);
            launch.busy = true;
            new_jobs.clear();
            ready_jobs.clear();
            progressed = true;
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    if (stream) (void)hipStreamSynchronize(stream);
    if (status == RQ_OK && job_count) {
        std::vector<unsigned long long> host(job_count);
        if (hipMemcpy(host.data(), digests, dbytes, hipMemcpyDeviceToHost) == hipSuccess)
            for (uint32_t i = 0; i < job_count; ++i)
                if (results[i].status == RQ_OK) results[i].digest = host[i];
    }
    for (hipEvent_t e : starts) (void)hipEventDestroy(e);
    if (launch.end) (void)hipEventDestroy(launch.end);
    if (stream) (void)rq_stream_destroy(stream);
    if (counts_host) (void)hipHostFree(counts_host);
    (void)hipFree(digests);
    (void)hipFree(arena_base);
    return status;
}
