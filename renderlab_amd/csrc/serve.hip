// Serving loop: a stream of render-query jobs, admitted by an HBM budget and
// run with continuous batching. No counterpart in the reference, whose
// renderer traces once per frame on one queue (src/gi_cache.rs:5150-5190).
//
// Each job (scene, camera seed, resolution, ray mix) runs its whole pipeline
// on one of `slots` HIP streams: primary rays, closest hits, sun-shadow and
// GI rays from the hits, any-hit shadow visibility, closest GI hits. The ray
// counts of the secondary traces are read on the device (rq_trace_indirect),
// so a job is enqueued in one go and jobs on different streams overlap on the
// GPU. A job is admitted when it has arrived, a slot is free and its buffers
// fit in the budget arena (first fit); completion is polled on events.
// Every job's outputs are folded on the device into a 64-bit digest
// (position-keyed, order-free sum), so a job's result can be compared alone
// and in the mix without copying frames back. Primary rays are generated in
// 8x8 screen-tile order when the job's size is a multiple of 8 (positions
// are ray indices).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <thread>

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

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, end = nullptr;
    int job = -1;
    size_t at = 0, bytes = 0;
};

}  // namespace serve
}  // namespace rq

using namespace rq;

static double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

extern "C" rq_status rq_serve(rq_scene* const* scenes, uint32_t scene_count, const rq_serve_config* cfg,
                              const rq_job* jobs, uint32_t job_count, rq_job_result* results) {
    if (!scenes || !cfg || (!jobs && job_count) || (!results && job_count) || cfg->slots == 0 || cfg->slots > 16 ||
        (int)cfg->layout < 0 || (int)cfg->layout > 2 || scene_count == 0) {
        set_error("rq_serve: bad argument");
        return RQ_ERR_INVALID_INPUT;
    }
    for (uint32_t i = 0; i < job_count; ++i)
        if (jobs[i].scene >= scene_count || !scenes[jobs[i].scene] || scenes[jobs[i].scene]->kind != RQ_SCENE_GI ||
            jobs[i].n_gi > 1 || jobs[i].width == 0 ||
            jobs[i].height == 0 || !scenes[jobs[i].scene]->dev[cfg->layout].ready) {
            set_error("rq_serve: job " + std::to_string(i) + " is malformed or names a scene without the layout");
            return RQ_ERR_INVALID_INPUT;
        }
    char* arena_base = nullptr;
    if (hipMalloc(&arena_base, cfg->hbm_budget) != hipSuccess) {
        set_error("rq_serve: cannot allocate the HBM budget arena");
        return RQ_ERR_OUT_OF_MEMORY;
    }
    unsigned long long* digests = nullptr;  // per job: digest, secondary ray count
    const size_t dbytes = 2 * sizeof(unsigned long long) * std::max<uint32_t>(job_count, 1);
    if (hipMalloc(&digests, dbytes) != hipSuccess || hipMemset(digests, 0, dbytes) != hipSuccess) {
        (void)hipFree(arena_base);
        set_error("rq_serve: cannot allocate digests");
        return RQ_ERR_OUT_OF_MEMORY;
    }
    serve::Arena arena;
    arena.init(cfg->hbm_budget);
    std::vector<serve::Slot> slots(cfg->slots);
    rq_status status = RQ_OK;
    for (auto& s : slots) {
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&s.start) != hipSuccess || hipEventCreate(&s.end) != hipSuccess) {
            set_error("rq_serve: stream/event creation failed");
            status = RQ_ERR_DEVICE;
        }
    }
    uint32_t max_rays = 1;
    for (uint32_t i = 0; i < job_count; ++i) max_rays = std::max(max_rays, jobs[i].width * jobs[i].height);
    for (auto& sl : slots)
        if (status == RQ_OK && sl.stream) status = rq_stream_reserve(sl.stream, max_rays);
    std::vector<uint32_t> order(job_count);
    for (uint32_t i = 0; i < job_count; ++i) {
        order[i] = i;
        std::memset(&results[i], 0, sizeof(rq_job_result));
        results[i].arrival_s = jobs[i].arrival_s;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return jobs[a].arrival_s < jobs[b].arrival_s; });
    std::deque<uint32_t> queue(order.begin(), order.end());
    uint32_t done = 0;
    auto t0 = std::chrono::steady_clock::now();

    auto launch = [&](serve::Slot& sl, uint32_t ji) -> rq_status {
        const rq_job& j = jobs[ji];
        rq_scene* sc = scenes[j.scene];
        size_t n = (size_t)j.width * j.height;
        char* b = arena_base + sl.at;
        rq_ray* prim = reinterpret_cast<rq_ray*>(b);
        rq_hit* hits = reinterpret_cast<rq_hit*>(b + n * 32);
        rq_ray* shadow = reinterpret_cast<rq_ray*>(b + n * 80);
        uint32_t* vis = reinterpret_cast<uint32_t*>(b + n * 112);
        size_t vis_bytes = ((n + 31) / 32) * 4;
        char* gi_base = b + n * 112 + ((vis_bytes + 255) & ~(size_t)255);
        rq_ray* gi = reinterpret_cast<rq_ray*>(gi_base);
        rq_hit* gi_hits = reinterpret_cast<rq_hit*>(gi_base + (size_t)j.n_gi * n * 32);
        uint32_t* counts = reinterpret_cast<uint32_t*>(b + sl.bytes - 256);
        float inv[16];
        rq_status st;
        if ((st = rq_gen_camera(j.scene_kind, j.seed, j.width, j.height, inv)) != RQ_OK) return st;
        void* s = sl.stream;
        (void)hipEventRecord(sl.start, sl.stream);
        // screen-tile order when the size allows (a wave's rays through one 8x8 tile)
        const bool tiled = j.width % 8 == 0 && j.height % 8 == 0;
        if ((st = tiled ? rq_gen_primary_tiled(inv, j.width, j.height, prim, s)
                        : rq_gen_primary(inv, j.width, j.height, prim, s)) != RQ_OK)
            return st;
        if ((st = rq_trace(sc, cfg->layout, prim, (uint32_t)n, RQ_MODE_CLOSEST, hits, nullptr, s)) != RQ_OK) return st;
        if ((st = rq_gen_secondary(prim, hits, (uint32_t)n, cfg->sun, (uint32_t)j.seed, j.n_gi, shadow, gi, counts, s)) !=
            RQ_OK)
            return st;
        if ((st = rq_trace_indirect(sc, cfg->layout, shadow, counts, (uint32_t)n, RQ_MODE_ANY, nullptr, vis, s)) != RQ_OK)
            return st;
        if (j.n_gi &&
            (st = rq_trace_indirect(sc, cfg->layout, gi, counts + 1, (uint32_t)n, RQ_MODE_CLOSEST, gi_hits, nullptr, s)) !=
                RQ_OK)
            return st;
        unsigned long long* d = digests + 2ull * ji;
        hipLaunchKernelGGL(serve::count_kernel, dim3(1), dim3(64), 0, sl.stream, counts, j.n_gi + 1u, d + 1);
        hipLaunchKernelGGL(serve::digest_kernel, dim3(1024), dim3(256), 0, sl.stream,
                           reinterpret_cast<const uint32_t*>(hits), 12u, nullptr, (uint32_t)n, 1ull, d);
        hipLaunchKernelGGL(serve::digest_kernel, dim3(256), dim3(256), 0, sl.stream, vis, 1u, nullptr,
                           (uint32_t)((n + 31) / 32), 2ull, d);
        if (j.n_gi)
            hipLaunchKernelGGL(serve::digest_kernel, dim3(1024), dim3(256), 0, sl.stream,
                               reinterpret_cast<const uint32_t*>(gi_hits), 12u, counts + 1, (uint32_t)n, 3ull, d);
        if (hipEventRecord(sl.end, sl.stream) != hipSuccess || hipGetLastError() != hipSuccess) {
            set_error("rq_serve: launch failed");
            return RQ_ERR_DEVICE;
        }
        return RQ_OK;
    };

    while (status == RQ_OK && done < job_count) {
        bool progressed = false;
        // retire finished jobs
        for (auto& sl : slots) {
            if (sl.job < 0) continue;
            hipError_t q = hipEventQuery(sl.end);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                set_error(std::string("rq_serve: job failed: ") + hipGetErrorString(q));
                status = RQ_ERR_DEVICE;
                break;
            }
            rq_job_result& r = results[sl.job];
            r.end_s = now_s(t0);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, sl.start, sl.end);
            r.gpu_ms = ms;
            arena.release(sl.at, sl.bytes);
            sl.job = -1;
            done++;
            progressed = true;
        }
        if (status != RQ_OK) break;
        // admit arrived jobs in arrival order while a slot and memory are free
        double t = now_s(t0);
        while (!queue.empty() && jobs[queue.front()].arrival_s <= t) {
            uint32_t ji = queue.front();
            size_t need = serve::job_bytes(jobs[ji]);
            if (need > cfg->hbm_budget) {
                results[ji].status = RQ_ERR_BUDGET;
                results[ji].start_s = results[ji].end_s = t;
                queue.pop_front();
                done++;
                progressed = true;
                continue;
            }
            serve::Slot* free_slot = nullptr;
            for (auto& sl : slots)
                if (sl.job < 0) {
                    free_slot = &sl;
                    break;
                }
            size_t at = 0;
            if (!free_slot || !arena.alloc(need, at)) break;
            free_slot->at = at;
            free_slot->bytes = (need + 4095) & ~(size_t)4095;
            free_slot->job = (int)ji;
            results[ji].start_s = t;
            {
                const rq_job& j = jobs[ji];
                size_t n = (size_t)j.width * j.height;
                results[ji].primary = n;
            }
            rq_status st = launch(*free_slot, ji);
            if (st != RQ_OK) {
                status = st;
                break;
            }
            queue.pop_front();
            progressed = true;
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    for (auto& sl : slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    }
    if (status == RQ_OK && job_count) {
        std::vector<unsigned long long> host(2ull * job_count);
        if (hipMemcpy(host.data(), digests, dbytes, hipMemcpyDeviceToHost) == hipSuccess)
            for (uint32_t i = 0; i < job_count; ++i) {
                results[i].digest = host[2ull * i];
                results[i].secondary = host[2ull * i + 1];
            }
    }
    for (auto& sl : slots) {
        if (sl.start) (void)hipEventDestroy(sl.start);
        if (sl.end) (void)hipEventDestroy(sl.end);
        if (sl.stream) (void)rq_stream_destroy(sl.stream);
    }
    (void)hipFree(digests);
    (void)hipFree(arena_base);
    return status;
}
