// server.hip -- request serving over resident scenes (crt_server_*).
//
// No reference counterpart (the reference renders one call at a time on the
// CPU).  Design:
//   * scenes stay resident in HBM (their BVH is built once, crt_scene_create);
//   * submitted requests wait in a FIFO; a request is ADMITTED when the
//     device working memory it needs fits in the remaining budget (requests
//     that could never fit are rejected at submission with CRT_ERR_ADMISSION);
//   * N worker threads, each with its own HIP stream (default priority) and
//     its own grow-only workspaces, loop: take the head request plus every
//     queued request that can share its launches (same scene, same kind,
//     same lights/bounces for renders) up to the batch limits and the memory
//     budget, run them, publish results, release the memory.  New requests
//     join the next batch as soon as a worker is free, while the other
//     workers' batches are still on the device: continuous batching.
//   * renders of a batch go through render_batch (one set of launches for all
//     cameras); the flash-LiDAR scans of a batch are enqueued back to back
//     with one synchronisation; flux requests run back to back.  Images and
//     scans return through the worker's pinned staging buffer.
// Every result is a pure function of its request (counter-based RNG keyed by
// the request seed), so it is the same alone and in any mix.
// `batching = 0` gives the plain baseline: one worker, one request per
// launch sequence, in arrival order.
#include "crt_internal.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace crt {
namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    uint64_t ticket;
    crt_request req;
    size_t mem;
    Clock::time_point submitted, done_at;
    bool done = false;
    int status = CRT_OK;
    std::string error;
    uint64_t rays = 0;
};

// Page-locked host staging, grow-only (powers of two), one per worker.
// Results go device -> staging by DMA, then into the request's own (pageable)
// buffer by a plain memcpy on the worker's thread: first-touch page faults
// of fresh output arrays are then taken outside the HIP runtime's pageable
// copy path, which the workers would otherwise share.
struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    template <typename T>
    T* get(size_t bytes) {
        if (bytes > cap) {
            size_t c = (size_t)1 << 20;
            while (c < bytes) c <<= 1;
            if (p) (void)hipHostFree(p);
            p = nullptr;
            cap = 0;
            CRT_HIP_CHECK(hipHostMalloc(&p, c, hipHostMallocDefault));
            cap = c;
        }
        return static_cast<T*>(p);
    }
};

size_t request_bytes(const crt_request& r) {
    const uint64_t px = (uint64_t)std::floor(r.camera.resolution[0]) * (uint64_t)std::floor(r.camera.resolution[1]);
    switch (r.kind) {
        case CRT_REQ_RENDER:
            return render_bytes_estimate(px, r.spp, (int)r.n_lights) + px * 16;
        case CRT_REQ_LIDAR:
            return (size_t)px * 24;   // per-beam outputs; sub-ray ranges stay in registers
        case CRT_REQ_FLUX: {
            const uint64_t n = (uint64_t)r.flux.grid * r.flux.grid;
            return (size_t)(n / LANE_BLOCK + 1) * 7 * 16 + 256;
        }
        default:
            return 0;
    }
}

// Renders batch together when they share the scene, the bounce count and
// the number of lights (every camera carries its own lights).
bool same_render_setup(const crt_request& a, const crt_request& b) {
    if (a.kind != CRT_REQ_RENDER || b.kind != CRT_REQ_RENDER) return false;
    return a.scene == b.scene && a.bounces == b.bounces && a.n_lights == b.n_lights;
}

}  // namespace
}  // namespace crt

struct crt_server {
    crt_server_opts opts{};
    std::vector<crt_scene*> scenes;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<crt::Job>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<crt::Job>> jobs;
    uint64_t next_ticket = 1;
    size_t reserved = 0;
    bool stop = false;
    int device = 0;
    std::vector<std::thread> workers;
    // statistics
    std::vector<double> latencies_ms;
    uint64_t completed = 0, rejected = 0, rays = 0, batches = 0, inflight = 0;
    bool have_first = false;
    crt::Clock::time_point first_submit, last_done;
};

namespace crt {
namespace {

void run_batch(crt_server* S, std::vector<std::shared_ptr<Job>>& batch, hipStream_t st, Workspace& ws,
               RenderWorkspace* rws, PinnedBuf& pin) {
    const crt_request& head = batch[0]->req;
    crt_scene* scene = S->scenes[head.scene];
    const auto pixels = [](const crt_request& r) {
        return (uint64_t)std::floor(r.camera.resolution[0]) * (uint64_t)std::floor(r.camera.resolution[1]);
    };
    if (head.kind == CRT_REQ_RENDER) {
        std::vector<crt_camera> cams;
        std::vector<int32_t> spp;
        std::vector<uint64_t> seeds;
        std::vector<crt_light> lights;
        uint64_t total = 0;
        for (auto& j : batch) {
            cams.push_back(j->req.camera);
            spp.push_back(j->req.spp);
            seeds.push_back(j->req.seed);
            for (uint32_t l = 0; l < head.n_lights; ++l) lights.push_back(j->req.lights[l]);
            total += pixels(j->req);
        }
        crt_render_opts o{};
        o.min_samples = 1;
        o.max_samples = 1;
        o.noise_threshold = 0.0;
        o.num_bounces = head.bounces;
        crt_batch_opts b = S->opts.render;
        // the batch's images (packed in camera order) stay on the device,
        // then come back through the pinned staging in one copy
        ws.a.reserve((size_t)total * 16);
        float* d_rgba = ws.a.as<float>();
        crt_timing tm{};
        render_batch(scene, (uint32_t)batch.size(), cams.data(), spp.data(), seeds.data(),
                     lights.empty() ? nullptr : lights.data(), head.n_lights, true, &o, &b, d_rgba, nullptr,
                     nullptr, CRT_OUT_DEVICE, st, &tm, rws);
        float* h = pin.get<float>((size_t)total * 16);
        CRT_HIP_CHECK(hipMemcpyAsync(h, d_rgba, (size_t)total * 16, hipMemcpyDeviceToHost, st));
        CRT_HIP_CHECK(hipStreamSynchronize(st));
        uint64_t off = 0;
        for (auto& j : batch) {
            const uint64_t px = pixels(j->req);
            if (j->req.rgba) std::memcpy(j->req.rgba, h + off * 4, px * 16);
            j->rays = (tm.rays_traced + tm.shadow_rays) * px / std::max<uint64_t>(total, 1);
            off += px;
        }
    } else if (head.kind == CRT_REQ_LIDAR) {
        // every scan of the batch is enqueued (device outputs single / first
        // / last, 3 x px doubles per request, back to back), one copy back
        uint64_t total = 0;
        for (auto& j : batch) total += pixels(j->req);
        ws.b.reserve((size_t)total * 24);
        double* d = ws.b.as<double>();
        uint64_t off = 0;
        for (auto& j : batch) {
            const crt_request& r = j->req;
            crt_flash_opts fo{};
            fo.beams = r.camera;
            fo.sub_rays = r.sub_rays;
            fo.sigma = r.sigma;
            fo.seed = r.seed;
            const uint64_t px = pixels(r);
            double* dr = d + 3 * off;
            flash_lidar_impl(scene, &fo, dr, dr + px, dr + 2 * px, nullptr, CRT_OUT_DEVICE, st, nullptr, ws);
            j->rays = px * (uint64_t)(r.sub_rays + 1);
            off += px;
        }
        double* h = pin.get<double>((size_t)total * 24);
        CRT_HIP_CHECK(hipMemcpyAsync(h, d, (size_t)total * 24, hipMemcpyDeviceToHost, st));
        CRT_HIP_CHECK(hipStreamSynchronize(st));
        off = 0;
        for (auto& j : batch) {
            const uint64_t px = pixels(j->req);
            if (j->req.lidar) std::memcpy(j->req.lidar, h + 3 * off, px * 24);
            off += px;
        }
    } else {
        for (auto& j : batch) {
            crt_request& r = j->req;
            crt_timing tm{};
            double f[3], t[3];
            radiation_pressure_impl(scene, &r.flux, f, t, nullptr, st, &tm, ws);
            if (r.force_torque) {
                std::memcpy(r.force_torque, f, sizeof(f));
                std::memcpy(r.force_torque + 3, t, sizeof(t));
            }
            j->rays = tm.rays_traced + tm.shadow_rays;
        }
    }
}

void worker_main(crt_server* S, int index) {
    (void)index;
    (void)hipSetDevice(S->device);
    hipStream_t st = nullptr;
    (void)hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    Workspace ws;
    std::unique_ptr<RenderWorkspace, void (*)(RenderWorkspace*)> rws(new_render_workspace(), delete_render_workspace);
    PinnedBuf pin;
    const bool batching = S->opts.batching != 0;
    while (true) {
        std::vector<std::shared_ptr<Job>> batch;
        size_t mem = 0;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] {
                return S->stop || (!S->queue.empty() && S->reserved + S->queue.front()->mem <= S->opts.memory_budget);
            });
            if (S->stop && S->queue.empty()) break;
            if (S->queue.empty()) continue;
            auto head = S->queue.front();
            S->queue.pop_front();
            batch.push_back(head);
            mem = head->mem;
            if (batching && head->req.kind == CRT_REQ_RENDER) {
                uint64_t paths = (uint64_t)std::floor(head->req.camera.resolution[0]) *
                                 (uint64_t)std::floor(head->req.camera.resolution[1]) * head->req.spp;
                for (auto it = S->queue.begin(); it != S->queue.end() && batch.size() < S->opts.max_batch;) {
                    const crt_request& r = (*it)->req;
                    const uint64_t p = (uint64_t)std::floor(r.camera.resolution[0]) *
                                       (uint64_t)std::floor(r.camera.resolution[1]) * r.spp;
                    if (same_render_setup(head->req, r) && paths + p <= S->opts.max_batch_paths &&
                        S->reserved + mem + (*it)->mem <= S->opts.memory_budget) {
                        paths += p;
                        mem += (*it)->mem;
                        batch.push_back(*it);
                        it = S->queue.erase(it);
                    } else {
                        ++it;
                    }
                }
            } else if (batching) {
                // LiDAR / flux requests of the same scene and kind ride along;
                // a LiDAR batch holds at most max_batch_paths rays (it shares
                // one device buffer and one pinned staging copy)
                const auto lidar_rays = [](const crt_request& r) {
                    return r.kind == CRT_REQ_LIDAR ? (uint64_t)std::floor(r.camera.resolution[0]) *
                                                         (uint64_t)std::floor(r.camera.resolution[1]) *
                                                         (uint64_t)(r.sub_rays + 1)
                                                   : 0u;
                };
                uint64_t rays = lidar_rays(head->req);
                for (auto it = S->queue.begin(); it != S->queue.end() && batch.size() < S->opts.max_batch;) {
                    const crt_request& r = (*it)->req;
                    if (r.kind == head->req.kind && r.scene == head->req.scene &&
                        (r.kind != CRT_REQ_LIDAR || rays + lidar_rays(r) <= S->opts.max_batch_paths) &&
                        S->reserved + mem + (*it)->mem <= S->opts.memory_budget) {
                        rays += lidar_rays(r);
                        mem += (*it)->mem;
                        batch.push_back(*it);
                        it = S->queue.erase(it);
                    } else {
                        ++it;
                    }
                }
            }
            S->reserved += mem;
            S->inflight += batch.size();
        }
        int status = CRT_OK;
        std::string err;
        try {
            run_batch(S, batch, st, ws, rws.get(), pin);
        } catch (const Error& e) {
            status = e.code;
            err = e.what();
        } catch (const std::exception& e) {
            status = CRT_ERR_HIP;
            err = e.what();
        }
        const auto now = Clock::now();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->reserved -= mem;
            S->inflight -= batch.size();
            ++S->batches;
            for (auto& j : batch) {
                j->done = true;
                j->done_at = now;
                j->status = status;
                j->error = err;
                ++S->completed;
                S->rays += j->rays;
                S->latencies_ms.push_back(std::chrono::duration<double, std::milli>(now - j->submitted).count());
            }
            S->last_done = now;
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
    if (st) (void)hipStreamDestroy(st);
}

}  // namespace
}  // namespace crt

using namespace crt;

extern "C" {

int crt_server_create(const crt_server_opts* opts, crt_server** out) {
    return guarded([&] {
        CRT_REQUIRE(opts && out, CRT_ERR_INVALID, "NULL argument");
        *out = nullptr;
        std::unique_ptr<crt_server> S(new crt_server());
        S->opts = *opts;
        if (S->opts.memory_budget == 0) S->opts.memory_budget = (uint64_t)16 << 30;
        if (S->opts.n_workers == 0) S->opts.n_workers = 3;
        if (S->opts.max_batch == 0) S->opts.max_batch = 64;
        if (S->opts.max_batch_paths == 0) S->opts.max_batch_paths = (uint64_t)1 << 23;
        if (!S->opts.batching) {
            S->opts.n_workers = 1;
            S->opts.max_batch = 1;
        }
        CRT_HIP_CHECK(hipGetDevice(&S->device));
        crt_server* raw = S.get();
        for (uint32_t k = 0; k < S->opts.n_workers; ++k) S->workers.emplace_back(worker_main, raw, (int)k);
        *out = S.release();
    });
}

int crt_server_add_scene(crt_server* S, crt_scene* scene, uint32_t* id) {
    return guarded([&] {
        CRT_REQUIRE(S && scene && id, CRT_ERR_INVALID, "NULL argument");
        std::lock_guard<std::mutex> lk(S->mu);
        S->scenes.push_back(scene);
        *id = (uint32_t)(S->scenes.size() - 1);
    });
}

int crt_server_submit(crt_server* S, const crt_request* req, uint64_t* ticket) {
    return guarded([&] {
        CRT_REQUIRE(S && req && ticket, CRT_ERR_INVALID, "NULL argument");
        CRT_REQUIRE(req->kind <= CRT_REQ_FLUX, CRT_ERR_TYPE, "unknown request kind");
        auto j = std::make_shared<Job>();
        j->req = *req;
        j->mem = request_bytes(*req);
        j->submitted = Clock::now();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            CRT_REQUIRE(req->scene < S->scenes.size(), CRT_ERR_INVALID, "unknown scene id");
            if (req->kind != CRT_REQ_FLUX) check_camera_args(&req->camera);
            if (req->kind == CRT_REQ_RENDER)
                CRT_REQUIRE(req->spp >= 1 && req->n_lights <= 2, CRT_ERR_INVALID, "bad render request");
            if (j->mem > S->opts.memory_budget) {
                ++S->rejected;
                throw Error(CRT_ERR_ADMISSION, "request needs more device memory than the server budget");
            }
            j->ticket = S->next_ticket++;
            if (!S->have_first) {
                S->first_submit = j->submitted;
                S->have_first = true;
            }
            S->jobs[j->ticket] = j;
            S->queue.push_back(j);
            *ticket = j->ticket;
        }
        S->cv_work.notify_one();
    });
}

int crt_server_wait(crt_server* S, uint64_t ticket, double* latency_ms) {
    return guarded([&] {
        CRT_REQUIRE(S != nullptr, CRT_ERR_INVALID, "NULL server");
        std::shared_ptr<Job> j;
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->jobs.find(ticket);
        CRT_REQUIRE(it != S->jobs.end(), CRT_ERR_INVALID, "unknown ticket");
        j = it->second;
        S->cv_done.wait(lk, [&] { return j->done; });
        if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(j->done_at - j->submitted).count();
        S->jobs.erase(ticket);
        if (j->status != CRT_OK) throw Error(j->status, j->error);
    });
}

int crt_server_drain(crt_server* S) {
    return guarded([&] {
        CRT_REQUIRE(S != nullptr, CRT_ERR_INVALID, "NULL server");
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_done.wait(lk, [&] { return S->queue.empty() && S->inflight == 0; });
    });
}

int crt_server_stats_get(crt_server* S, crt_server_stats* out) {
    return guarded([&] {
        CRT_REQUIRE(S && out, CRT_ERR_INVALID, "NULL argument");
        std::lock_guard<std::mutex> lk(S->mu);
        std::memset(out, 0, sizeof(*out));
        out->completed = S->completed;
        out->rejected = S->rejected;
        out->rays = S->rays;
        out->batches = S->batches;
        std::vector<double> l = S->latencies_ms;
        if (!l.empty()) {
            std::sort(l.begin(), l.end());
            auto pct = [&](double p) {
                const double x = p * (double)(l.size() - 1);
                const size_t i = (size_t)x;
                const double f = x - (double)i;
                return i + 1 < l.size() ? l[i] * (1 - f) + l[i + 1] * f : l[i];
            };
            out->p50_ms = pct(0.50);
            out->p99_ms = pct(0.99);
            double s = 0;
            for (double v : l) s += v;
            out->mean_ms = s / (double)l.size();
        }
        if (S->have_first && S->completed > 0) {
            out->wall_s = std::chrono::duration<double>(S->last_done - S->first_submit).count();
            out->mrays_per_s = out->wall_s > 0 ? (double)S->rays / out->wall_s / 1e6 : 0.0;
        }
    });
}

void crt_server_destroy(crt_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& t : S->workers) t.join();
    delete S;
}

}  // extern "C"
