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
//     cameras); the flash-LiDAR scans of a batch run as one launch over a
//     segment table (flash_lidar_multi), its flux grids as one launch
//     sequence (radiation_pressure_multi); one synchronisation per batch.
//     Images, scans and forces return through the worker's pinned staging.
//   * every option is validated at submission; a batch that still fails at
//     run time is rerun request by request, so an error reaches only its own
//     request; the scene of a request is resolved at submission.
//   * the memory budget covers the admitted requests and the device
//     workspaces the workers keep between batches (a worker frees its
//     workspaces when keeping them would not let the next batch fit).
// Every result is a pure function of its request (counter-based RNG keyed by
// the request seed), so it is the same alone and in any mix.
// `batching = 0` gives the plain baseline: one worker, one request per
// launch sequence, in arrival order.
#include "crt_internal.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
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
    crt_scene* scene = nullptr;       // resolved at submission
    std::vector<crt_light> lights;    // the request's lights (copied at submission)
    size_t mem;
    Clock::time_point submitted, done_at;
    bool done = false;
    int status = CRT_OK;
    std::string error;
    uint64_t rays = 0;
    uint32_t passed_over = 0;         // times a later request was taken first (scene affinity)
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
            return (size_t)px * 24 + 512;   // per-beam outputs; sub-ray ranges stay in registers
        case CRT_REQ_FLUX: {
            const uint64_t n = (uint64_t)r.flux.grid * r.flux.grid;
            return (size_t)(n / LANE_BLOCK + 1) * 7 * 8 + 7 * 64 * 16 + 512;
        }
        default:
            return 0;
    }
}

// Everything a request's batch would check, at submission (raises).
void validate_request(const crt_request& r, uint32_t n_scenes) {
    CRT_REQUIRE(r.kind <= CRT_REQ_FLUX, CRT_ERR_TYPE, "unknown request kind");
    CRT_REQUIRE(r.scene < n_scenes, CRT_ERR_INVALID, "unknown scene id");
    if (r.kind == CRT_REQ_RENDER) {
        check_camera_args(&r.camera);
        CRT_REQUIRE(r.spp >= 1 && r.bounces >= 0, CRT_ERR_INVALID, "bad render request (spp >= 1, bounces >= 0)");
        CRT_REQUIRE(r.light_list != nullptr || r.n_lights <= 2, CRT_ERR_INVALID,
                    "more than 2 lights need light_list");
        CRT_REQUIRE(r.n_lights <= (1u << 20), CRT_ERR_INVALID, "too many lights");
        const crt_light* L = r.light_list ? r.light_list : r.lights;
        for (uint32_t l = 0; l < r.n_lights; ++l) check_light(L[l]);
        const uint64_t px = (uint64_t)std::floor(r.camera.resolution[0]) * (uint64_t)std::floor(r.camera.resolution[1]);
        CRT_REQUIRE(px < (1ull << 31), CRT_ERR_INVALID, "camera too large");
    } else if (r.kind == CRT_REQ_LIDAR) {
        crt_flash_opts fo{};
        fo.beams = r.camera;
        fo.sub_rays = r.sub_rays;
        fo.sigma = r.sigma;
        check_flash_opts(&fo);
    } else {
        check_flux_opts(&r.flux);
    }
}

// Renders batch together when they share the scene, the bounce count and
// the number of lights (every camera carries its own lights).
bool same_render_setup(const Job& a, const Job& b) {
    if (a.req.kind != CRT_REQ_RENDER || b.req.kind != CRT_REQ_RENDER) return false;
    return a.scene == b.scene && a.req.bounces == b.req.bounces && a.req.n_lights == b.req.n_lights;
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
    bool stop = false;
    int device = 0;
    std::vector<std::thread> workers;
    // per worker: its stream, device workspaces and pinned staging (another
    // worker may free the workspaces of an idle worker, under `mu`), the
    // device bytes they keep between batches and the reservation of the
    // batch it runs (admission counts the max of the two)
    struct Worker {
        hipStream_t st = nullptr;
        crt::Workspace ws;
        crt::RenderWorkspace* rws = nullptr;
        crt::PinnedBuf pin;
        ~Worker() {
            if (rws) crt::delete_render_workspace(rws);
            if (st) (void)hipStreamDestroy(st);
        }
    };
    std::vector<std::unique_ptr<Worker>> wstate;
    std::vector<size_t> held, running;
    std::vector<const crt_scene*> running_scene;   // per worker, null when idle
    // statistics
    std::vector<double> latencies_ms;
    uint64_t completed = 0, rejected = 0, cancelled = 0, rays = 0, batches = 0, inflight = 0;
    bool have_first = false;
    crt::Clock::time_point first_submit, last_done;
};

namespace crt {
namespace {

void run_batch(std::vector<std::shared_ptr<Job>>& batch, hipStream_t st, Workspace& ws, RenderWorkspace* rws,
               PinnedBuf& pin, const crt_batch_opts& ropts) {
    const crt_request& head = batch[0]->req;
    crt_scene* scene = batch[0]->scene;
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
            lights.insert(lights.end(), j->lights.begin(), j->lights.end());
            total += pixels(j->req);
        }
        crt_render_opts o{};
        o.min_samples = 1;
        o.max_samples = 1;
        o.noise_threshold = 0.0;
        o.num_bounces = head.bounces;
        crt_batch_opts b = ropts;
        // the batch's images (packed in camera order) stay on the device,
        // then come back through the pinned staging in one copy
        ws.a.reserve((size_t)total * 16);
        float* d_rgba = ws.a.as<float>();
        crt_timing tm{};
        const int rc = render_batch(scene, (uint32_t)batch.size(), cams.data(), spp.data(), seeds.data(),
                                    lights.empty() ? nullptr : lights.data(), head.n_lights, true, &o, &b, d_rgba,
                                    nullptr, nullptr, CRT_OUT_DEVICE, st, &tm, rws);
        if (rc != CRT_OK) throw Error(rc, g_last_error);
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
        // every scan of the batch in one launch (single / first / last, 3 x px
        // doubles per request, back to back), one copy back
        uint64_t total = 0;
        for (auto& j : batch) total += pixels(j->req);
        ws.b.reserve((size_t)total * 24);
        double* d = ws.b.as<double>();
        std::vector<crt_flash_opts> fo(batch.size());
        std::vector<double*> outs(3 * batch.size());
        uint64_t off = 0;
        for (size_t i = 0; i < batch.size(); ++i) {
            const crt_request& r = batch[i]->req;
            fo[i] = crt_flash_opts{};
            fo[i].beams = r.camera;
            fo[i].sub_rays = r.sub_rays;
            fo[i].sigma = r.sigma;
            fo[i].seed = r.seed;
            const uint64_t px = pixels(r);
            double* dr = d + 3 * off;
            outs[3 * i] = dr;
            outs[3 * i + 1] = dr + px;
            outs[3 * i + 2] = dr + 2 * px;
            batch[i]->rays = px * (uint64_t)(r.sub_rays + 1);
            off += px;
        }
        std::vector<uint8_t> table;
        flash_lidar_multi(scene, (uint32_t)batch.size(), fo.data(), outs.data(), st, ws, table);
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
        // every flux grid of the batch in one launch sequence, one copy back
        std::vector<crt_flux_opts> fo(batch.size());
        for (size_t i = 0; i < batch.size(); ++i) {
            fo[i] = batch[i]->req.flux;
            batch[i]->rays = (uint64_t)fo[i].grid * fo[i].grid;
        }
        ws.f.reserve(batch.size() * 7 * sizeof(double));
        std::vector<uint8_t> table;
        radiation_pressure_multi(scene, (uint32_t)batch.size(), fo.data(), ws.f.as<double>(), st, ws, table);
        double* h = pin.get<double>(batch.size() * 7 * sizeof(double));
        CRT_HIP_CHECK(hipMemcpyAsync(h, ws.f.p, batch.size() * 7 * sizeof(double), hipMemcpyDeviceToHost, st));
        CRT_HIP_CHECK(hipStreamSynchronize(st));
        for (size_t i = 0; i < batch.size(); ++i)
            if (batch[i]->req.force_torque) std::memcpy(batch[i]->req.force_torque, h + 7 * i, 6 * sizeof(double));
    }
}

size_t workspace_bytes(const Workspace& ws, const RenderWorkspace* rws) {
    return ws.a.bytes + ws.b.bytes + ws.c.bytes + ws.d.bytes + ws.e.bytes + ws.f.bytes + ws.g.bytes + ws.h.bytes +
           ws.counters.bytes + render_workspace_bytes(rws);
}

// Admission of `need` bytes on worker w (under S->mu): the other workers'
// footprints (a running one: max of what it keeps and its batch's
// reservation; an idle one: what it keeps) plus w's own.  Returns 0 (does
// not fit), 1 (fits keeping w's workspaces), 2 (fits once w frees them) or 3
// (fits once w and the idle workers free theirs).
int admit(const crt_server* S, int w, size_t need) {
    size_t busy = 0, idle = 0;
    for (size_t v = 0; v < S->held.size(); ++v) {
        if ((int)v == w) continue;
        if (S->running[v]) busy += std::max(S->held[v], S->running[v]);
        else idle += S->held[v];
    }
    const size_t B = S->opts.memory_budget;
    if (busy + idle + std::max(S->held[w], need) <= B) return 1;
    if (busy + idle + need <= B) return 2;
    if (busy + need <= B) return 3;
    return 0;
}

// The next batch's head (under S->mu), or queue.end() to wait.  The oldest
// request, unless it is on a scene no other worker is tracing and one of the
// next AFFINITY_WINDOW requests is on such a scene: taking that one keeps the
// concurrent batches on one BVH, which then stays in L2 / MALL instead of
// two large meshes evicting each other.  The oldest request is passed over
// at most AFFINITY_SKIPS times, and a request that does not fit the memory
// budget is never passed (it waits at the front, as in FIFO order).
constexpr int AFFINITY_WINDOW = 8;
constexpr uint32_t AFFINITY_SKIPS = 8;

std::deque<std::shared_ptr<Job>>::iterator pick_head(crt_server* S, int w) {
    auto front = S->queue.begin();
    if (front == S->queue.end() || admit(S, w, (*front)->mem) == 0) return S->queue.end();
    const auto traced = [&](const crt_scene* sc) {
        for (size_t v = 0; v < S->running_scene.size(); ++v)
            if ((int)v != w && S->running_scene[v] == sc) return true;
        return false;
    };
    bool others = false;
    for (size_t v = 0; v < S->running_scene.size(); ++v) others |= (int)v != w && S->running_scene[v] != nullptr;
    if (!others || (*front)->passed_over >= AFFINITY_SKIPS || traced((*front)->scene)) return front;
    int k = 0;
    for (auto it = std::next(front); it != S->queue.end() && k < AFFINITY_WINDOW; ++it, ++k)
        if (traced((*it)->scene) && admit(S, w, (*it)->mem) != 0) return it;
    return front;
}

void worker_main(crt_server* S, int index) {
    (void)hipSetDevice(S->device);
    crt_server::Worker& me = *S->wstate[index];
    const bool batching = S->opts.batching != 0;
    const auto pixels = [](const crt_request& r) {
        return (uint64_t)std::floor(r.camera.resolution[0]) * (uint64_t)std::floor(r.camera.resolution[1]);
    };
    while (true) {
        std::vector<std::shared_ptr<Job>> batch;
        std::vector<Workspace> drop_ws;
        std::vector<RenderWorkspace*> drop_rws;
        size_t mem = 0;
        if (!retired_bufs().p.empty()) {
            // buffers this worker's workspaces outgrew during its batches:
            // freed once it has no work, where waiting for the device costs
            // nothing
            bool idle;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                idle = pick_head(S, index) == S->queue.end();
            }
            if (idle) {
                free_retired();
                std::lock_guard<std::mutex> lk(S->mu);
                S->held[index] = workspace_bytes(me.ws, me.rws);
            }
        }
        {
            std::unique_lock<std::mutex> lk(S->mu);
            // on stop, queued requests still run (close drains the queue)
            S->cv_work.wait(lk, [&] {
                return (S->stop && S->queue.empty()) || pick_head(S, index) != S->queue.end();
            });
            if (S->stop && S->queue.empty()) break;
            const auto pick = pick_head(S, index);
            if (pick == S->queue.end()) continue;
            auto head = *pick;
            if (pick != S->queue.begin()) ++S->queue.front()->passed_over;
            S->queue.erase(pick);
            batch.push_back(head);
            mem = head->mem;
            if (batching) {
                // requests that share the head's launches join it, within the
                // batch limits and the memory budget
                const bool render = head->req.kind == CRT_REQ_RENDER;
                const auto work = [&](const Job& j) {
                    return render ? pixels(j.req) * (uint64_t)j.req.spp
                                  : (j.req.kind == CRT_REQ_LIDAR ? pixels(j.req) * (uint64_t)(j.req.sub_rays + 1) : 0u);
                };
                uint64_t amount = work(*head);
                for (auto it = S->queue.begin(); it != S->queue.end() && batch.size() < S->opts.max_batch;) {
                    const Job& j = **it;
                    const bool joins = render ? same_render_setup(*head, j)
                                              : (j.req.kind == head->req.kind && j.scene == head->scene);
                    if (joins && amount + work(j) <= S->opts.max_batch_paths && admit(S, index, mem + j.mem) != 0) {
                        amount += work(j);
                        mem += j.mem;
                        batch.push_back(*it);
                        it = S->queue.erase(it);
                    } else {
                        ++it;
                    }
                }
            }
            // free what this worker and, if needed, the idle ones keep (an
            // idle worker has nothing in flight on its stream); the buffers
            // are only taken here and released after the lock, because
            // hipFree may wait for the device
            const int a = admit(S, index, mem);
            const auto take = [&](crt_server::Worker& w, size_t v) {
                drop_ws.push_back(std::move(w.ws));
                w.ws = Workspace();
                drop_rws.push_back(w.rws);
                w.rws = new_render_workspace();
                S->held[v] = 0;
            };
            if (a >= 2) take(me, (size_t)index);
            if (a == 3)
                for (size_t v = 0; v < S->wstate.size(); ++v)
                    if ((int)v != index && S->running[v] == 0 && S->held[v] > 0) take(*S->wstate[v], v);
            S->running[index] = std::max<size_t>(mem, 1);
            S->running_scene[index] = head->scene;
            S->inflight += batch.size();
        }
        drop_ws.clear();
        for (RenderWorkspace* r : drop_rws) delete_render_workspace(r);
        if (!drop_rws.empty()) free_retired();   // the room is needed now
        // run the batch; if it fails, rerun its requests one by one so an
        // error reaches only its own request
        std::vector<int> status(batch.size(), CRT_OK);
        std::vector<std::string> err(batch.size());
        const auto run = [&](std::vector<std::shared_ptr<Job>>& b, size_t first) {
            try {
                run_batch(b, me.st, me.ws, me.rws, me.pin, S->opts.render);
                return true;
            } catch (const Error& e) {
                for (size_t i = 0; i < b.size(); ++i) { status[first + i] = e.code; err[first + i] = e.what(); }
            } catch (const std::exception& e) {
                for (size_t i = 0; i < b.size(); ++i) { status[first + i] = CRT_ERR_HIP; err[first + i] = e.what(); }
            }
            (void)hipGetLastError();
            (void)hipStreamSynchronize(me.st);
            return false;
        };
        retired_bufs().defer = true;
        if (!run(batch, 0) && batch.size() > 1)
            for (size_t i = 0; i < batch.size(); ++i) {
                std::vector<std::shared_ptr<Job>> one{batch[i]};
                status[i] = CRT_OK;
                err[i].clear();
                run(one, i);
            }
        retired_bufs().defer = false;
        const size_t keep = workspace_bytes(me.ws, me.rws) + retired_bufs().bytes;
        const auto now = Clock::now();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->running[index] = 0;
            S->running_scene[index] = nullptr;
            S->held[index] = keep;
            S->inflight -= batch.size();
            ++S->batches;
            for (size_t i = 0; i < batch.size(); ++i) {
                auto& j = batch[i];
                j->done = true;
                j->done_at = now;
                j->status = status[i];
                j->error = err[i];
                ++S->completed;
                if (status[i] == CRT_OK) S->rays += j->rays;
                S->latencies_ms.push_back(std::chrono::duration<double, std::milli>(now - j->submitted).count());
            }
            S->last_done = now;
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
    free_retired();
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
        check_batch_opts(S->opts.render);
        if (S->opts.memory_budget == 0) S->opts.memory_budget = (uint64_t)16 << 30;
        if (S->opts.n_workers == 0) S->opts.n_workers = 3;
        if (S->opts.max_batch == 0) S->opts.max_batch = 64;
        if (S->opts.max_batch_paths == 0) S->opts.max_batch_paths = (uint64_t)1 << 23;
        if (!S->opts.batching) {
            S->opts.n_workers = 1;
            S->opts.max_batch = 1;
        }
        CRT_HIP_CHECK(hipGetDevice(&S->device));
        for (uint32_t k = 0; k < S->opts.n_workers; ++k) {
            std::unique_ptr<crt_server::Worker> w(new crt_server::Worker());
            CRT_HIP_CHECK(hipStreamCreateWithFlags(&w->st, hipStreamNonBlocking));   // default priority
            w->rws = new_render_workspace();
            S->wstate.push_back(std::move(w));
        }
        S->held.assign(S->opts.n_workers, 0);
        S->running.assign(S->opts.n_workers, 0);
        S->running_scene.assign(S->opts.n_workers, nullptr);
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
        if (req->kind == CRT_REQ_RENDER && req->n_lights) {
            const crt_light* L = req->light_list ? req->light_list : req->lights;
            CRT_REQUIRE(req->light_list != nullptr || req->n_lights <= 2, CRT_ERR_INVALID,
                        "more than 2 lights need light_list");
            j->lights.assign(L, L + req->n_lights);
        }
        j->req.light_list = nullptr;
        j->mem = request_bytes(*req);
        j->submitted = Clock::now();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            validate_request(*req, (uint32_t)S->scenes.size());
            j->scene = S->scenes[req->scene];
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

int crt_server_wait_for(crt_server* S, uint64_t ticket, double timeout_ms, double* latency_ms) {
    return guarded([&] {
        CRT_REQUIRE(S != nullptr, CRT_ERR_INVALID, "NULL server");
        CRT_REQUIRE(timeout_ms >= 0.0, CRT_ERR_INVALID, "timeout_ms must be >= 0 (infinity: no limit)");
        std::shared_ptr<Job> j;
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->jobs.find(ticket);
        CRT_REQUIRE(it != S->jobs.end(), CRT_ERR_INVALID, "unknown ticket");
        j = it->second;
        if (std::isinf(timeout_ms)) {
            S->cv_done.wait(lk, [&] { return j->done; });
        } else if (!S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms),
                                        [&] { return j->done; })) {
            throw Error(CRT_ERR_TIMEOUT, "request not finished within the timeout");
        }
        if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(j->done_at - j->submitted).count();
        S->jobs.erase(ticket);
        if (j->status != CRT_OK) throw Error(j->status, j->error);
    });
}

int crt_server_wait(crt_server* S, uint64_t ticket, double* latency_ms) {
    return crt_server_wait_for(S, ticket, std::numeric_limits<double>::infinity(), latency_ms);
}

int crt_server_cancel(crt_server* S, uint64_t ticket) {
    return guarded([&] {
        CRT_REQUIRE(S != nullptr, CRT_ERR_INVALID, "NULL server");
        {
            std::lock_guard<std::mutex> lk(S->mu);
            auto it = S->jobs.find(ticket);
            CRT_REQUIRE(it != S->jobs.end(), CRT_ERR_INVALID, "unknown ticket");
            const std::shared_ptr<Job> j = it->second;
            CRT_REQUIRE(!j->done, CRT_ERR_STATE, "request already finished");
            auto q = std::find(S->queue.begin(), S->queue.end(), j);
            CRT_REQUIRE(q != S->queue.end(), CRT_ERR_STATE, "request is running");
            S->queue.erase(q);
            j->done = true;
            j->done_at = Clock::now();
            j->status = CRT_ERR_CANCELLED;
            j->error = "request cancelled before it ran";
            ++S->cancelled;
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();   // the queue's head changed
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
        out->cancelled = S->cancelled;
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

int crt_server_stats_reset(crt_server* S) {
    return guarded([&] {
        CRT_REQUIRE(S, CRT_ERR_INVALID, "NULL argument");
        std::lock_guard<std::mutex> lk(S->mu);
        CRT_REQUIRE(S->queue.empty() && S->inflight == 0, CRT_ERR_STATE,
                    "statistics are reset only while nothing is queued or running");
        S->completed = S->rejected = S->cancelled = S->rays = S->batches = 0;
        S->latencies_ms.clear();
        S->have_first = false;
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
