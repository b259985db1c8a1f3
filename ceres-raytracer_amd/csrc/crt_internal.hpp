// crt_internal.hpp -- host-side internals shared by the .hip translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <stdint.h>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <new>
#include <utility>
#include <vector>

#include "crt_device.hpp"
#include "../../include/crt_hip.h"

namespace crt {

// Error carried from deep inside the library to the C-ABI boundary, where it
// becomes a status code + crt_last_error() text.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CRT_HIP_CHECK(expr)                                                              \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            (void)hipGetLastError();                                                     \
            throw ::crt::Error(e_ == hipErrorOutOfMemory ? CRT_ERR_OOM : CRT_ERR_HIP,    \
                               std::string(#expr) + ": " + hipGetErrorString(e_));       \
        }                                                                                \
    } while (0)

#define CRT_REQUIRE(cond, code, msg)                     \
    do {                                                 \
        if (!(cond)) throw ::crt::Error((code), (msg));  \
    } while (0)

// Thread-local text of the last error, returned by crt_last_error().
inline thread_local std::string g_last_error;

// Run `f`, mapping exceptions to C-ABI status codes.
template <typename F>
int guarded(F&& f) {
    try {
        f();
        return CRT_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return CRT_ERR_OOM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return CRT_ERR_HIP;
    }
}

// Owning device buffer.
// Device buffers released on a thread that must not wait for the device
// (a server worker inside a batch: hipFree waits for the whole device and
// so stalls the other workers' streams).  While `defer` is set, released
// buffers are kept here and freed by free_retired() at a point where the
// wait costs nothing.
struct RetiredBufs {
    std::vector<void*> p;
    size_t bytes = 0;
    bool defer = false;
};
inline RetiredBufs& retired_bufs() {
    static thread_local RetiredBufs r;
    return r;
}
inline void free_retired() {
    RetiredBufs& r = retired_bufs();
    for (void* q : r.p) (void)hipFree(q);
    r.p.clear();
    r.bytes = 0;
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            RetiredBufs& r = retired_bufs();
            if (r.defer) {
                r.p.push_back(p);
                r.bytes += bytes;
            } else {
                (void)hipFree(p);
            }
        }
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b) {
        release();
        if (b == 0) return;
        CRT_HIP_CHECK(hipMalloc(&p, b));
        bytes = b;
    }
    // grow-only reallocation (contents not preserved).  A buffer that has to
    // grow takes at least 1.5 times its size: workspaces reused by calls of
    // varying sizes (request batches, repeated API calls) then reallocate a
    // few times only, and each reallocation's hipFree waits for the whole
    // device, stalling every other stream's work.
    void reserve(size_t b) {
        if (b > bytes) alloc(bytes ? std::max(b, bytes + bytes / 2) : b);
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

struct BuildResult {
    DevBuf nodes;        // wide nodes (256 B)
    DevBuf cnodes;       // the same nodes compressed (96 B)
    DevBuf tris;         // leaf-ordered triangles (3 float4 each)
    uint32_t n_nodes = 0;
    uint32_t n_leaves = 0;
    uint32_t max_depth = 0;
    uint32_t ploc_rounds = 0;   // 0: LBVH (Karras) tree
    double build_ms = 0.0;
    double sah_cost = 0.0;
    double sah_at_build = 0.0;        // SAH when the topology was last built
    double refit_ms = 0.0;            // device time of the last refit
    std::vector<uint32_t> level_end;  // wide node ids are level-ordered: level l ends at level_end[l]
    float bounds[6] = {0, 0, 0, 0, 0, 0};
};

// GPU BVH build: LBVH (63-bit Morton codes, LDS radix sort, Karras
// hierarchy, bottom-up refit) collapsed into 8-wide nodes.
void build_bvh(const float* d_vertices, const uint32_t* d_indices, uint32_t n_tris,
               uint32_t max_leaf, uint32_t build_flags, hipStream_t stream, BuildResult& out);
// Refit: same topology, boxes recomputed from new (already uploaded) vertices.
void refit_bvh(const float* d_vertices, const uint32_t* d_indices, uint32_t n_tris, hipStream_t stream,
               BuildResult& b, DevBuf& scratch);
double wide_sah(const BuildResult& b, hipStream_t stream);   // node cost 1, triangle cost 1, per root area

// Shared helpers of the trace / render / sensor translation units.
// Camera in the scene frame of `s` (position moved by the scene origin).
CamView make_cam_view(const crt_camera& c, const crt_scene* s);
void check_camera_args(const crt_camera* c);
size_t trace_lds_bytes(const SceneView& s);   // octet kernels
size_t lane_lds_bytes(const SceneView& s);    // lane kernels
size_t quad_lds_bytes(const SceneView& s);    // quad (any-hit) kernels
int device_cus();                             // compute units of the current device
float self_offset(const crt_scene* s);

// Render entry shared by crt_render, crt_render_batch and the server.
struct RenderWorkspace;
RenderWorkspace* new_render_workspace();
void delete_render_workspace(RenderWorkspace* w);
size_t render_bytes_estimate(uint64_t pixels, int spp, int n_lights);
size_t render_workspace_bytes(const RenderWorkspace* w);   // device bytes it holds
void check_light(const crt_light& L);                      // raise on bad light parameters
void check_batch_opts(const crt_batch_opts& B);            // raise on bad launch parameters
// `lights` holds n_lights lights shared by every camera, or, with
// per_camera_lights, n_cams x n_lights (camera c's lights at c * n_lights).
int render_batch(crt_scene* scene, uint32_t n_cams, const crt_camera* cams, const int32_t* max_samples,
                 const uint64_t* seeds, const crt_light* lights, uint32_t n_lights, bool per_camera_lights,
                 const crt_render_opts* opts, const crt_batch_opts* bopts, float* rgba, uint8_t* rgba_u8,
                 int32_t* spp_used, uint32_t flags, void* stream, crt_timing* timing, RenderWorkspace* ws);

// Scratch buffers reused across calls on one scene (grow-only).
struct Workspace {
    DevBuf a, b, c, d, e, f, g, h;
    DevBuf counters;
};

void flash_lidar_impl(crt_scene* scene, const crt_flash_opts* opts, double* single, double* first, double* last,
                      uint8_t* n_returns, uint32_t flags, void* stream, crt_timing* timing, Workspace& ws);
void radiation_pressure_impl(crt_scene* scene, const crt_flux_opts* opts, double force[3], double torque[3],
                             uint64_t* n_hits, void* stream, crt_timing* timing, Workspace& ws);
void check_flash_opts(const crt_flash_opts* opts);   // raise CRT_ERR_INVALID on bad options
void check_flux_opts(const crt_flux_opts* opts);
// Batched scans / flux grids of one scene, one launch (sequence) each; the
// device segment table sits in ws.e, `table_bytes` keeps its host copy until
// the caller synchronises.  outs: 3 device pointers (single, first, last) per
// scan; out7: device, 7 doubles per flux request.
void flash_lidar_multi(crt_scene* scene, uint32_t n, const crt_flash_opts* opts, double* const* outs,
                       hipStream_t st, Workspace& ws, std::vector<uint8_t>& table_bytes);
void radiation_pressure_multi(crt_scene* scene, uint32_t n, const crt_flux_opts* opts, double* out7,
                              hipStream_t st, Workspace& ws, std::vector<uint8_t>& table_bytes);

}  // namespace crt

struct crt_scene {
    int device = 0;
    uint32_t n_tris = 0, n_vertices = 0, n_instances = 0;
    crt::DevBuf vertices, indices, tri_vidx, vnormals, tri_inst, inst_color, inst_smooth;
    // materials (crt_scene_set_materials; empty: ColoredLambertian with inst_color)
    crt::DevBuf mats, vuv, texels, tex_info;
    uint32_t n_textures = 0;
    // refit (crt_scene_refit): build parameters for a rebuild, exact boxes
    uint32_t max_leaf = 8, build_flags = 0;
    crt::DevBuf refit_scratch;
    crt::BuildResult bvh;
    crt::SceneView view{};
    crt_bvh_info info{};
    // Grow-only device buffers of the scene's own entry points, one set per
    // stream: `ws` for passes, ray queries and sensors, `rws` for renders.
    // A call holds its stream's slot lock for its whole duration, so calls
    // on one stream of a scene run one at a time, while calls on different
    // streams use different buffers and run concurrently (host and device).
    // `mu` guards only the slot table.  The server's workers bring their
    // own buffers.
    struct StreamSlot {
        void* stream = nullptr;
        crt::Workspace ws;
        crt::RenderWorkspace* rws = nullptr;
        std::mutex mu;
    };
    std::vector<std::unique_ptr<StreamSlot>> slots;
    std::mutex mu;
    StreamSlot& slot(void* stream) {
        std::lock_guard<std::mutex> lk(mu);
        for (auto& e : slots)
            if (e->stream == stream) return *e;
        slots.emplace_back(new StreamSlot());
        slots.back()->stream = stream;
        return *slots.back();
    }
    crt_scene() = default;
    crt_scene(const crt_scene&) = delete;
    crt_scene& operator=(const crt_scene&) = delete;
    ~crt_scene() {
        for (auto& e : slots)
            if (e->rws) crt::delete_render_workspace(e->rws);
    }
};
