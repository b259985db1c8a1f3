// trace.hip -- scene lifetime, camera passes, general ray queries and the
// primary + sun-shadow pass, exported through the C ABI of crt_hip.h.
//
// Closest-hit kernels here trace one ray per lane over the compressed wide
// nodes (trace_lane, one 64-lane wave per block with its group stacks in
// LDS); the sun pass's shadow rays use one quad per ray over the 256-B nodes
// (trace_quad_any, 16 rays per wave and per one-wave block).
#include "crt_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

namespace crt {
namespace {

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

size_t trace_lds_bytes(const SceneView& s) { return (size_t)OCTETS_PER_BLOCK * s.stack_cap * 8u; }
size_t lane_lds_bytes(const SceneView& s) { return (size_t)LANE_BLOCK * s.lane_cap * 8u; }
size_t quad_lds_bytes(const SceneView& s) { return (size_t)(TRACE_BLOCK / QUAD) * s.stack_cap * 4u; }

int device_cus() {
    static std::mutex mu;
    static int cached[64] = {0};
    int dev = 0;
    CRT_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if (dev < 0 || dev >= 64) dev = 0;
    if (cached[dev] == 0) {
        int v = 0;
        CRT_HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
        cached[dev] = std::max(v, 1);
    }
    return cached[dev];
}

CamView make_cam_view(const crt_camera& c, const crt_scene* s) {
    CamView v;
    for (int k = 0; k < 3; ++k) v.pos[k] = c.position[k] - s->view.origin[k];
    for (int k = 0; k < 9; ++k) v.rot[k] = c.rotation[k];
    // src/crt/cameras/simple_camera.hpp:23-26: centre = res/2, scale = res/sensor
    v.cx = c.resolution[0] / 2.0;
    v.cy = c.resolution[1] / 2.0;
    v.sx = c.resolution[0] / c.sensor_size[0];
    v.sy = c.resolution[1] / c.sensor_size[1];
    v.f = c.z_positive ? c.focal_length : -c.focal_length;
    // src/pybind.cpp:534-535: (size_t) floor(resolution)
    v.w = (int)std::floor(c.resolution[0]);
    v.h = (int)std::floor(c.resolution[1]);
    return v;
}

void check_camera_args(const crt_camera* c) {
    CRT_REQUIRE(c != nullptr, CRT_ERR_INVALID, "camera is NULL");
    CRT_REQUIRE(c->resolution[0] >= 1 && c->resolution[1] >= 1, CRT_ERR_INVALID, "camera resolution must be >= 1");
    CRT_REQUIRE(c->resolution[0] * c->resolution[1] <= 1.0e9, CRT_ERR_INVALID, "camera resolution too large");
    CRT_REQUIRE(c->sensor_size[0] > 0 && c->sensor_size[1] > 0, CRT_ERR_INVALID, "sensor size must be > 0");
    CRT_REQUIRE(std::isfinite(c->focal_length), CRT_ERR_INVALID, "focal length must be finite");
}

// Self-intersection offset: 1e-4 absolute like the reference
// (src/crt/path_tracing/unidirectional.hpp:61), raised to 1e-6 of the scene
// extent so that it stays ~8 FP32 ulp above the coordinates' resolution.
float self_offset(const crt_scene* s) {
    double ext = 0.0;
    for (int k = 0; k < 6; ++k) ext = std::max(ext, std::fabs((double)s->bvh.bounds[k]));   // scene frame
    return (float)std::max(1e-4, 1e-6 * ext);
}

namespace {

// Closest/any hit of one FP64 ray (unit direction d, origin in the scene
// frame): rebase in FP64, trace in FP32 (one ray per lane), then t, u, v of
// the hit triangle re-evaluated in FP64.  Returns false on a miss.
template <bool ANY, bool STATS = false>
__device__ inline bool trace_f64(const SceneView& s, const double o[3], const double d[3], double tmin,
                                 double tmax, const LaneStack& stk, double& t_out, double& u, double& v, int& slot,
                                 TraversalStats* ts = nullptr) {
    double t0;
    if (!rebase_t0(s, o, d, t0)) return false;
    if (t0 > tmax) return false;
    const double ro[3] = {o[0] + t0 * d[0], o[1] + t0 * d[1], o[2] + t0 * d[2]};
    const RayPre r = make_ray_pre((float)ro[0], (float)ro[1], (float)ro[2], (float)d[0], (float)d[1], (float)d[2]);
    const float ltmin = (float)fmax(tmin - t0, 0.0);
    const float ltmax = tmax >= 3.0e38 ? __int_as_float(0x7f800000) : (float)(tmax - t0);
    const HitRec h = trace_lane<ANY, STATS>(s, r, ltmin, ltmax, stk, ts);
    if (h.slot < 0) return false;
    t_out = t0 + (double)h.t;
    u = h.u;
    v = h.v;
    slot = h.slot;
    refine_hit(s, slot, o, d, t_out, u, v);
    return true;
}

// Hit point from barycentrics in FP64 (src/crt/passes.hpp:53-58).
__device__ inline void bary_point(const SceneView& s, int slot, double u, double v, double p[3]) {
    const float4 a = s.tris[3 * slot], b = s.tris[3 * slot + 1], c = s.tris[3 * slot + 2];
    const double du = u, dv = v, dw = 1.0 - du - dv;
    p[0] = du * b.x + dv * c.x + dw * a.x;
    p[1] = du * b.y + dv * c.y + dw * a.y;
    p[2] = du * b.z + dv * c.z + dw * a.z;
}

// Normal pass value (src/crt/passes.hpp:176-181): flat = normalize((p0-p1)x(p2-p0))
// (inward for CCW-outward meshes), smooth = normalize(u vn1 + v vn2 + w vn0).
__device__ inline void hit_normal(const SceneView& s, int slot, double u, double v, double n[3]) {
    const float4 a = s.tris[3 * slot], b = s.tris[3 * slot + 1], c = s.tris[3 * slot + 2];
    const uint32_t orig = __float_as_uint(a.w);
    const uint32_t inst = s.tri_inst[orig];
    const bool smooth = s.vnormals != nullptr && inst >= 1 && inst <= s.n_instances && s.inst_smooth[inst - 1];
    double x, y, z;
    if (smooth) {
        const uint32_t i0 = s.tri_vidx[3 * orig], i1 = s.tri_vidx[3 * orig + 1], i2 = s.tri_vidx[3 * orig + 2];
        const double du = u, dv = v, dw = 1.0 - du - dv;
        x = du * s.vnormals[3 * i1] + dv * s.vnormals[3 * i2] + dw * s.vnormals[3 * i0];
        y = du * s.vnormals[3 * i1 + 1] + dv * s.vnormals[3 * i2 + 1] + dw * s.vnormals[3 * i0 + 1];
        z = du * s.vnormals[3 * i1 + 2] + dv * s.vnormals[3 * i2 + 2] + dw * s.vnormals[3 * i0 + 2];
    } else {
        const double e1x = (double)a.x - b.x, e1y = (double)a.y - b.y, e1z = (double)a.z - b.z;
        const double e2x = (double)c.x - a.x, e2y = (double)c.y - a.y, e2z = (double)c.z - a.z;
        x = e1y * e2z - e1z * e2y;
        y = e1z * e2x - e1x * e2z;
        z = e1x * e2y - e1y * e2x;
    }
    const double inv = 1.0 / sqrt(x * x + y * y + z * z);
    n[0] = x * inv; n[1] = y * inv; n[2] = z * inv;
}

// Pixel of lane `t` in block `blk`: a block (one wave) covers an 8x8 pixel
// tile, tiles run row-major.
constexpr int PIX_TILE = 8;
__device__ inline void tile_pixel(int blk, int t, int w, int& px, int& py) {
    const int tiles_x = (w + PIX_TILE - 1) / PIX_TILE;
    px = (blk % tiles_x) * PIX_TILE + (t & 7);
    py = (blk / tiles_x) * PIX_TILE + (t >> 3);
}

struct PassOut {
    double* xyz;
    double* normal;
    uint32_t* instance;
    double* depth;
    int32_t* tri;
};

CRT_LANE_KERNEL void k_camera_pass(SceneView s, CamView cam, PassOut out) {
    const LaneStack stk = lane_stack(s);
    int px, py;
    tile_pixel(blockIdx.x, threadIdx.x, cam.w, px, py);
    if (px >= cam.w || py >= cam.h) return;
    double d[3];
    // integer pixel coordinates, like src/crt/passes.hpp:46
    camera_dir(cam, (double)px, (double)py, d);
    double t;
    double u = 0.0, v = 0.0;
    int slot = -1;
    const bool hit = trace_f64<false>(s, cam.pos, d, 0.0, 1.0e300, stk, t, u, v, slot);
    const size_t pix = (size_t)py * cam.w + px;
    if (out.instance) out.instance[pix] = hit ? s.tri_inst[__float_as_uint(s.tris[3 * slot].w)] : 0u;
    if (out.tri) out.tri[pix] = hit ? (int32_t)__float_as_uint(s.tris[3 * slot].w) : -1;
    if (out.xyz || out.depth) {
        double p[3] = {0.0, 0.0, 0.0};
        if (hit) bary_point(s, slot, u, v, p);
        if (out.xyz) {
            out.xyz[3 * pix] = hit ? p[0] + s.origin[0] : 0.0;
            out.xyz[3 * pix + 1] = hit ? p[1] + s.origin[1] : 0.0;
            out.xyz[3 * pix + 2] = hit ? p[2] + s.origin[2] : 0.0;
        }
        if (out.depth) {
            double dd = 0.0;
            if (hit) {
                const double dx = p[0] - cam.pos[0], dy = p[1] - cam.pos[1], dz = p[2] - cam.pos[2];
                dd = sqrt(dx * dx + dy * dy + dz * dz);
            }
            out.depth[pix] = dd;
        }
    }
    if (out.normal) {
        double n[3] = {0.0, 0.0, 0.0};
        if (hit) hit_normal(s, slot, u, v, n);
        out.normal[3 * pix] = n[0];
        out.normal[3 * pix + 1] = n[1];
        out.normal[3 * pix + 2] = n[2];
    }
}

struct RayIn {
    const double* o;
    const double* d;
    const double* tmin;
    const double* tmax;
};
struct RayOut {
    double* t;
    int32_t* tri;
    double* u;
    double* v;
    uint32_t* counts;   // traversal statistics: nodes, boxes, triangles per ray (k_trace_rays<ANY, true>)
};

template <bool ANY, bool STATS = false>
CRT_LANE_KERNEL void k_trace_rays(SceneView s, RayIn in, RayOut out, uint64_t n) {
    const LaneStack stk = lane_stack(s);
    const uint64_t ray = (uint64_t)blockIdx.x * LANE_BLOCK + threadIdx.x;
    if (ray >= n) return;
    const double ow[3] = {in.o[3 * ray], in.o[3 * ray + 1], in.o[3 * ray + 2]};
    double o[3];
    to_scene(s, ow, o);
    double d[3] = {in.d[3 * ray], in.d[3 * ray + 1], in.d[3 * ray + 2]};
    const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double tmin = in.tmin ? in.tmin[ray] : 0.0;
    double tmax = in.tmax ? in.tmax[ray] : 1.0e300;
    bool hit = false;
    double t = 0.0;
    double u = 0.0, v = 0.0;
    int slot = -1;
    TraversalStats ts{0u, 0u, 0u};
    if (len > 0.0 && !(tmax < tmin)) {
        d[0] /= len; d[1] /= len; d[2] /= len;
        // distances along the unit direction
        tmin *= len;
        tmax = tmax >= 1.0e300 ? tmax : tmax * len;
        hit = trace_f64<ANY, STATS>(s, o, d, tmin, tmax, stk, t, u, v, slot, &ts);
        if (hit) t /= len;
    }
    if (STATS) {
        out.counts[3 * ray] = ts.nodes;
        out.counts[3 * ray + 1] = ts.boxes;
        out.counts[3 * ray + 2] = ts.tris;
    }
    if (out.t) out.t[ray] = hit ? t : -1.0;
    if (out.tri) out.tri[ray] = hit ? (int32_t)__float_as_uint(s.tris[3 * slot].w) : -1;
    if (out.u) out.u[ray] = hit ? u : 0.0;
    if (out.v) out.v[ray] = hit ? v : 0.0;
}

// ---- primary + sun shadow (wavefront: closest hit -> compacted shadow queue -> any hit)
struct ShadowRay {
    float4 o;       // origin xyz, pixel index (bits)
};
// The queue buffer holds n ShadowRay records followed by n uint32 origin
// triangles (leaf order), tested before the traversal (hint_blocks).
inline uint32_t* shadow_slots(ShadowRay* q, size_t n) { return reinterpret_cast<uint32_t*>(q + n); }

CRT_LANE_KERNEL void k_primary_sun(SceneView s, CamView cam, double* depth, int32_t* tri,
                                                             uint8_t* lit, ShadowRay* queue, uint32_t* qslot,
                                                             uint32_t* qcount, float offset) {
    const LaneStack stk = lane_stack(s);
    int px, py;
    tile_pixel(blockIdx.x, threadIdx.x, cam.w, px, py);
    const bool active = px < cam.w && py < cam.h;
    bool hit = false;
    double p[3] = {0.0, 0.0, 0.0};
    float nx = 0.f, ny = 0.f, nz = 0.f;
    size_t pix = 0;
    int slot = -1;
    if (active) {
        double d[3];
        camera_dir(cam, (double)px, (double)py, d);
        double t;
        double u = 0.0, v = 0.0;
        hit = trace_f64<false>(s, cam.pos, d, 0.0, 1.0e300, stk, t, u, v, slot);
        pix = (size_t)py * cam.w + px;
        if (hit) {
            bary_point(s, slot, u, v, p);
            double n[3];
            hit_normal(s, slot, u, v, n);
            nx = (float)n[0]; ny = (float)n[1]; nz = (float)n[2];
        }
        if (depth) {
            double dd = 0.0;
            if (hit) {
                const double dx = p[0] - cam.pos[0], dy = p[1] - cam.pos[1], dz = p[2] - cam.pos[2];
                dd = sqrt(dx * dx + dy * dy + dz * dz);
            }
            depth[pix] = dd;
        }
        if (tri) tri[pix] = hit ? (int32_t)__float_as_uint(s.tris[3 * slot].w) : -1;
        if (lit) lit[pix] = 0;
    }
    // one shadow ray per hit, compacted with a wave ballot + one atomic per wave
    const uint64_t m = __ballot(active && hit);
    if (m == 0ull) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(qcount, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, leader, 64);
    if (active && hit) {
        const uint32_t slot_q = base + (uint32_t)__popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
        // offset off the surface along the outward normal (the geometric normal is inward)
        const float ox = (float)p[0] - offset * nx, oy = (float)p[1] - offset * ny, oz = (float)p[2] - offset * nz;
        queue[slot_q].o = make_float4(ox, oy, oz, __uint_as_float((uint32_t)pix));
        qslot[slot_q] = (uint32_t)slot;
    }
}

__global__ void k_zero_u32(uint32_t* p) { *p = 0u; }

// Shadow rays are any-hit: quad traversal (16 rays per wave, two children
// per lane) beats both one ray per lane, whose waves wait on their longest
// unoccluded ray, and one ray per octet.  One wave (16 rays) per block.
__global__ __launch_bounds__(TRACE_BLOCK) void k_shadow_sun(SceneView s, const ShadowRay* queue,
                                                            const uint32_t* qslot, const uint32_t* qcount, float3 sun,
                                                            uint8_t* lit) {
    extern __shared__ uint32_t quad_stacks[];
    const int lane = threadIdx.x & 63;
    const int j = lane & (QUAD - 1), lane_base = lane & ~(QUAD - 1), local = threadIdx.x / QUAD;
    uint32_t* stk = quad_stacks + (size_t)local * s.stack_cap;
    const uint32_t n = *qcount;
    const uint32_t q = blockIdx.x * (TRACE_BLOCK / QUAD) + local;
    if (q >= n) return;   // quad-uniform
    const float4 o = queue[q].o;
    const RayPre r = make_ray_pre(o.x, o.y, o.z, sun.x, sun.y, sun.z);
    const float inf = __int_as_float(0x7f800000);
    const bool occluded = hint_blocks(s, r, qslot[q], 0.0f, inf) || trace_quad_any(s, r, 0.0f, inf, stk, j, lane_base);
    if (j == 0 && lit) lit[__float_as_uint(o.w)] = occluded ? 0 : 1;
}

// Events only when the caller asked for timing: a call with device outputs
// and no timing only enqueues work (no events, no synchronisation).
struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr, c = nullptr, d = nullptr;
    explicit EventTimer(bool on) {
        if (!on) return;
        CRT_HIP_CHECK(hipEventCreate(&a));
        CRT_HIP_CHECK(hipEventCreate(&b));
        CRT_HIP_CHECK(hipEventCreate(&c));
        CRT_HIP_CHECK(hipEventCreate(&d));
    }
    ~EventTimer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        if (c) (void)hipEventDestroy(c);
        if (d) (void)hipEventDestroy(d);
    }
    static void rec(hipEvent_t e, hipStream_t st) {
        if (e) CRT_HIP_CHECK(hipEventRecord(e, st));
    }
    static double ms(hipEvent_t x, hipEvent_t y) {
        float f = 0.f;
        CRT_HIP_CHECK(hipEventElapsedTime(&f, x, y));
        return f;
    }
};

// Kernel timing log (crt_timing_log_*): while enabled, entry points record
// a hipEvent pair around each of their dominant kernels on the launch
// stream, without synchronising, so a benchmark can time kernels inside its
// own timed region (streams overlapping) the way a profiler sees them.
struct TlogEntry {
    const char* kernel;
    hipEvent_t a, b;
};
std::mutex g_tlog_mu;
bool g_tlog_on = false;
std::vector<TlogEntry> g_tlog;

hipEvent_t tlog_event(hipStream_t st) {
    hipEvent_t e = nullptr;
    CRT_HIP_CHECK(hipEventCreate(&e));
    CRT_HIP_CHECK(hipEventRecord(e, st));
    return e;
}

// Device output staging: use the caller's device pointer directly, or a
// workspace buffer followed by a copy to the host pointer.
struct OutSlot {
    void* host = nullptr;
    void* dev = nullptr;
    size_t bytes = 0;
};

}  // namespace

// ---- host helpers shared with other translation units ----
void scene_make_view(crt_scene* s, const float* host_vertices);
void scene_bind_view(crt_scene* s);
uint64_t scene_device_bytes(const crt_scene* s);
void noalloc_setup_current_device();   // sensors.hip

}  // namespace crt

using namespace crt;

// Pointers, counts and stack sizes of s->view; the frame origin, sphere
// centre and radius must already be set.
void crt::scene_bind_view(crt_scene* s) {
    SceneView& v = s->view;
    v.nodes = s->bvh.nodes.as<float4>();
    v.cnodes = s->bvh.cnodes.as<float4>();
    v.tris = s->bvh.tris.as<float4>();
    v.tri_vidx = s->tri_vidx.as<uint32_t>();
    v.vnormals = s->vnormals.p ? s->vnormals.as<float>() : nullptr;
    v.tri_inst = s->tri_inst.as<uint32_t>();
    v.inst_color = s->inst_color.as<float>();
    v.inst_smooth = s->inst_smooth.as<uint8_t>();
    v.mats = s->mats.p ? s->mats.as<MatDev>() : nullptr;
    v.vuv = s->vuv.p ? s->vuv.as<float2>() : nullptr;
    v.texels = s->texels.p ? s->texels.as<uint32_t>() : nullptr;
    v.tex_info = s->tex_info.p ? s->tex_info.as<uint4>() : nullptr;
    v.n_tris = s->n_tris;
    v.n_instances = s->n_instances;
    v.stack_cap = std::max<uint32_t>(8u, 7u * s->bvh.max_depth + 1u);
    v.lane_cap = s->bvh.max_depth + 1u;
    for (int k = 0; k < 3; ++k) s->info.sphere[k] = v.center[k] + v.origin[k];   // world
    s->info.sphere[3] = v.radius;
    s->info.stack_capacity = v.stack_cap;
}

uint64_t crt::scene_device_bytes(const crt_scene* s) {
    return s->vertices.bytes + s->indices.bytes + s->tri_vidx.bytes + s->tri_inst.bytes + s->inst_color.bytes +
           s->inst_smooth.bytes + s->vnormals.bytes + s->bvh.nodes.bytes + s->bvh.cnodes.bytes + s->bvh.tris.bytes +
           s->mats.bytes + s->vuv.bytes + s->texels.bytes + s->tex_info.bytes;
}

void crt::scene_make_view(crt_scene* s, const float* host_vertices) {
    SceneView& v = s->view;
    // Bounding sphere for FP64 rebasing: centre of the box, radius = the
    // farthest vertex, padded so FP32 rounding of rebased origins stays
    // outside every triangle.
    const float* bb = s->bvh.bounds;   // scene frame
    const double cx = 0.5 * ((double)bb[0] + bb[3]), cy = 0.5 * ((double)bb[1] + bb[4]),
                 cz = 0.5 * ((double)bb[2] + bb[5]);
    double r2 = 0.0;
    for (uint64_t i = 0; i < s->n_vertices; ++i) {
        const double dx = host_vertices[3 * i] - cx, dy = host_vertices[3 * i + 1] - cy,
                     dz = host_vertices[3 * i + 2] - cz;
        r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
    }
    double r = std::sqrt(r2);
    r = r * (1.0 + 1e-3) + 1e-6 * (std::fabs(cx) + std::fabs(cy) + std::fabs(cz) + r) + 1e-9;
    v.center[0] = cx; v.center[1] = cy; v.center[2] = cz;
    v.radius = r;
    scene_bind_view(s);
}

extern "C" {

int crt_version(void) { return 100; }

int crt_timing_log_enable(int on) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_tlog_mu);
        g_tlog_on = on != 0;
    });
}

int crt_timing_log_read(const char* kernel, uint64_t* count, double* mean_ms) {
    return guarded([&] {
        CRT_REQUIRE(kernel && count && mean_ms, CRT_ERR_INVALID, "NULL argument");
        std::vector<TlogEntry> mine;
        {
            std::lock_guard<std::mutex> lk(g_tlog_mu);
            std::vector<TlogEntry> rest;
            for (const TlogEntry& e : g_tlog) (std::strcmp(e.kernel, kernel) == 0 ? mine : rest).push_back(e);
            g_tlog.swap(rest);
        }
        double sum = 0.0;
        for (const TlogEntry& e : mine) {
            CRT_HIP_CHECK(hipEventSynchronize(e.b));
            float f = 0.f;
            CRT_HIP_CHECK(hipEventElapsedTime(&f, e.a, e.b));
            sum += f;
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
        *count = mine.size();
        *mean_ms = mine.empty() ? 0.0 : sum / (double)mine.size();
    });
}

const char* crt_last_error(void) { return g_last_error.c_str(); }

int crt_device_init(int device) {
    return guarded([&] {
        int n = 0;
        CRT_HIP_CHECK(hipGetDeviceCount(&n));
        CRT_REQUIRE(n > 0, CRT_ERR_HIP, "no HIP device available");
        CRT_REQUIRE(device >= 0 && device < n, CRT_ERR_INVALID, "device index out of range");
        CRT_HIP_CHECK(hipSetDevice(device));
        noalloc_setup_current_device();
    });
}

}  // extern "C"

namespace {

// Scene frame: a body whose box contains the world origin keeps origin 0
// (its FP32 vertices stay exact; moving them by the box centre could push
// coordinates across a binade and round them, tilting small triangles'
// normals by ulp / edge); any other body takes its box centre rounded to a
// multiple of the FP32 ulp of its largest coordinate, which makes
// vertex - origin exact for FP32 input vertices of the origin's magnitude (a
// body far from the world origin).  FP64 input is rounded once, after the
// subtraction.  `local` receives the vertices in the frame.
template <typename VT>
void choose_frame(const VT* vertices, uint32_t n_vertices, double origin[3], std::vector<float>& local) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double amax = 0.0;
    for (uint64_t i = 0; i < n_vertices; ++i)
        for (int k = 0; k < 3; ++k) {
            const double x = vertices[3 * i + k];
            lo[k] = std::min(lo[k], x);
            hi[k] = std::max(hi[k], x);
            amax = std::max(amax, std::fabs(x));
        }
    int ex = 0;
    std::frexp(amax, &ex);
    const double q = std::ldexp(1.0, ex - 24);   // >= ulp of every vertex coordinate
    const bool around_zero = lo[0] <= 0.0 && hi[0] >= 0.0 && lo[1] <= 0.0 && hi[1] >= 0.0 && lo[2] <= 0.0 &&
                             hi[2] >= 0.0;
    for (int k = 0; k < 3; ++k)
        origin[k] = (amax > 0.0 && !around_zero) ? std::round(0.5 * (lo[k] + hi[k]) / q) * q : 0.0;
    local.resize((size_t)n_vertices * 3);
    for (uint64_t i = 0; i < 3ull * n_vertices; ++i) local[i] = (float)((double)vertices[i] - origin[i % 3]);
}

// Shared by crt_scene_create (FP32 vertices) and crt_scene_create_f64 (FP64
// vertices): the frame origin is chosen from the vertices at their own
// precision and subtracted before the one rounding to FP32, so an FP64 mesh
// far from the world origin keeps the FP32 resolution of its own size.
template <typename VT>
int scene_create_impl(const VT* vertices, uint32_t n_vertices, const uint32_t* indices, uint32_t n_tris,
                      const uint32_t* tri_instance, const float* vertex_normals, uint32_t n_instances,
                      const float* instance_color, const uint8_t* instance_smooth, const crt_build_opts* opts,
                      crt_scene** out_scene) {
    return guarded([&] {
        CRT_REQUIRE(out_scene != nullptr, CRT_ERR_INVALID, "out_scene is NULL");
        *out_scene = nullptr;
        CRT_REQUIRE(vertices && indices, CRT_ERR_INVALID, "vertices/indices are NULL");
        CRT_REQUIRE(n_tris > 0, CRT_ERR_INVALID, "scene has no triangles");
        CRT_REQUIRE(n_vertices > 0, CRT_ERR_INVALID, "scene has no vertices");
        for (uint64_t i = 0; i < 3ull * n_tris; ++i)
            CRT_REQUIRE(indices[i] < n_vertices, CRT_ERR_INVALID, "triangle index out of range");
        for (uint64_t i = 0; i < 3ull * n_vertices; ++i)
            CRT_REQUIRE(std::isfinite((double)vertices[i]) && std::fabs((double)vertices[i]) <= 1e18, CRT_ERR_INVALID,
                        "vertex coordinate is not finite or exceeds 1e18");
        const uint32_t ni = std::max(n_instances, 1u);
        if (tri_instance)
            for (uint32_t i = 0; i < n_tris; ++i)
                CRT_REQUIRE(tri_instance[i] >= 1 && tri_instance[i] <= ni, CRT_ERR_INVALID,
                            "triangle instance id out of range 1..n_instances");
        std::unique_ptr<crt_scene> s(new crt_scene());
        CRT_HIP_CHECK(hipGetDevice(&s->device));
        s->n_tris = n_tris;
        s->n_vertices = n_vertices;
        s->n_instances = ni;
        std::vector<float> local;
        choose_frame(vertices, n_vertices, s->view.origin, local);
        s->vertices.alloc((size_t)n_vertices * 12);
        s->indices.alloc((size_t)n_tris * 12);
        s->tri_vidx.alloc((size_t)n_tris * 12);
        s->tri_inst.alloc((size_t)n_tris * 4);
        s->inst_color.alloc((size_t)ni * 12);
        s->inst_smooth.alloc((size_t)ni);
        CRT_HIP_CHECK(hipMemcpy(s->vertices.p, local.data(), (size_t)n_vertices * 12, hipMemcpyHostToDevice));
        CRT_HIP_CHECK(hipMemcpy(s->indices.p, indices, (size_t)n_tris * 12, hipMemcpyHostToDevice));
        CRT_HIP_CHECK(hipMemcpy(s->tri_vidx.p, indices, (size_t)n_tris * 12, hipMemcpyHostToDevice));
        if (tri_instance) {
            CRT_HIP_CHECK(hipMemcpy(s->tri_inst.p, tri_instance, (size_t)n_tris * 4, hipMemcpyHostToDevice));
        } else {
            std::vector<uint32_t> ones(n_tris, 1u);
            CRT_HIP_CHECK(hipMemcpy(s->tri_inst.p, ones.data(), (size_t)n_tris * 4, hipMemcpyHostToDevice));
        }
        std::vector<float> col((size_t)ni * 3, 1.0f);
        if (instance_color) std::memcpy(col.data(), instance_color, (size_t)n_instances * 12);
        CRT_HIP_CHECK(hipMemcpy(s->inst_color.p, col.data(), col.size() * 4, hipMemcpyHostToDevice));
        std::vector<uint8_t> sm(ni, 0);
        if (instance_smooth) std::memcpy(sm.data(), instance_smooth, n_instances);
        CRT_HIP_CHECK(hipMemcpy(s->inst_smooth.p, sm.data(), sm.size(), hipMemcpyHostToDevice));
        if (vertex_normals) {
            s->vnormals.alloc((size_t)n_vertices * 12);
            CRT_HIP_CHECK(hipMemcpy(s->vnormals.p, vertex_normals, (size_t)n_vertices * 12, hipMemcpyHostToDevice));
        }
        const uint32_t max_leaf = opts && opts->max_leaf_size ? opts->max_leaf_size : 8u;
        CRT_REQUIRE(max_leaf <= 8u, CRT_ERR_INVALID, "max_leaf_size must be in 1..8");
        const uint32_t bflags = opts ? opts->flags : 0u;
        CRT_REQUIRE((bflags & ~(CRT_BUILD_PLOC | CRT_BUILD_GREEDY_WIDE | CRT_BUILD_LBVH)) == 0u, CRT_ERR_INVALID,
                    "unknown build flags");
        CRT_REQUIRE((bflags & (CRT_BUILD_PLOC | CRT_BUILD_LBVH)) != (CRT_BUILD_PLOC | CRT_BUILD_LBVH), CRT_ERR_INVALID,
                    "CRT_BUILD_PLOC and CRT_BUILD_LBVH exclude each other");
        s->max_leaf = max_leaf;
        s->build_flags = bflags;
        build_bvh(s->vertices.as<float>(), s->indices.as<uint32_t>(), n_tris, max_leaf, bflags, nullptr, s->bvh);
        crt_bvh_info& inf = s->info;
        inf.n_tris = n_tris;
        inf.n_wide_nodes = s->bvh.n_nodes;
        inf.n_leaves = s->bvh.n_leaves;
        inf.max_depth = s->bvh.max_depth;
        inf.build_ms = s->bvh.build_ms;
        inf.sah_cost = s->bvh.sah_cost;
        for (int k = 0; k < 6; ++k) inf.bounds[k] = s->bvh.bounds[k] + s->view.origin[k % 3];   // world, min/max xyz
        inf.device_bytes = scene_device_bytes(s.get());
        scene_make_view(s.get(), local.data());
        inf.self_offset = self_offset(s.get());
        CRT_REQUIRE(s->bvh.n_nodes < (1u << 23), CRT_ERR_INVALID, "BVH exceeds 2^23 wide nodes");
        CRT_REQUIRE(trace_lds_bytes(s->view) <= 64u * 1024u && lane_lds_bytes(s->view) <= 64u * 1024u &&
                        quad_lds_bytes(s->view) <= 64u * 1024u,
                    CRT_ERR_INVALID,
                    "BVH too deep for the LDS traversal stack");
        *out_scene = s.release();
    });
}

}  // namespace

extern "C" {

int crt_scene_create(const float* vertices, uint32_t n_vertices, const uint32_t* indices, uint32_t n_tris,
                     const uint32_t* tri_instance, const float* vertex_normals, uint32_t n_instances,
                     const float* instance_color, const uint8_t* instance_smooth, const crt_build_opts* opts,
                     crt_scene** out_scene) {
    return scene_create_impl(vertices, n_vertices, indices, n_tris, tri_instance, vertex_normals, n_instances,
                             instance_color, instance_smooth, opts, out_scene);
}

int crt_scene_create_f64(const double* vertices, uint32_t n_vertices, const uint32_t* indices, uint32_t n_tris,
                         const uint32_t* tri_instance, const float* vertex_normals, uint32_t n_instances,
                         const float* instance_color, const uint8_t* instance_smooth, const crt_build_opts* opts,
                         crt_scene** out_scene) {
    return scene_create_impl(vertices, n_vertices, indices, n_tris, tri_instance, vertex_normals, n_instances,
                             instance_color, instance_smooth, opts, out_scene);
}

int crt_scene_set_materials(crt_scene* scene, uint32_t n_materials, const crt_material* materials,
                            const float* vertex_uv, uint32_t n_textures, const crt_texture* textures) {
    return guarded([&] {
        CRT_REQUIRE(scene != nullptr, CRT_ERR_INVALID, "scene is NULL");
        CRT_REQUIRE(materials != nullptr && n_materials == scene->n_instances, CRT_ERR_INVALID,
                    "one material per instance is required");
        CRT_REQUIRE(n_textures == 0 || textures != nullptr, CRT_ERR_INVALID, "textures is NULL");
        // work already enqueued on any stream may still read the old tables
        CRT_HIP_CHECK(hipDeviceSynchronize());
        std::vector<MatDev> m(n_materials);
        for (uint32_t i = 0; i < n_materials; ++i) {
            const crt_material& x = materials[i];
            CRT_REQUIRE(x.type <= CRT_MAT_MIRROR, CRT_ERR_INVALID, "unknown material type");
            CRT_REQUIRE(x.texture >= -1 && x.texture < (int64_t)n_textures && x.spec_texture >= -1 &&
                            x.spec_texture < (int64_t)n_textures,
                        CRT_ERR_INVALID, "material texture index out of range");
            CRT_REQUIRE(std::isfinite(x.alpha), CRT_ERR_INVALID, "material alpha must be finite");
            for (int k = 0; k < 3; ++k)
                CRT_REQUIRE(std::isfinite(x.color[k]) && std::isfinite(x.coeffs[k]), CRT_ERR_INVALID,
                            "material colour / coefficients must be finite");
            MatDev& d = m[i];
            d.type = x.type;
            d.tex = x.texture;
            d.spec_tex = x.spec_texture;
            d.alpha = x.alpha;
            for (int k = 0; k < 3; ++k) { d.color[k] = x.color[k]; d.coeffs[k] = x.coeffs[k]; }
            d.color[3] = d.coeffs[3] = 0.f;
        }
        std::vector<uint4> info(n_textures);
        uint64_t total = 0;
        for (uint32_t t = 0; t < n_textures; ++t) {
            const crt_texture& x = textures[t];
            CRT_REQUIRE(x.width >= 1 && x.height >= 1 && x.rgba != nullptr, CRT_ERR_INVALID,
                        "texture must be at least 1 x 1 with pixel data");
            CRT_REQUIRE(total + (uint64_t)x.width * x.height < (1ull << 32), CRT_ERR_INVALID,
                        "textures exceed 2^32 texels");
            info[t] = make_uint4((uint32_t)total, x.width, x.height, 0u);
            total += (uint64_t)x.width * x.height;
        }
        std::vector<uint32_t> texels((size_t)total);
        for (uint32_t t = 0; t < n_textures; ++t)
            std::memcpy(texels.data() + info[t].x, textures[t].rgba, (size_t)info[t].y * info[t].z * 4);
        if (vertex_uv)
            for (uint64_t i = 0; i < 2ull * scene->n_vertices; ++i)
                CRT_REQUIRE(std::isfinite(vertex_uv[i]), CRT_ERR_INVALID, "vertex UVs must be finite");
        crt::DevBuf dm, duv, dtx, dti;
        dm.alloc(m.size() * sizeof(MatDev));
        CRT_HIP_CHECK(hipMemcpy(dm.p, m.data(), m.size() * sizeof(MatDev), hipMemcpyHostToDevice));
        if (vertex_uv) {
            duv.alloc((size_t)scene->n_vertices * 8);
            CRT_HIP_CHECK(hipMemcpy(duv.p, vertex_uv, (size_t)scene->n_vertices * 8, hipMemcpyHostToDevice));
        }
        if (n_textures) {
            dti.alloc(info.size() * sizeof(uint4));
            CRT_HIP_CHECK(hipMemcpy(dti.p, info.data(), info.size() * sizeof(uint4), hipMemcpyHostToDevice));
            dtx.alloc(std::max<size_t>(texels.size(), 1) * 4);
            if (!texels.empty())
                CRT_HIP_CHECK(hipMemcpy(dtx.p, texels.data(), texels.size() * 4, hipMemcpyHostToDevice));
        }
        scene->mats = std::move(dm);
        scene->vuv = std::move(duv);
        scene->texels = std::move(dtx);
        scene->tex_info = std::move(dti);
        scene->n_textures = n_textures;
        scene_bind_view(scene);
        scene->info.device_bytes = scene_device_bytes(scene);
    });
}

int crt_scene_refit(crt_scene* scene, const double* vertices, const float* vertex_normals, double max_sah_growth,
                    uint32_t* rebuilt) {
    return guarded([&] {
        CRT_REQUIRE(scene != nullptr && vertices != nullptr, CRT_ERR_INVALID, "NULL argument");
        crt_scene* s = scene;
        const uint32_t nv = s->n_vertices;
        for (uint64_t i = 0; i < 3ull * nv; ++i)
            CRT_REQUIRE(std::isfinite(vertices[i]) && std::fabs(vertices[i]) <= 1e18, CRT_ERR_INVALID,
                        "vertex coordinate is not finite or exceeds 1e18");
        const double growth = max_sah_growth > 0.0 ? max_sah_growth : 1.5;
        // work already enqueued on any stream may still read the old nodes
        // and triangles
        CRT_HIP_CHECK(hipDeviceSynchronize());
        std::vector<float> local;
        double origin[3];
        choose_frame(vertices, nv, origin, local);
        CRT_HIP_CHECK(hipMemcpy(s->vertices.p, local.data(), (size_t)nv * 12, hipMemcpyHostToDevice));
        if (vertex_normals) {
            if (!s->vnormals.p) s->vnormals.alloc((size_t)nv * 12);
            CRT_HIP_CHECK(hipMemcpy(s->vnormals.p, vertex_normals, (size_t)nv * 12, hipMemcpyHostToDevice));
        }
        for (int k = 0; k < 3; ++k) s->view.origin[k] = origin[k];
        refit_bvh(s->vertices.as<float>(), s->indices.as<uint32_t>(), s->n_tris, nullptr, s->bvh, s->refit_scratch);
        bool rb = false;
        if (s->bvh.sah_cost > growth * s->bvh.sah_at_build) {
            // boxes have grown apart (entities moved relative to each other): rebuild
            const uint32_t max_leaf = s->max_leaf, flags = s->build_flags;
            build_bvh(s->vertices.as<float>(), s->indices.as<uint32_t>(), s->n_tris, max_leaf, flags, nullptr, s->bvh);
            rb = true;
        } else {
            // scene-frame bounds = the root's children (exact boxes of the refit)
            float4 ex[16];
            CRT_HIP_CHECK(hipMemcpy(ex, s->refit_scratch.p, sizeof(ex), hipMemcpyDeviceToHost));
            float4 root[16];
            CRT_HIP_CHECK(hipMemcpy(root, s->bvh.nodes.p, sizeof(root), hipMemcpyDeviceToHost));
            float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
            for (int k = 0; k < 8; ++k) {
                uint32_t ref;
                std::memcpy(&ref, &root[2 * k].w, 4);
                if (ref == REF_EMPTY) continue;
                b[0] = std::min(b[0], ex[2 * k].x); b[1] = std::min(b[1], ex[2 * k].y); b[2] = std::min(b[2], ex[2 * k].z);
                b[3] = std::max(b[3], ex[2 * k + 1].x); b[4] = std::max(b[4], ex[2 * k + 1].y);
                b[5] = std::max(b[5], ex[2 * k + 1].z);
            }
            for (int k = 0; k < 6; ++k) s->bvh.bounds[k] = b[k];
        }
        crt_bvh_info& inf = s->info;
        inf.n_wide_nodes = s->bvh.n_nodes;
        inf.n_leaves = s->bvh.n_leaves;
        inf.max_depth = s->bvh.max_depth;
        if (rb) inf.build_ms = s->bvh.build_ms;
        inf.refit_ms = rb ? 0.0 : s->bvh.refit_ms;
        inf.sah_cost = s->bvh.sah_cost;
        for (int k = 0; k < 6; ++k) inf.bounds[k] = s->bvh.bounds[k] + s->view.origin[k % 3];
        scene_make_view(s, local.data());
        inf.self_offset = self_offset(s);
        inf.device_bytes = scene_device_bytes(s);
        CRT_REQUIRE(trace_lds_bytes(s->view) <= 64u * 1024u && lane_lds_bytes(s->view) <= 64u * 1024u &&
                        quad_lds_bytes(s->view) <= 64u * 1024u,
                    CRT_ERR_INVALID, "BVH too deep for the LDS traversal stack");
        if (rebuilt) *rebuilt = rb ? 1u : 0u;
    });
}

int crt_scene_info(const crt_scene* scene, crt_bvh_info* info) {
    return guarded([&] {
        CRT_REQUIRE(scene && info, CRT_ERR_INVALID, "NULL argument");
        *info = scene->info;
    });
}

void crt_scene_destroy(crt_scene* scene) {
    if (!scene) return;
    (void)hipDeviceSynchronize();   // enqueued work may still read the scene
    delete scene;
}

int crt_pass(crt_scene* scene, const crt_camera* camera, double* xyz, double* normal, uint32_t* instance,
             double* depth, int32_t* tri_id, uint32_t flags, void* stream, crt_timing* timing) {
    return guarded([&] {
        CRT_REQUIRE(scene != nullptr, CRT_ERR_INVALID, "scene is NULL");
        check_camera_args(camera);
        hipStream_t st = (hipStream_t)stream;
        const CamView cam = make_cam_view(*camera, scene);
        const size_t npix = (size_t)cam.w * cam.h;
        const bool dev = (flags & CRT_OUT_DEVICE) != 0;
        crt_scene::StreamSlot& slot = scene->slot(stream);
        std::lock_guard<std::mutex> slot_lock(slot.mu);
        Workspace& ws = slot.ws;
        PassOut o{nullptr, nullptr, nullptr, nullptr, nullptr};
        if (xyz) { if (dev) o.xyz = xyz; else { ws.a.reserve(npix * 24); o.xyz = ws.a.as<double>(); } }
        if (normal) { if (dev) o.normal = normal; else { ws.b.reserve(npix * 24); o.normal = ws.b.as<double>(); } }
        if (instance) { if (dev) o.instance = instance; else { ws.c.reserve(npix * 4); o.instance = ws.c.as<uint32_t>(); } }
        if (depth) { if (dev) o.depth = depth; else { ws.d.reserve(npix * 8); o.depth = ws.d.as<double>(); } }
        if (tri_id) { if (dev) o.tri = tri_id; else { ws.e.reserve(npix * 4); o.tri = ws.e.as<int32_t>(); } }
        EventTimer tm(timing != nullptr);
        const uint32_t blocks = cdiv((uint64_t)cam.w, PIX_TILE) * cdiv((uint64_t)cam.h, PIX_TILE);
        EventTimer::rec(tm.a, st);
        hipLaunchKernelGGL(k_camera_pass, dim3(blocks), dim3(LANE_BLOCK), lane_lds_bytes(scene->view), st,
                           scene->view, cam, o);
        CRT_HIP_CHECK(hipGetLastError());
        EventTimer::rec(tm.b, st);
        if (!dev) {
            if (xyz) CRT_HIP_CHECK(hipMemcpyAsync(xyz, o.xyz, npix * 24, hipMemcpyDeviceToHost, st));
            if (normal) CRT_HIP_CHECK(hipMemcpyAsync(normal, o.normal, npix * 24, hipMemcpyDeviceToHost, st));
            if (instance) CRT_HIP_CHECK(hipMemcpyAsync(instance, o.instance, npix * 4, hipMemcpyDeviceToHost, st));
            if (depth) CRT_HIP_CHECK(hipMemcpyAsync(depth, o.depth, npix * 8, hipMemcpyDeviceToHost, st));
            if (tri_id) CRT_HIP_CHECK(hipMemcpyAsync(tri_id, o.tri, npix * 4, hipMemcpyDeviceToHost, st));
        }
        if (!dev || timing) CRT_HIP_CHECK(hipStreamSynchronize(st));
        if (timing) {
            std::memset(timing, 0, sizeof(*timing));
            timing->trace_ms = EventTimer::ms(tm.a, tm.b);
            timing->total_ms = timing->trace_ms;
            timing->rays_traced = npix;
        }
    });
}

}  // extern "C"

namespace crt {
namespace {
int trace_rays_impl(crt_scene* scene, uint64_t n, const double* origins, const double* dirs, const double* tmin,
                    const double* tmax, double* t, int32_t* tri, double* u, double* v, uint32_t* counts,
                    uint32_t flags, void* stream, crt_timing* timing) {
    return guarded([&] {
        CRT_REQUIRE(scene != nullptr, CRT_ERR_INVALID, "scene is NULL");
        CRT_REQUIRE(origins && dirs, CRT_ERR_INVALID, "origins/dirs are NULL");
        if (n == 0) return;
        CRT_REQUIRE(n <= (1ull << 31), CRT_ERR_INVALID, "too many rays in one call");
        hipStream_t st = (hipStream_t)stream;
        const bool dev = (flags & CRT_OUT_DEVICE) != 0;
        const bool any = (flags & CRT_ANY_HIT) != 0;
        crt_scene::StreamSlot& slot = scene->slot(stream);
        std::lock_guard<std::mutex> slot_lock(slot.mu);
        Workspace& ws = slot.ws;
        RayIn in{origins, dirs, tmin, tmax};
        RayOut out{t, tri, u, v, counts};
        if (!dev) {
            // stage inputs: o(3n) d(3n) tmin(n) tmax(n) in ws.f; outputs in ws.g
            const size_t in_bytes = n * 8 * 8;
            ws.f.reserve(in_bytes);
            double* base = ws.f.as<double>();
            CRT_HIP_CHECK(hipMemcpyAsync(base, origins, n * 24, hipMemcpyHostToDevice, st));
            CRT_HIP_CHECK(hipMemcpyAsync(base + 3 * n, dirs, n * 24, hipMemcpyHostToDevice, st));
            in.o = base;
            in.d = base + 3 * n;
            if (tmin) { CRT_HIP_CHECK(hipMemcpyAsync(base + 6 * n, tmin, n * 8, hipMemcpyHostToDevice, st)); in.tmin = base + 6 * n; }
            if (tmax) { CRT_HIP_CHECK(hipMemcpyAsync(base + 7 * n, tmax, n * 8, hipMemcpyHostToDevice, st)); in.tmax = base + 7 * n; }
            ws.g.reserve(n * 40);
            char* ob = ws.g.as<char>();
            out.t = t ? reinterpret_cast<double*>(ob) : nullptr;
            out.u = u ? reinterpret_cast<double*>(ob + n * 8) : nullptr;
            out.v = v ? reinterpret_cast<double*>(ob + n * 16) : nullptr;
            out.tri = tri ? reinterpret_cast<int32_t*>(ob + n * 24) : nullptr;
            out.counts = counts ? reinterpret_cast<uint32_t*>(ob + n * 28) : nullptr;
        }
        EventTimer tm(timing != nullptr);
        const uint32_t blocks = cdiv(n, LANE_BLOCK);
        EventTimer::rec(tm.a, st);
        const size_t lds = lane_lds_bytes(scene->view);
        if (counts && any)
            hipLaunchKernelGGL((k_trace_rays<true, true>), dim3(blocks), dim3(LANE_BLOCK), lds, st, scene->view, in,
                               out, n);
        else if (counts)
            hipLaunchKernelGGL((k_trace_rays<false, true>), dim3(blocks), dim3(LANE_BLOCK), lds, st, scene->view, in,
                               out, n);
        else if (any)
            hipLaunchKernelGGL(k_trace_rays<true>, dim3(blocks), dim3(LANE_BLOCK), lds, st, scene->view, in, out, n);
        else
            hipLaunchKernelGGL(k_trace_rays<false>, dim3(blocks), dim3(LANE_BLOCK), lds, st, scene->view, in, out, n);
        CRT_HIP_CHECK(hipGetLastError());
        EventTimer::rec(tm.b, st);
        if (!dev) {
            if (t) CRT_HIP_CHECK(hipMemcpyAsync(t, out.t, n * 8, hipMemcpyDeviceToHost, st));
            if (u) CRT_HIP_CHECK(hipMemcpyAsync(u, out.u, n * 8, hipMemcpyDeviceToHost, st));
            if (v) CRT_HIP_CHECK(hipMemcpyAsync(v, out.v, n * 8, hipMemcpyDeviceToHost, st));
            if (tri) CRT_HIP_CHECK(hipMemcpyAsync(tri, out.tri, n * 4, hipMemcpyDeviceToHost, st));
            if (counts) CRT_HIP_CHECK(hipMemcpyAsync(counts, out.counts, n * 12, hipMemcpyDeviceToHost, st));
        }
        if (!dev || timing) CRT_HIP_CHECK(hipStreamSynchronize(st));
        if (timing) {
            std::memset(timing, 0, sizeof(*timing));
            const double ms = EventTimer::ms(tm.a, tm.b);
            if (any) { timing->shadow_ms = ms; timing->shadow_rays = n; }
            else { timing->trace_ms = ms; timing->rays_traced = n; }
            timing->total_ms = ms;
        }
    });
}
}  // namespace
}  // namespace crt

extern "C" {

int crt_trace_rays(crt_scene* scene, uint64_t n, const double* origins, const double* dirs, const double* tmin,
                   const double* tmax, double* t, int32_t* tri, double* u, double* v, uint32_t flags, void* stream,
                   crt_timing* timing) {
    return trace_rays_impl(scene, n, origins, dirs, tmin, tmax, t, tri, u, v, nullptr, flags, stream, timing);
}

int crt_trace_stats(crt_scene* scene, uint64_t n, const double* origins, const double* dirs, uint32_t flags,
                    uint32_t* counts, double* t, int32_t* tri, void* stream) {
    if (counts == nullptr) return guarded([] { CRT_REQUIRE(false, CRT_ERR_INVALID, "counts is NULL"); });
    return trace_rays_impl(scene, n, origins, dirs, nullptr, nullptr, t, tri, nullptr, nullptr, counts,
                           flags & ~(uint32_t)CRT_OUT_DEVICE, stream, nullptr);
}

int crt_sun_pass(crt_scene* scene, const crt_camera* camera, const double sun_dir[3], double* depth, int32_t* tri_id,
                 uint8_t* lit, uint32_t flags, void* stream, crt_timing* timing) {
    return guarded([&] {
        CRT_REQUIRE(scene != nullptr, CRT_ERR_INVALID, "scene is NULL");
        check_camera_args(camera);
        CRT_REQUIRE(sun_dir != nullptr, CRT_ERR_INVALID, "sun_dir is NULL");
        const double sl = std::sqrt(sun_dir[0] * sun_dir[0] + sun_dir[1] * sun_dir[1] + sun_dir[2] * sun_dir[2]);
        CRT_REQUIRE(sl > 0 && std::isfinite(sl), CRT_ERR_INVALID, "sun_dir must be a finite non-zero vector");
        const float3 sun = make_float3((float)(sun_dir[0] / sl), (float)(sun_dir[1] / sl), (float)(sun_dir[2] / sl));
        hipStream_t st = (hipStream_t)stream;
        const CamView cam = make_cam_view(*camera, scene);
        const size_t npix = (size_t)cam.w * cam.h;
        const bool dev = (flags & CRT_OUT_DEVICE) != 0;
        crt_scene::StreamSlot& slot = scene->slot(stream);
        std::lock_guard<std::mutex> slot_lock(slot.mu);
        Workspace& ws = slot.ws;
        double* d_depth = depth;
        int32_t* d_tri = tri_id;
        uint8_t* d_lit = lit;
        if (!dev) {
            if (depth) { ws.d.reserve(npix * 8); d_depth = ws.d.as<double>(); }
            if (tri_id) { ws.e.reserve(npix * 4); d_tri = ws.e.as<int32_t>(); }
            if (lit) { ws.c.reserve(npix); d_lit = ws.c.as<uint8_t>(); }
        }
        ws.h.reserve(npix * (sizeof(ShadowRay) + sizeof(uint32_t)));
        ws.counters.reserve(256);
        uint32_t* qcount = ws.counters.as<uint32_t>();
        const float offset = self_offset(scene);
        EventTimer tm(timing != nullptr);
        const uint32_t blocks = cdiv((uint64_t)cam.w, PIX_TILE) * cdiv((uint64_t)cam.h, PIX_TILE);
        // cleared by a kernel, not a memset node, so that the call can be
        // captured in a graph
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, st, qcount);
        bool log;
        {
            std::lock_guard<std::mutex> lk(g_tlog_mu);
            log = g_tlog_on;
        }
        hipEvent_t la = log ? tlog_event(st) : nullptr;
        EventTimer::rec(tm.a, st);
        hipLaunchKernelGGL(k_primary_sun, dim3(blocks), dim3(LANE_BLOCK), lane_lds_bytes(scene->view), st,
                           scene->view, cam, d_depth, d_tri, d_lit, ws.h.as<ShadowRay>(),
                           shadow_slots(ws.h.as<ShadowRay>(), npix), qcount, offset);
        EventTimer::rec(tm.b, st);
        hipEvent_t lb = log ? tlog_event(st) : nullptr, lc = log ? tlog_event(st) : nullptr;
        // one one-wave block per 16 pixels, though only hits have a shadow
        // ray: blocks past the queue exit after one load, and blocks
        // dispatched in order keep the rays in flight one contiguous stretch
        // of the queue (neighbouring pixels sharing BVH nodes in L2).  On the
        // 2048^2 L8 pass (1.8M shadow rays, 57 % of the blocks empty) this
        // takes 0.35 ms; a grid of 32 blocks per CU striding over the queue
        // 0.48 ms (the rays in flight span the image), with an atomic claim
        // of 16 rays per wave 1.42 ms (one contended counter).
        hipLaunchKernelGGL(k_shadow_sun, dim3(cdiv(npix, TRACE_BLOCK / QUAD)), dim3(TRACE_BLOCK),
                           quad_lds_bytes(scene->view), st, scene->view, ws.h.as<ShadowRay>(),
                           shadow_slots(ws.h.as<ShadowRay>(), npix), qcount, sun, d_lit);
        CRT_HIP_CHECK(hipGetLastError());
        EventTimer::rec(tm.c, st);
        if (log) {
            hipEvent_t ld = tlog_event(st);
            std::lock_guard<std::mutex> lk(g_tlog_mu);
            g_tlog.push_back({"k_primary_sun", la, lb});
            g_tlog.push_back({"k_shadow_sun", lc, ld});
        }
        uint32_t nshadow = 0;
        if (timing) CRT_HIP_CHECK(hipMemcpyAsync(&nshadow, qcount, 4, hipMemcpyDeviceToHost, st));
        if (!dev) {
            if (depth) CRT_HIP_CHECK(hipMemcpyAsync(depth, d_depth, npix * 8, hipMemcpyDeviceToHost, st));
            if (tri_id) CRT_HIP_CHECK(hipMemcpyAsync(tri_id, d_tri, npix * 4, hipMemcpyDeviceToHost, st));
            if (lit) CRT_HIP_CHECK(hipMemcpyAsync(lit, d_lit, npix, hipMemcpyDeviceToHost, st));
        }
        if (!dev || timing) CRT_HIP_CHECK(hipStreamSynchronize(st));
        if (timing) {
            std::memset(timing, 0, sizeof(*timing));
            timing->trace_ms = EventTimer::ms(tm.a, tm.b);
            timing->shadow_ms = EventTimer::ms(tm.b, tm.c);
            timing->total_ms = EventTimer::ms(tm.a, tm.c);
            timing->rays_traced = npix;
            timing->shadow_rays = nshadow;
        }
    });
}

}  // extern "C"
