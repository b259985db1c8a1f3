// render.hip -- wavefront path tracer (crt_render).
//
// Semantics of do_render + unidirectional (src/crt/do_render.hpp:65-134,
// src/crt/path_tracing/unidirectional.hpp:24-96), Lambertian material
// (src/crt/materials/material.hpp:90-111, brdfs.hpp:17-42):
//   * per pixel, samples s = 1..max_samples; jitter U(-1/2,1/2) unless
//     max_samples == 1; running mean m += (x - m) / s in float; stop after
//     s >= min_samples when |m_s - m_{s-1}| < noise_threshold
//   * path weight starts at 2*pi; per bounce: geometric normal n (inward),
//     shading normal (smooth ? interpolated : n), point p - eps*n; for every
//     light: colour * (-L.n_s) * irradiance if the shadow ray is clear; the
//     sum is clamped to [0,1] from bounce 1 on; radiance += sum * weight;
//     then a Malley cosine sample around n_s and weight *= (1 - r1).
//
// MI355X structure: samples are processed in passes; each pass is a
// wavefront over (pixel, sample) paths:
//   k_raygen  -> FP64 camera ray, rebased, FP32 path state
//   per bounce: k_extend (octet closest hit) -> k_shade (thread per path:
//     light samples + shadow rays + next direction) -> k_shadow (octet any
//     hit) -> k_accumulate (thread per path; survivors are compacted into the
//     next queue with a wave ballot + one atomic per wave)
//   k_fold    -> per pixel, folds the pass's samples into the running mean in
//     sample order and applies the adaptive stop, so results do not depend on
//     how samples are batched.
// Randomness is counter-based: u01(seed, key_offset + pixel, sample,
// 64 * bounce + dim), no shared generator state.
#include "crt_internal.hpp"
#include "crt_rng.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace crt {
namespace {

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

constexpr int MAX_LIGHTS = 8;
constexpr int TB = 256;

struct LightDev {
    int type;
    float intensity;
    float pos[3];
    float rot[9];
    float size[2];
    float dir[3];
};

struct PathState {
    float4 o;          // origin xyz, unused
    float4 d;          // direction xyz, unused
    float4 w;          // weight rgb, pending bounce factor
    float4 L;          // radiance rgb, unused
    uint32_t pixel;    // pixel index
    uint32_t sample;   // 1-based sample index
    int32_t slot;      // hit triangle (leaf order) or -1
    float hu, hv;      // hit barycentrics
    uint32_t pad[3];
};

struct RenderParams {
    SceneView s;
    CamView cam;
    LightDev lights[MAX_LIGHTS];
    int nl;
    int nb;
    int jitter;
    float offset;
    uint64_t seed;
    uint64_t key_offset;
};

__device__ inline float3 f3(float4 a) { return make_float3(a.x, a.y, a.z); }
__device__ inline float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline float3 norm3(float3 a) {
    const float inv = 1.0f / sqrtf(dot3(a, a));
    return make_float3(a.x * inv, a.y * inv, a.z * inv);
}

// Wave-aggregated append: returns the slot of this lane (valid only if `want`).
__device__ inline uint32_t wave_append(bool want, uint32_t* counter) {
    const uint64_t m = __ballot(want);
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (m != 0ull && lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader < 0 ? 0 : leader, 64);
    return base + (uint32_t)__popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
}

// Paths of a pass: list of (pixel, first sample, count); path k of the pass
// = pixel list[k / count], sample first + k % count.
__global__ __launch_bounds__(TB) void k_raygen(RenderParams P, const uint32_t* __restrict__ pixels, uint32_t npix,
                                               uint32_t s_first, uint32_t s_count, PathState* __restrict__ st,
                                               uint32_t* __restrict__ queue, uint32_t* __restrict__ qcount,
                                               float4* __restrict__ sample_out) {
    const uint64_t k = (uint64_t)blockIdx.x * TB + threadIdx.x;
    const uint64_t n = (uint64_t)npix * s_count;
    bool alive = false;
    if (k < n) {
        const uint32_t pix = pixels[k / s_count];
        const uint32_t smp = s_first + (uint32_t)(k % s_count);
        const int px = (int)(pix % (uint32_t)P.cam.w), py = (int)(pix / (uint32_t)P.cam.w);
        double u = px, v = py;
        const uint64_t a = P.key_offset + pix;
        if (P.jitter) {
            u += rng_u01(P.seed, a, smp, rng_dim(0, RNG_JITTER)) - 0.5;
            v += rng_u01(P.seed, a, smp, rng_dim(0, RNG_JITTER + 1)) - 0.5;
        }
        double d[3];
        camera_dir(P.cam, u, v, d);
        double t0 = 0.0;
        PathState ps;
        ps.pixel = pix;
        ps.sample = smp;
        ps.slot = -1;
        ps.hu = ps.hv = 0.f;
        const float tp = (float)(2.0 * 3.14159265358979323846);
        ps.w = make_float4(tp, tp, tp, 1.0f);
        ps.L = make_float4(0.f, 0.f, 0.f, 0.f);
        if (rebase_t0(P.s, P.cam.pos, d, t0)) {
            ps.o = make_float4((float)(P.cam.pos[0] + t0 * d[0]), (float)(P.cam.pos[1] + t0 * d[1]),
                               (float)(P.cam.pos[2] + t0 * d[2]), 0.f);
            ps.d = make_float4((float)d[0], (float)d[1], (float)d[2], 0.f);
            alive = true;
        } else {
            ps.o = make_float4(0.f, 0.f, 0.f, 0.f);
            ps.d = make_float4(0.f, 0.f, 1.f, 0.f);
        }
        st[k] = ps;
        if (!alive) sample_out[k] = make_float4(0.f, 0.f, 0.f, 0.f);   // camera ray misses the scene
    }
    const uint32_t slot = wave_append(alive, qcount);
    if (alive) queue[slot] = (uint32_t)k;
}

__global__ __launch_bounds__(TRACE_BLOCK) void k_extend(SceneView s, PathState* __restrict__ st,
                                                        const uint32_t* __restrict__ queue,
                                                        const uint32_t* __restrict__ qcount) {
    extern __shared__ uint32_t smem[];
    const int lane = threadIdx.x & 63;
    const int sub = lane & 7, lane_base = lane & ~7, local = threadIdx.x >> 3;
    uint32_t* sref = smem + (size_t)local * s.stack_cap;
    float* sd = reinterpret_cast<float*>(smem + (size_t)OCTETS_PER_BLOCK * s.stack_cap) + (size_t)local * s.stack_cap;
    const uint32_t q = blockIdx.x * OCTETS_PER_BLOCK + local;
    if (q >= *qcount) return;
    const uint32_t k = queue[q];
    const float4 o = st[k].o, d = st[k].d;
    const RayPre r = make_ray_pre(o.x, o.y, o.z, d.x, d.y, d.z);
    const HitRec h = trace_octet<false>(s, r, 0.0f, __int_as_float(0x7f800000), sref, sd, sub, lane_base);
    if (sub == 0) {
        st[k].slot = h.slot;
        st[k].hu = h.u;
        st[k].hv = h.v;
    }
}

struct ShadowRec {
    float4 o;      // origin, tmax
    float4 d;      // direction, unused
    float4 c;      // colour * cos * irradiance, occluded flag (set by k_shadow)
};

__device__ inline void light_sample(const RenderParams& P, int l, float3 p, uint64_t a, uint32_t smp, int bounce,
                                    float3& dir, float& tmax, float& irr) {
    const LightDev& L = P.lights[l];
    if (L.type == CRT_LIGHT_POINT) {
        // src/crt/lights/point_light.hpp:32-39
        const float3 dd = make_float3(L.pos[0] - p.x, L.pos[1] - p.y, L.pos[2] - p.z);
        const float d2 = dot3(dd, dd);
        const float len = sqrtf(d2);
        dir = make_float3(dd.x / len, dd.y / len, dd.z / len);
        tmax = len;
        irr = fminf(L.intensity / d2, 10000.0f);
    } else if (L.type == CRT_LIGHT_AREA) {
        // src/crt/lights/area_light.hpp:54-72: uniform point of the local XY
        // rectangle, placed with R (transform.hpp:31-40); I / d^2, no cosine
        const float r1 = (float)rng_u01(P.seed, a, smp, rng_dim(bounce, RNG_LIGHT + 2 * l));
        const float r2 = (float)rng_u01(P.seed, a, smp, rng_dim(bounce, RNG_LIGHT + 2 * l + 1));
        const float x = L.size[0] * r1 - L.size[0] / 2, y = L.size[1] * r2 - L.size[1] / 2;
        const float3 sp = make_float3(L.rot[0] * x + L.rot[1] * y + L.pos[0], L.rot[3] * x + L.rot[4] * y + L.pos[1],
                                      L.rot[6] * x + L.rot[7] * y + L.pos[2]);
        const float3 dd = make_float3(sp.x - p.x, sp.y - p.y, sp.z - p.z);
        const float d2 = dot3(dd, dd);
        const float len = sqrtf(d2);
        dir = make_float3(dd.x / len, dd.y / len, dd.z / len);
        tmax = len;
        irr = L.intensity / d2;
    } else {
        dir = make_float3(L.dir[0], L.dir[1], L.dir[2]);
        tmax = __int_as_float(0x7f800000);
        irr = L.intensity;
    }
}

// Malley cosine sample around the (inward) shading normal, brdfs.hpp:17-42.
__device__ inline float3 cosine_sample(float3 n, float r1, float r2) {
    const float r = sqrtf(r1);
    const float th = r2 * 2.0f * 3.1415926f;
    float sn, cs;
    sincosf(th, &sn, &cs);
    const float x = -r * cs, y = -r * sn;
    const float z = -sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y));
    float3 nx;
    if (fabsf(n.x) > fabsf(n.y)) {
        const float inv = 1.0f / sqrtf(n.x * n.x + n.z * n.z);
        nx = make_float3(n.z * inv, 0.0f, -n.x * inv);
    } else {
        const float inv = 1.0f / sqrtf(n.y * n.y + n.z * n.z);
        nx = make_float3(0.0f, n.z * inv, -n.y * inv);
    }
    const float3 ny = cross3(n, nx);
    return make_float3(nx.x * x + ny.x * y + n.x * z, nx.y * x + ny.y * y + n.y * z, nx.z * x + ny.z * y + n.z * z);
}

__global__ __launch_bounds__(TB) void k_shade(RenderParams P, int bounce, PathState* __restrict__ st,
                                              const uint32_t* __restrict__ queue, const uint32_t* __restrict__ qcount,
                                              ShadowRec* __restrict__ shadow, uint32_t* __restrict__ squeue,
                                              uint32_t* __restrict__ scount, float4* __restrict__ sample_out,
                                              uint32_t* __restrict__ live_count, uint32_t* __restrict__ live) {
    const uint32_t q = blockIdx.x * TB + threadIdx.x;
    const uint32_t n = *qcount;
    bool has_hit = false;
    uint32_t k = 0;
    PathState ps;
    if (q < n) {
        k = queue[q];
        ps = st[k];
        if (ps.slot < 0) {
            sample_out[k] = ps.L;     // path ends: no hit
        } else {
            has_hit = true;
        }
    }
    const uint32_t ls = wave_append(has_hit, live_count);
    if (has_hit) live[ls] = k;
    if (!has_hit) return;
    const SceneView& s = P.s;
    const float4 a = s.tris[3 * ps.slot], b = s.tris[3 * ps.slot + 1], c = s.tris[3 * ps.slot + 2];
    const uint32_t orig = __float_as_uint(a.w);
    const uint32_t inst = s.tri_inst[orig];
    const float u = ps.hu, v = ps.hv, w = 1.0f - u - v;
    // geometric normal (inward): normalize((p0 - p1) x (p2 - p0))
    const float3 e1 = make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
    const float3 e2 = make_float3(c.x - a.x, c.y - a.y, c.z - a.z);
    const float3 gn = norm3(cross3(e1, e2));
    float3 sn = gn;
    if (s.vnormals && inst >= 1 && inst <= s.n_instances && s.inst_smooth[inst - 1]) {
        const uint32_t i0 = s.tri_vidx[3 * orig], i1 = s.tri_vidx[3 * orig + 1], i2 = s.tri_vidx[3 * orig + 2];
        sn = norm3(make_float3(u * s.vnormals[3 * i1] + v * s.vnormals[3 * i2] + w * s.vnormals[3 * i0],
                               u * s.vnormals[3 * i1 + 1] + v * s.vnormals[3 * i2 + 1] + w * s.vnormals[3 * i0 + 1],
                               u * s.vnormals[3 * i1 + 2] + v * s.vnormals[3 * i2 + 2] + w * s.vnormals[3 * i0 + 2]));
    }
    float3 p = make_float3(u * b.x + v * c.x + w * a.x, u * b.y + v * c.y + w * a.y, u * b.z + v * c.z + w * a.z);
    p = make_float3(p.x - P.offset * gn.x, p.y - P.offset * gn.y, p.z - P.offset * gn.z);
    const float* col = s.inst_color + 3 * (inst >= 1 && inst <= s.n_instances ? inst - 1 : 0);
    const uint64_t key = P.key_offset + ps.pixel;
    for (int l = 0; l < P.nl; ++l) {
        float3 ld;
        float tmax, irr;
        light_sample(P, l, p, key, ps.sample, bounce, ld, tmax, irr);
        // ColoredLambertianMaterial::compute: colour * (-L.n_s), no max(0, .)
        const float ldn = -dot3(ld, sn);
        ShadowRec r;
        r.o = make_float4(p.x, p.y, p.z, tmax);
        r.d = make_float4(ld.x, ld.y, ld.z, 0.f);
        r.c = make_float4(col[0] * ldn * irr, col[1] * ldn * irr, col[2] * ldn * irr, 0.f);
        shadow[(size_t)k * P.nl + l] = r;
    }
    if (P.nl > 0) {
        const uint32_t base = atomicAdd(scount, (uint32_t)P.nl);
        for (int l = 0; l < P.nl; ++l) squeue[base + l] = (uint32_t)((size_t)k * P.nl + l);
    }
    if (bounce < P.nb - 1) {
        const float r1 = (float)rng_u01(P.seed, key, ps.sample, rng_dim(bounce, RNG_BRDF));
        const float r2 = (float)rng_u01(P.seed, key, ps.sample, rng_dim(bounce, RNG_BRDF + 1));
        const float3 nd = cosine_sample(sn, r1, r2);
        st[k].o = make_float4(p.x, p.y, p.z, 0.f);
        st[k].d = make_float4(nd.x, nd.y, nd.z, 0.f);
        st[k].w.w = 1.0f - r1;
    }
}

__global__ __launch_bounds__(TRACE_BLOCK) void k_shadow(SceneView s, ShadowRec* __restrict__ shadow,
                                                        const uint32_t* __restrict__ squeue,
                                                        const uint32_t* __restrict__ scount) {
    extern __shared__ uint32_t smem[];
    const int lane = threadIdx.x & 63;
    const int sub = lane & 7, lane_base = lane & ~7, local = threadIdx.x >> 3;
    uint32_t* sref = smem + (size_t)local * s.stack_cap;
    float* sd = reinterpret_cast<float*>(smem + (size_t)OCTETS_PER_BLOCK * s.stack_cap) + (size_t)local * s.stack_cap;
    const uint32_t q = blockIdx.x * OCTETS_PER_BLOCK + local;
    if (q >= *scount) return;
    const uint32_t i = squeue[q];
    const float4 o = shadow[i].o, d = shadow[i].d;
    const RayPre r = make_ray_pre(o.x, o.y, o.z, d.x, d.y, d.z);
    const HitRec h = trace_octet<true>(s, r, 0.0f, o.w, sref, sd, sub, lane_base);
    if (sub == 0) shadow[i].c.w = h.slot >= 0 ? 1.0f : 0.0f;
}

__global__ __launch_bounds__(TB) void k_accumulate(RenderParams P, int bounce, PathState* __restrict__ st,
                                                   const uint32_t* __restrict__ live, const uint32_t* __restrict__ lcount,
                                                   const ShadowRec* __restrict__ shadow,
                                                   uint32_t* __restrict__ next_queue, uint32_t* __restrict__ next_count,
                                                   float4* __restrict__ sample_out) {
    const uint32_t q = blockIdx.x * TB + threadIdx.x;
    const uint32_t n = *lcount;
    bool cont = false;
    uint32_t k = 0;
    if (q < n) {
        k = live[q];
        PathState& ps = st[k];
        float lr[3] = {0.f, 0.f, 0.f};
        for (int l = 0; l < P.nl; ++l) {
            const float4 c = shadow[(size_t)k * P.nl + l].c;
            if (c.w == 0.0f) { lr[0] += c.x; lr[1] += c.y; lr[2] += c.z; }
        }
        if (bounce >= 1)
            for (int j = 0; j < 3; ++j) lr[j] = fminf(fmaxf(lr[j], 0.0f), 1.0f);
        float4 L = ps.L;
        float4 w = ps.w;
        L.x += lr[0] * w.x;
        L.y += lr[1] * w.y;
        L.z += lr[2] * w.z;
        ps.L = L;
        if (bounce == P.nb - 1) {
            sample_out[k] = L;
        } else {
            w.x *= w.w; w.y *= w.w; w.z *= w.w;
            ps.w = w;
            cont = true;
        }
    }
    const uint32_t slot = wave_append(cont, next_count);
    if (cont) next_queue[slot] = k;
}

// Fold the pass's samples into the running mean (do_render.hpp:100-108).
__global__ __launch_bounds__(TB) void k_fold(const uint32_t* __restrict__ pixels, uint32_t npix, uint32_t s_first,
                                             uint32_t s_count, int min_samples, float thr,
                                             const float4* __restrict__ sample_out, float4* __restrict__ mean,
                                             int32_t* __restrict__ spp_used, uint8_t* __restrict__ active) {
    const uint32_t i = blockIdx.x * TB + threadIdx.x;
    if (i >= npix) return;
    const uint32_t pix = pixels[i];
    float4 m = mean[pix];
    int used = spp_used[pix];
    bool act = true;
    for (uint32_t j = 0; j < s_count; ++j) {
        const uint32_t smp = s_first + j;
        const float4 x = sample_out[(size_t)i * s_count + j];
        const float inv = 1.0f / (float)smp;
        const float rx = (x.x - m.x) * inv, ry = (x.y - m.y) * inv, rz = (x.z - m.z) * inv;
        m.x += rx; m.y += ry; m.z += rz;
        used = (int)smp;
        if ((int)smp >= min_samples) {
            const float noise = sqrtf(rx * rx + ry * ry + rz * rz);
            if (noise < thr) { act = false; break; }
        }
    }
    mean[pix] = m;
    spp_used[pix] = used;
    active[pix] = act ? 1 : 0;
}

__global__ __launch_bounds__(TB) void k_finish(const float4* __restrict__ mean, uint32_t npix, float* __restrict__ rgba,
                                               uint8_t* __restrict__ u8) {
    const uint32_t i = blockIdx.x * TB + threadIdx.x;
    if (i >= npix) return;
    const float4 m = mean[i];
    const float v[4] = {m.x, m.y, m.z, 1.0f};
    for (int c = 0; c < 4; ++c) {
        if (rgba) rgba[4 * (size_t)i + c] = v[c];
        // (uint8_t) std::clamp(x * 256, 0, 255), do_render.hpp:122-134
        if (u8) u8[4 * (size_t)i + c] = (uint8_t)fminf(fmaxf(v[c] * 256.0f, 0.0f), 255.0f);
    }
}

__global__ void k_iota(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * TB + threadIdx.x;
    if (i < n) p[i] = i;
}

// compact active pixels into a list
__global__ __launch_bounds__(TB) void k_active_list(const uint8_t* __restrict__ active, uint32_t npix,
                                                    uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
    const uint32_t i = blockIdx.x * TB + threadIdx.x;
    const bool a = i < npix && active[i];
    const uint32_t slot = wave_append(a, count);
    if (a) list[slot] = i;
}

}  // namespace

}  // namespace crt

using namespace crt;

extern "C" int crt_render(crt_scene* scene, const crt_camera* camera, const crt_light* lights, uint32_t n_lights,
                          const crt_render_opts* opts, float* rgba, uint8_t* rgba_u8, int32_t* spp_used,
                          uint32_t flags, void* stream, crt_timing* timing) {
  return guarded([&] {
    CRT_REQUIRE(scene != nullptr, CRT_ERR_INVALID, "scene is NULL");
    check_camera_args(camera);
    CRT_REQUIRE(opts != nullptr, CRT_ERR_INVALID, "render options are NULL");
    CRT_REQUIRE(n_lights <= (uint32_t)MAX_LIGHTS, CRT_ERR_INVALID, "at most 8 lights per render");
    CRT_REQUIRE(n_lights == 0 || lights != nullptr, CRT_ERR_INVALID, "lights is NULL");
    CRT_REQUIRE(opts->max_samples >= 1, CRT_ERR_INVALID, "max_samples must be >= 1");
    CRT_REQUIRE(opts->num_bounces >= 0, CRT_ERR_INVALID, "num_bounces must be >= 0");
    hipStream_t st = (hipStream_t)stream;
    RenderParams P;
    std::memset(&P, 0, sizeof(P));
    P.s = scene->view;
    P.cam = make_cam_view(*camera);
    for (uint32_t l = 0; l < n_lights; ++l) {
        const crt_light& L = lights[l];
        CRT_REQUIRE(L.type >= 0 && L.type <= 2, CRT_ERR_TYPE, "unknown light type");
        LightDev& D = P.lights[l];
        D.type = L.type;
        D.intensity = (float)L.intensity;
        for (int k = 0; k < 3; ++k) D.pos[k] = (float)L.position[k];
        for (int k = 0; k < 9; ++k) D.rot[k] = (float)L.rotation[k];
        D.size[0] = (float)L.size[0];
        D.size[1] = (float)L.size[1];
        if (L.type == CRT_LIGHT_DIRECTIONAL) {
            const double n = std::sqrt(L.direction[0] * L.direction[0] + L.direction[1] * L.direction[1] +
                                       L.direction[2] * L.direction[2]);
            CRT_REQUIRE(n > 0, CRT_ERR_INVALID, "directional light needs a non-zero direction");
            for (int k = 0; k < 3; ++k) D.dir[k] = (float)(L.direction[k] / n);
        }
    }
    P.nl = (int)n_lights;
    P.nb = opts->num_bounces;
    P.jitter = opts->max_samples != 1;
    P.seed = opts->seed;
    P.key_offset = 0;
    P.offset = self_offset(scene);
    const uint32_t npix = (uint32_t)P.cam.w * (uint32_t)P.cam.h;
    const int min_s = std::max(1, std::min(opts->min_samples, opts->max_samples));
    const int max_s = opts->max_samples;
    const size_t budget = (size_t)1 << 22;   // paths per wavefront pass

    // persistent per-call buffers
    DevBuf mean, spp, active, plist, counters;
    mean.alloc((size_t)npix * 16);
    spp.alloc((size_t)npix * 4);
    active.alloc(npix);
    plist.alloc((size_t)npix * 4);
    counters.alloc(64);
    CRT_HIP_CHECK(hipMemsetAsync(mean.p, 0, mean.bytes, st));
    CRT_HIP_CHECK(hipMemsetAsync(spp.p, 0, spp.bytes, st));
    uint32_t* cnt = counters.as<uint32_t>();   // [0] queue, [1] next queue, [2] shadow, [3] live, [4] active list
    hipLaunchKernelGGL(k_iota, dim3(cdiv(npix, TB)), dim3(TB), 0, st, plist.as<uint32_t>(), npix);
    hipEvent_t e0, e1;
    CRT_HIP_CHECK(hipEventCreate(&e0));
    CRT_HIP_CHECK(hipEventCreate(&e1));
    CRT_HIP_CHECK(hipEventRecord(e0, st));
    uint64_t rays = 0, srays = 0;
    uint32_t n_active = npix;
    int s_done = 0;
    DevBuf pstate, queue[2], live, shadow, squeue, samples;
    while (s_done < max_s && n_active > 0) {
        // samples of this pass
        int s_count = (s_done == 0) ? min_s : std::min(max_s - s_done, std::max(1, (int)(budget / n_active)));
        if (s_done == 0 && (size_t)n_active * s_count > budget * 4)
            s_count = std::max(1, (int)(budget * 4 / n_active));
        const uint32_t s_first = (uint32_t)s_done + 1;
        const uint64_t npath = (uint64_t)n_active * s_count;
        CRT_REQUIRE(npath < (1ull << 32), CRT_ERR_INVALID, "render pass too large");
        pstate.reserve(npath * sizeof(PathState));
        queue[0].reserve(npath * 4);
        queue[1].reserve(npath * 4);
        live.reserve(npath * 4);
        shadow.reserve(npath * std::max(1, P.nl) * sizeof(ShadowRec));
        squeue.reserve(npath * std::max(1, P.nl) * 4);
        samples.reserve(npath * 16);
        CRT_HIP_CHECK(hipMemsetAsync(cnt, 0, 64, st));
        hipLaunchKernelGGL(k_raygen, dim3(cdiv(npath, TB)), dim3(TB), 0, st, P, plist.as<uint32_t>(), n_active,
                           s_first, (uint32_t)s_count, pstate.as<PathState>(), queue[0].as<uint32_t>(), cnt + 0,
                           samples.as<float4>());
        int cur = 0;
        for (int b = 0; b < P.nb; ++b) {
            uint32_t* qc = cnt + (cur == 0 ? 0 : 1);
            uint32_t* nc = cnt + (cur == 0 ? 1 : 0);
            CRT_HIP_CHECK(hipMemsetAsync(nc, 0, 4, st));
            CRT_HIP_CHECK(hipMemsetAsync(cnt + 2, 0, 8, st));   // shadow + live counters
            hipLaunchKernelGGL(k_extend, dim3(cdiv(npath, OCTETS_PER_BLOCK)), dim3(TRACE_BLOCK), trace_lds_bytes(P.s),
                               st, P.s, pstate.as<PathState>(), queue[cur].as<uint32_t>(), qc);
            hipLaunchKernelGGL(k_shade, dim3(cdiv(npath, TB)), dim3(TB), 0, st, P, b, pstate.as<PathState>(),
                               queue[cur].as<uint32_t>(), qc, shadow.as<ShadowRec>(), squeue.as<uint32_t>(), cnt + 2,
                               samples.as<float4>(), cnt + 3, live.as<uint32_t>());
            if (P.nl > 0)
                hipLaunchKernelGGL(k_shadow, dim3(cdiv(npath * P.nl, OCTETS_PER_BLOCK)), dim3(TRACE_BLOCK),
                                   trace_lds_bytes(P.s), st, P.s, shadow.as<ShadowRec>(), squeue.as<uint32_t>(),
                                   cnt + 2);
            hipLaunchKernelGGL(k_accumulate, dim3(cdiv(npath, TB)), dim3(TB), 0, st, P, b, pstate.as<PathState>(),
                               live.as<uint32_t>(), cnt + 3, shadow.as<ShadowRec>(), queue[cur ^ 1].as<uint32_t>(), nc,
                               samples.as<float4>());
            uint32_t hc[4];
            CRT_HIP_CHECK(hipMemcpyAsync(hc, cnt, 16, hipMemcpyDeviceToHost, st));
            CRT_HIP_CHECK(hipStreamSynchronize(st));
            rays += hc[cur == 0 ? 0 : 1];
            srays += hc[2];
            cur ^= 1;
            if (hc[cur == 0 ? 0 : 1] == 0) break;   // no path continues
        }
        if (P.nb == 0) {
            // no bounce: every path contributes zero radiance
            CRT_HIP_CHECK(hipMemsetAsync(samples.p, 0, npath * 16, st));
        }
        hipLaunchKernelGGL(k_fold, dim3(cdiv(n_active, TB)), dim3(TB), 0, st, plist.as<uint32_t>(), n_active, s_first,
                           (uint32_t)s_count, min_s, (float)opts->noise_threshold, samples.as<float4>(),
                           mean.as<float4>(), spp.as<int32_t>(), active.as<uint8_t>());
        s_done += s_count;
        if (s_done >= max_s) break;
        // next pass: remaining active pixels
        CRT_HIP_CHECK(hipMemsetAsync(cnt + 4, 0, 4, st));
        hipLaunchKernelGGL(k_active_list, dim3(cdiv(npix, TB)), dim3(TB), 0, st, active.as<uint8_t>(), npix,
                           plist.as<uint32_t>(), cnt + 4);
        uint32_t na = 0;
        CRT_HIP_CHECK(hipMemcpyAsync(&na, cnt + 4, 4, hipMemcpyDeviceToHost, st));
        CRT_HIP_CHECK(hipStreamSynchronize(st));
        // keep the pixel list sorted so that pass-to-pass order is stable
        if (na > 0) {
            std::vector<uint32_t> hl(na);
            CRT_HIP_CHECK(hipMemcpy(hl.data(), plist.p, (size_t)na * 4, hipMemcpyDeviceToHost));
            std::sort(hl.begin(), hl.end());
            CRT_HIP_CHECK(hipMemcpy(plist.p, hl.data(), (size_t)na * 4, hipMemcpyHostToDevice));
        }
        n_active = na;
    }
    const bool dev = (flags & CRT_OUT_DEVICE) != 0;
    DevBuf o_rgba, o_u8;
    float* d_rgba = rgba;
    uint8_t* d_u8 = rgba_u8;
    if (!dev) {
        if (rgba) { o_rgba.alloc((size_t)npix * 16); d_rgba = o_rgba.as<float>(); }
        if (rgba_u8) { o_u8.alloc((size_t)npix * 4); d_u8 = o_u8.as<uint8_t>(); }
    }
    hipLaunchKernelGGL(k_finish, dim3(cdiv(npix, TB)), dim3(TB), 0, st, mean.as<float4>(), npix, d_rgba, d_u8);
    CRT_HIP_CHECK(hipGetLastError());
    CRT_HIP_CHECK(hipEventRecord(e1, st));
    if (!dev) {
        if (rgba) CRT_HIP_CHECK(hipMemcpyAsync(rgba, d_rgba, (size_t)npix * 16, hipMemcpyDeviceToHost, st));
        if (rgba_u8) CRT_HIP_CHECK(hipMemcpyAsync(rgba_u8, d_u8, (size_t)npix * 4, hipMemcpyDeviceToHost, st));
        if (spp_used) CRT_HIP_CHECK(hipMemcpyAsync(spp_used, spp.p, (size_t)npix * 4, hipMemcpyDeviceToHost, st));
    } else if (spp_used) {
        CRT_HIP_CHECK(hipMemcpyAsync(spp_used, spp.p, (size_t)npix * 4, hipMemcpyDeviceToDevice, st));
    }
    CRT_HIP_CHECK(hipStreamSynchronize(st));
    float ms = 0.f;
    CRT_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (timing) {
        std::memset(timing, 0, sizeof(*timing));
        timing->total_ms = ms;
        timing->rays_traced = rays;
        timing->shadow_rays = srays;
    }
  });
}
