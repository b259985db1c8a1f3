// render.hip -- wavefront path tracer for one camera or a batch of cameras
// (crt_render, crt_render_batch).
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
// MI355X structure.  Work is organised in *passes*; a pass is a list of
// segments (camera, pixels, samples [s_first, s_first + s_count)) and every
// (pixel, sample) of every segment is one path; a camera's paths follow the
// 8x8-tile pixel order of tile_order with a pixel's samples adjacent.  The
// default (fused) pass runs, per bounce:
//   k_extend_shade_lane   camera rays of bounce 0 made in the kernel, one
//                         lane per path: closest hit + shading (light
//                         samples -> shadow records, next direction)
//   k_extend_shade_octet  later bounces, an octet traces R paths in turn,
//                         then lane p of the octet shades path p
//   k_shadow_accum        a quad tests the shadow rays of R live paths in
//                         turn (origin triangle first), then lane p adds path
//                         p's clamped direct light and queues it on
// plus k_fold per pixel at the end of the pass (running mean in sample order,
// adaptive stop); a pass with one sample per pixel has no k_fold: each path
// folds its sample into its pixel's mean in the kernel where it ends (one
// path per pixel, so no two paths touch a mean).  Small passes trace camera rays by octets (k_raygen, k_trace,
// k_shade).  The stage pipeline (separate_kernels, and the persistent forms
// whose octets / quads claim `grab` entries per atomic) runs k_raygen, then
// per bounce k_trace / k_extend_lane -> k_shade -> k_tally -> k_shadow_quad ->
// k_accumulate with the same arithmetic, so both give bit-identical images.
// Queues are compacted with a wave ballot + one atomic per wave, and are
// sharded (SHARDS counters per queue, see shard_append) because device-scope
// atomics on one address serialise across the XCDs.  All kernels
// read their sizes from device memory, so a pass can be captured once in a
// hipGraph and replayed (the host only rewrites the segment table between
// passes).  Counters are cleared by a kernel, never by a memset node.
//
// Randomness is counter-based: u01(seed_c, pixel, sample, 64*bounce + dim)
// with a per-camera seed, so a camera rendered inside a batch is bit-for-bit
// the same as the camera rendered alone with that seed.
#include "crt_internal.hpp"
#include "crt_rng.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace crt {
namespace {

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

constexpr int TB = 256;
constexpr uint32_t NO_LIST = 0xFFFFFFFFu;
// default extend mode: one lane per camera ray from this many paths per pass
// (2^20 lanes = 16 waves per SIMD); smaller passes trace camera rays by octets
constexpr uint64_t LANE_PASS_MIN = (uint64_t)1 << 20;

struct LightDev {
    int type;
    float intensity;
    float pos[3];
    float rot[9];
    float size[2];
    float dir[3];
};

struct CamDev {
    CamView v;
    uint64_t seed;
    uint32_t pix_base;     // first global pixel of this camera
    uint32_t jitter;       // max_samples != 1
    uint32_t light_base;   // this camera's lights (scene frame) start here in the light array;
    uint32_t pad_;         // the count is batch-wide
};

struct Segment {
    uint64_t path_begin;   // first path of the segment in the pass
    uint32_t fold_begin;   // first fold slot (pixel) of the segment in the pass
    uint32_t cam;
    uint32_t npix;
    uint32_t s_first;      // 1-based
    uint32_t s_count;
    uint32_t plist_off;    // offset in the pixel list, NO_LIST = pixels 0..npix-1
};

struct PassCtl {
    uint64_t npath;
    uint32_t nfold;
    uint32_t nseg;
};

struct PathState {
    float4 o;          // origin
    float4 d;          // direction
    float4 w;          // weight rgb, pending bounce factor (1 - r1)
    float4 L;          // radiance rgb
    uint32_t cam;
    uint32_t pix;      // camera-local pixel index
    uint32_t sample;   // 1-based
    int32_t slot;      // hit triangle (leaf order) or -1
    float hu, hv;
    uint32_t pad[2];
};

// Path state in HBM as planes (structure of arrays): a wave's paths are
// nearly consecutive ids, so each plane's 16-B loads of a wave touch a few
// contiguous cache lines instead of one line per lane for a 96-B struct.
struct PathPlanes {
    float4* o;     // origin
    float4* d;     // direction
    float4* w;     // weight rgb, pending bounce factor (1 - r1)
    float4* L;     // radiance rgb
    uint4* m;      // camera, pixel, sample (1-based), hit slot (int, -1 = none)
    float2* h;     // hit barycentrics u, v
};
constexpr size_t PATH_BYTES = 5 * sizeof(float4) + sizeof(float2);   // per path over all planes

struct ShadowRec {
    float4 o;      // origin, tmax
    float4 d;      // direction, origin triangle (leaf order, bits; tested first)
    float4 c;      // colour * cos * irradiance, occluded flag (written by k_shadow)
};

// Everything a pass kernel needs, passed by value.
struct Ctx {
    SceneView s;
    int nl;
    int nb;
    int min_samples;
    float thr;
    float offset;
    uint32_t grab;
    uint32_t tile_w;       // camera pixel tiles: 64 pixels, tile_w wide (tile_order)
    int direct_fold;       // one sample per pixel per pass: a path folds its own sample (no k_fold)
    const CamDev* cams;
    const LightDev* lights;
    const Segment* segs;
    const PassCtl* ctl;
    const uint32_t* plist;
    PathPlanes st;
    uint32_t region;       // entries per shard of the sharded buffers below
    uint32_t* queue[2];    // path ids to extend, sharded
    uint32_t* qcnt[2];     // their shard counters
    uint32_t* sorted;      // a bounce's queue ordered by ray-origin cell (coherence sort), dense
    uint32_t* bins;        // SORT_BINS counts, then SORT_BINS cursors
    uint32_t* live;        // paths with a hit at this bounce, sharded
    uint32_t* lcnt;        // their shard counters
    ShadowRec* shadow;     // nl records per live entry: live position p -> p*nl .. p*nl+nl-1
    unsigned long long* tally;   // [0] closest-hit rays, [1] shadow rays
    float4* samples;
    float4* mean;
    int32_t* spp;
    uint8_t* active;
    uint32_t* cnt;   // CNT_* below, then the shard counters
};

// Plain counters (C.cnt[...]).
constexpr int CNT_ACTIVE = 4;        // adaptive: active pixel list
constexpr int CNT_WORK_EXT = 5;      // persistent extension rays: next claim
constexpr int CNT_WORK_SHADOW = 6;   // persistent shadow rays: next claim
constexpr int CNT_SORTED = 7;        // entries of the sorted queue
constexpr int CNT_OVERFLOW = 8;      // a shard filled up (ruled out by the region size; reported)
constexpr int CNT_PLAIN = 32;

// Sharded append buffers.  A wave appends its ballot with one device-scope
// atomic; on a single address those serialise across the 8 XCDs (about 4 ns
// each: 3 ms per bounce of 2^23 paths, more than the shading).  So a queue
// has SHARDS counters, each in its own 128-B line, and shard s owns entries
// [s*region, s*region + region) of the buffer.  Block b appends to shard
// b % SHARDS; a consumer block b reads chunk b / SHARDS of shard b % SHARDS
// and appends what it keeps to the same shard.  A producer over n inputs
// with at most MAX_BLOCK_ITEMS per block fills a shard with at most
// n/SHARDS + MAX_BLOCK_ITEMS entries, and a consumer never appends more to a
// shard than it read from it, so region = ceil(cap/SHARDS) + MAX_BLOCK_ITEMS
// never overflows (the appends check it anyway).
constexpr uint32_t SHARDS = 64;
constexpr uint32_t SHARD_STRIDE = 32;        // counters 128 B apart
constexpr uint32_t MAX_BLOCK_ITEMS = 256;    // inputs of one block, any kernel
constexpr uint32_t NO_POS = 0xFFFFFFFFu;
constexpr int SET_Q0 = 0, SET_Q1 = 1, SET_LIVE = 2, N_SETS = 3;
static_assert(SHARDS == 64, "shard_total reduces one counter per lane of a wave");
static_assert(TB <= (int)MAX_BLOCK_ITEMS && LANE_BLOCK <= (int)MAX_BLOCK_ITEMS && TRACE_BLOCK <= (int)MAX_BLOCK_ITEMS,
              "block inputs exceed MAX_BLOCK_ITEMS");

inline uint32_t shard_region(uint64_t cap) { return cdiv(cap, SHARDS) + MAX_BLOCK_ITEMS; }
inline size_t counter_words() { return CNT_PLAIN + (size_t)N_SETS * SHARDS * SHARD_STRIDE; }

__device__ inline float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline float3 norm3(float3 a) {
    const float inv = 1.0f / sqrtf(dot3(a, a));
    return make_float3(a.x * inv, a.y * inv, a.z * inv);
}

// Wave-aggregated append: slot of this lane (valid only if `want`).
__device__ inline uint32_t wave_append(bool want, uint32_t* counter) {
    const uint64_t m = __ballot(want);
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (m != 0ull && lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader < 0 ? 0 : leader, 64);
    return base + (uint32_t)__popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
}

// Work claim of each lane group (octet or quad; a group's lanes are active
// together) still taking work from a shared counter: `amount` entries per
// group.  One ballot of the groups' first lanes, one atomic by the wave's
// first active lane (always a group's first lane) for all of them, and every
// lane derives its group's first entry from the ballot, so no lane reads
// another lane's register after a branch on the lane index: in a loop with
// group-dependent exits, such a read (a first lane's atomic, then a shuffle
// from it) let the compiler move the other lanes into an inner loop that
// went on reading the exited first lane's register, and never left it.
__device__ inline uint32_t group_grab(uint32_t* counter, bool group_first, int group_base, uint32_t amount = 1u) {
    const uint64_t lead = __ballot(group_first);
    uint32_t base = 0;
    if ((int)(threadIdx.x & 63) == __ffsll((long long)lead) - 1)
        base = atomicAdd(counter, (uint32_t)__popcll(lead) * amount);
    base = __builtin_amdgcn_readfirstlane(base);
    return base + (uint32_t)__popcll(lead & ((1ull << group_base) - 1ull)) * amount;
}

__device__ inline uint32_t block_shard() { return blockIdx.x & (SHARDS - 1); }

// Position in a sharded buffer of this lane's entry, appended to shard s of
// `counters`; NO_POS if !want (or if the shard were full).
__device__ inline uint32_t shard_append(const Ctx& C, bool want, uint32_t* counters, uint32_t s) {
    const uint32_t off = wave_append(want, counters + s * SHARD_STRIDE);
    if (!want) return NO_POS;
    if (off >= C.region) {
        atomicOr(C.cnt + CNT_OVERFLOW, 1u);
        return NO_POS;
    }
    return s * C.region + off;
}

// Input t (< items) of this block in a sharded buffer: chunk blockIdx.x /
// SHARDS of shard blockIdx.x % SHARDS.  False past the shard's count.
__device__ inline bool shard_entry(const Ctx& C, const uint32_t* counters, uint32_t items, uint32_t t,
                                   uint32_t& pos) {
    const uint32_t s = block_shard();
    const uint32_t off = (blockIdx.x / SHARDS) * items + t;
    pos = s * C.region + off;
    return off < counters[s * SHARD_STRIDE];
}

// Extension ray input t of this block: from the sharded queue qsel, or,
// when `sorted`, entry blockIdx.x * items + t of the dense sorted queue.
__device__ inline bool queue_entry(const Ctx& C, int qsel, int sorted, uint32_t items, uint32_t t, uint32_t& k) {
    if (sorted) {
        const uint32_t q = blockIdx.x * items + t;
        if (q >= C.cnt[CNT_SORTED]) return false;
        k = C.sorted[q];
        return true;
    }
    uint32_t pos;
    if (!shard_entry(C, C.qcnt[qsel], items, t, pos)) return false;
    k = C.queue[qsel][pos];
    return true;
}

// Sum of a set of shard counters, in every lane; all 64 lanes of the wave call.
__device__ inline uint32_t shard_total(const uint32_t* counters) {
    uint32_t v = counters[(threadIdx.x & 63) * SHARD_STRIDE];
    for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
    return v;
}

// largest segment index with key(seg) <= x
template <typename F>
__device__ inline uint32_t seg_search(uint32_t nseg, uint64_t x, F key) {
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (key(mid) <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// The same for callers whose lanes hold ascending x (consecutive slots of a
// pass): one wave-uniform search from the first active lane, whose loads
// are scalar, then each lane steps over the few segment starts inside its
// wave; a lane below the first lane's segment (not expected) searches alone.
template <typename F>
__device__ inline uint32_t seg_search_wave(uint32_t nseg, uint64_t x, F key) {
    const uint64_t x0 = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(x >> 32)) << 32) |
                        (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)x);
    uint32_t lo = seg_search(nseg, x0, key);
    if (key(lo) > x) return seg_search(nseg, x, key);
    while (lo + 1 < nseg && key(lo + 1) <= x) ++lo;
    return lo;
}

// Clears the plain counters in `mask` and the shard counter sets in `sets`
// (one block of SHARDS threads).
__global__ void k_clear(uint32_t* cnt, uint32_t mask, uint32_t sets) {
    const uint32_t i = threadIdx.x;
    if (i < (uint32_t)CNT_PLAIN && ((mask >> i) & 1u)) cnt[i] = 0u;
    for (int j = 0; j < N_SETS; ++j)
        if ((sets >> j) & 1u) cnt[CNT_PLAIN + ((size_t)j * SHARDS + i) * SHARD_STRIDE] = 0u;
}

// Camera path k of the pass: FP64 camera ray (jittered), rebased to the
// scene's bounding sphere, FP32 path state.  Returns false (and writes a
// zero sample) when the ray misses the sphere.
// Running-mean step of sample `smp` (1-based) of a pixel
// (src/crt/do_render.hpp:100-108); false when the pixel stops (adaptive
// threshold, from min_samples on).  Shared by k_fold and the paths that fold
// their own sample, with contraction off, so every kernel computes the same
// bits.
__device__ inline bool fold_step(float4& m, float4 x, uint32_t smp, int min_samples, float thr) {
#pragma clang fp contract(off)
    const float inv = 1.0f / (float)smp;
    const float rx = (x.x - m.x) * inv, ry = (x.y - m.y) * inv, rz = (x.z - m.z) * inv;
    m.x += rx; m.y += ry; m.z += rz;
    if ((int)smp >= min_samples) {
        const float noise = sqrtf(rx * rx + ry * ry + rz * rz);
        if (noise < thr) return false;
    }
    return true;
}

// End of path k (camera cam, camera-local pixel pix, sample smp) with
// radiance L: written for k_fold, or, in a pass with one sample per pixel
// (each pixel has exactly one path, so no other path touches its mean),
// folded into the pixel's running mean at once.
__device__ inline void end_path(const Ctx& C, uint64_t k, uint32_t cam, uint32_t pix, uint32_t smp, float4 L) {
    if (!C.direct_fold) {
        C.samples[k] = L;
        return;
    }
    const uint32_t g = C.cams[cam].pix_base + pix;
    float4 m = C.mean[g];
    const bool act = fold_step(m, L, smp, C.min_samples, C.thr);
    C.mean[g] = m;
    C.spp[g] = (int)smp;
    C.active[g] = act ? 1 : 0;
}

__device__ inline void end_path(const Ctx& C, uint64_t k, float4 L) {
    if (!C.direct_fold) {
        C.samples[k] = L;
        return;
    }
    const uint4 m = C.st.m[k];   // cam, pix, sample
    end_path(C, k, m.x, m.y, m.z, L);
}

__device__ inline bool gen_camera_path(const Ctx& C, const PassCtl& ctl, uint64_t k, PathState& ps) {
    bool alive = false;
    {
        const uint32_t si = seg_search_wave(ctl.nseg, k, [&](uint32_t i) { return C.segs[i].path_begin; });
        const Segment sg = C.segs[si];
        const uint32_t local = (uint32_t)(k - sg.path_begin);   // < paths per pass < 2^31
        const uint32_t i = local / sg.s_count;
        const uint32_t smp = sg.s_first + local % sg.s_count;
        const CamDev& cd = C.cams[sg.cam];
        const uint32_t pix = sg.plist_off == NO_LIST ? tile_order(i, (uint32_t)cd.v.w, (uint32_t)cd.v.h, C.tile_w)
                                                     : C.plist[sg.plist_off + i];
        const int px = (int)(pix % (uint32_t)cd.v.w), py = (int)(pix / (uint32_t)cd.v.w);
        double u = px, v = py;
        if (cd.jitter) {
            const uint64_t key = rng_key(cd.seed, pix, smp);
            u += rng_u01k(key, rng_dim(0, RNG_JITTER)) - 0.5;
            v += rng_u01k(key, rng_dim(0, RNG_JITTER + 1)) - 0.5;
        }
        double d[3];
        camera_dir(cd.v, u, v, d);
        double t0 = 0.0;
        ps.cam = sg.cam;
        ps.pix = pix;
        ps.sample = smp;
        ps.slot = -1;
        ps.hu = ps.hv = 0.f;
        const float tp = (float)(2.0 * 3.14159265358979323846);
        ps.w = make_float4(tp, tp, tp, 1.0f);
        ps.L = make_float4(0.f, 0.f, 0.f, 0.f);
        if (rebase_t0(C.s, cd.v.pos, d, t0)) {
            ps.o = make_float4((float)(cd.v.pos[0] + t0 * d[0]), (float)(cd.v.pos[1] + t0 * d[1]),
                               (float)(cd.v.pos[2] + t0 * d[2]), 0.f);
            ps.d = make_float4((float)d[0], (float)d[1], (float)d[2], 0.f);
            alive = true;
        } else {
            ps.o = make_float4(0.f, 0.f, 0.f, 0.f);
            ps.d = make_float4(0.f, 0.f, 1.f, 0.f);
            end_path(C, k, sg.cam, pix, smp, make_float4(0.f, 0.f, 0.f, 0.f));   // camera ray misses the scene
        }
    }
    return alive;
}

__global__ __launch_bounds__(TB) void k_raygen(Ctx C) {
    const uint64_t k = (uint64_t)blockIdx.x * TB + threadIdx.x;
    const PassCtl ctl = *C.ctl;
    bool alive = false;
    if (k < ctl.npath) {
        PathState ps;
        alive = gen_camera_path(C, ctl, k, ps);
        C.st.o[k] = ps.o;
        C.st.d[k] = ps.d;
        C.st.w[k] = ps.w;
        C.st.L[k] = ps.L;
        C.st.m[k] = make_uint4(ps.cam, ps.pix, ps.sample, (uint32_t)ps.slot);
        C.st.h[k] = make_float2(ps.hu, ps.hv);
    }
    const uint32_t pos = shard_append(C, alive, C.qcnt[0], block_shard());
    if (pos != NO_POS) C.queue[0][pos] = (uint32_t)k;
}

// ---- coherence sort (crt_batch_opts.coherence_sort) ----------------------
// Optionally, before the extension rays of a bounce after the first are
// traced, the bounce's queue is reordered by the cell of the ray origin (a
// 16^3 grid over the scene's bounding cube, Morton order), so that the rays
// of a wave start in the same part of the tree.  A counting sort (histogram,
// scan, scatter); paths of one cell keep no particular order, which changes
// nothing in the image: every path's result depends only on its own state
// and RNG keys.  Off by default: the queues already follow the pixel tiles
// of their paths, so origins are coherent without it (256 cameras of the L8
// body, 8 spp, 3 bounces: 191 ms without, 193 ms with).
constexpr uint32_t SORT_BINS = 4096;
constexpr int SORT_TB = 256;
constexpr int SORT_ITEMS = 16;
constexpr uint32_t SORT_TILE = (uint32_t)SORT_TB * SORT_ITEMS;

__device__ inline uint32_t origin_cell(const SceneView& s, float4 o) {
    const float r = (float)s.radius;
    const float inv = 16.0f / (2.0f * r);
    const float lo[3] = {(float)s.center[0] - r, (float)s.center[1] - r, (float)s.center[2] - r};
    const float p[3] = {o.x, o.y, o.z};
    uint32_t key = 0u;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t c = (uint32_t)fminf(fmaxf((p[a] - lo[a]) * inv, 0.0f), 15.0f);
        key |= ((c & 1u) | ((c & 2u) << 2) | ((c & 4u) << 4) | ((c & 8u) << 6)) << a;
    }
    return key;
}

__global__ __launch_bounds__(SORT_TB) void k_sort_zero(uint32_t* __restrict__ bins) {
    for (uint32_t i = threadIdx.x; i < SORT_BINS; i += SORT_TB) bins[i] = 0u;
}

// Both passes read the sharded queue like every consumer (SORT_TILE entries
// per block); the sorted queue is dense.
__global__ __launch_bounds__(SORT_TB) void k_sort_hist(Ctx C, int qsel) {
    __shared__ uint32_t h[SORT_BINS];
    for (uint32_t i = threadIdx.x; i < SORT_BINS; i += SORT_TB) h[i] = 0u;
    __syncthreads();
    for (int j = 0; j < SORT_ITEMS; ++j) {
        uint32_t pos;
        if (shard_entry(C, C.qcnt[qsel], SORT_TILE, (uint32_t)j * SORT_TB + threadIdx.x, pos))
            atomicAdd(&h[origin_cell(C.s, C.st.o[C.queue[qsel][pos]])], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < SORT_BINS; i += SORT_TB)
        if (h[i] != 0u) atomicAdd(&C.bins[i], h[i]);
}

// One block: cursors = exclusive prefix of the counts; *total = their sum.
__global__ __launch_bounds__(1024) void k_sort_scan(uint32_t* __restrict__ bins, uint32_t* __restrict__ total) {
    __shared__ uint32_t part[1024];
    constexpr uint32_t PER = SORT_BINS / 1024;
    const uint32_t t = threadIdx.x;
    uint32_t v[PER], sum = 0u;
    for (uint32_t j = 0; j < PER; ++j) { v[j] = bins[t * PER + j]; sum += v[j]; }
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t x = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t j = 0; j < PER; ++j) { bins[SORT_BINS + t * PER + j] = run; run += v[j]; }
    if (t == 1023) *total = part[t];
}

__global__ __launch_bounds__(SORT_TB) void k_sort_scatter(Ctx C, int qsel) {
    __shared__ uint32_t h[SORT_BINS];
    __shared__ uint32_t at[SORT_BINS];
    for (uint32_t i = threadIdx.x; i < SORT_BINS; i += SORT_TB) h[i] = 0u;
    __syncthreads();
    uint32_t path[SORT_ITEMS], slot[SORT_ITEMS];   // slot = cell << 12 | rank in the block's cell
    for (int j = 0; j < SORT_ITEMS; ++j) {
        uint32_t pos;
        path[j] = 0xFFFFFFFFu;
        slot[j] = 0u;
        if (shard_entry(C, C.qcnt[qsel], SORT_TILE, (uint32_t)j * SORT_TB + threadIdx.x, pos)) {
            path[j] = C.queue[qsel][pos];
            const uint32_t c = origin_cell(C.s, C.st.o[path[j]]);
            slot[j] = (c << 12) | atomicAdd(&h[c], 1u);   // rank < SORT_TILE = 4096
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < SORT_BINS; i += SORT_TB)
        if (h[i] != 0u) at[i] = atomicAdd(&C.bins[SORT_BINS + i], h[i]);
    __syncthreads();
    for (int j = 0; j < SORT_ITEMS; ++j)
        if (path[j] != 0xFFFFFFFFu) C.sorted[at[slot[j] >> 12] + (slot[j] & 0xFFFu)] = path[j];
}

// ---- traversal kernels -------------------------------------------------
__device__ inline void extend_one(const Ctx& C, uint32_t k, uint32_t* sref, float* sd, int sub, int lane_base) {
    const float4 o = C.st.o[k], d = C.st.d[k];
    const RayPre r = make_ray_pre(o.x, o.y, o.z, d.x, d.y, d.z);
    const HitRec h = trace_octet(C.s, r, 0.0f, __int_as_float(0x7f800000), sref, sd, sub, lane_base);
    if (sub == 0) {
        C.st.m[k].w = (uint32_t)h.slot;
        C.st.h[k] = make_float2(h.u, h.v);
    }
}

#define OCTET_SETUP(C)                                                                                      \
    uint32_t* smem = traversal_lds((C).s);                                                                  \
    const int lane = threadIdx.x & 63;                                                                      \
    const int sub = lane & 7, lane_base = lane & ~7, local = threadIdx.x >> 3;                              \
    uint32_t* sref = smem + (size_t)local * (C).s.stack_cap;                                                \
    float* sd = reinterpret_cast<float*>(smem + (size_t)OCTETS_PER_BLOCK * (C).s.stack_cap) +                \
                (size_t)local * (C).s.stack_cap;

// Extension rays: one per octet (closest hit).
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace(Ctx C, int qsel) {
    OCTET_SETUP(C)
    uint32_t k;
    if (!queue_entry(C, qsel, 0, OCTETS_PER_BLOCK, local, k)) return;   // octet-uniform
    extend_one(C, k, sref, sd, sub, lane_base);
}

// Shadow rays: one per quad (any hit, see trace_quad_any).
#define QUAD_SETUP(C)                                                                                       \
    extern __shared__ uint32_t quad_stacks[];                                                               \
    const int lane = threadIdx.x & 63;                                                                      \
    const int qj = lane & (QUAD - 1), qbase = lane & ~(QUAD - 1), qlocal = threadIdx.x / QUAD;              \
    uint32_t* qstk = quad_stacks + (size_t)qlocal * (C).s.stack_cap;

__device__ inline void shadow_quad(const Ctx& C, uint32_t i, uint32_t* stk, int j, int lane_base) {
    const float4 o = C.shadow[i].o, d = C.shadow[i].d;
    const RayPre r = make_ray_pre(o.x, o.y, o.z, d.x, d.y, d.z);
    const bool occ =
        hint_blocks(C.s, r, __float_as_uint(d.w), 0.0f, o.w) || trace_quad_any(C.s, r, 0.0f, o.w, stk, j, lane_base);
    if (j == 0) C.shadow[i].c.w = occ ? 1.0f : 0.0f;
}

// The shadow records of shard s are records s*region*nl .. + count(s)*nl;
// block b tests chunk b / SHARDS of shard b % SHARDS.
__global__ __launch_bounds__(TRACE_BLOCK) void k_shadow_quad(Ctx C) {
    QUAD_SETUP(C)
    const uint32_t s = block_shard(), nl = (uint32_t)C.nl;
    const uint32_t local = (blockIdx.x / SHARDS) * (TRACE_BLOCK / QUAD) + qlocal;
    if (local >= C.lcnt[s * SHARD_STRIDE] * nl) return;   // quad-uniform
    shadow_quad(C, s * C.region * nl + local, qstk, qj, qbase);
}

// Persistent form: quads claim `grab` consecutive records per atomic over
// the first `span` entries' records of every shard.
__global__ __launch_bounds__(TRACE_BLOCK) void k_shadow_quad_persistent(Ctx C, uint32_t span) {
    QUAD_SETUP(C)
    uint32_t* work = C.cnt + CNT_WORK_SHADOW;
    const uint32_t nl = (uint32_t)C.nl, per = span * nl, n = SHARDS * per;
    while (true) {
        const uint32_t base = group_grab(work, qj == 0, qbase, C.grab);   // quad-uniform
        if (base >= n) break;
        const uint32_t end = min(base + C.grab, n);
        for (uint32_t q = base; q < end; ++q) {
            const uint32_t s = q / per, local = q - s * per;
            if (local < C.lcnt[s * SHARD_STRIDE] * nl) shadow_quad(C, s * C.region * nl + local, qstk, qj, qbase);
        }
    }
}

// One extension ray per lane (coherent rays: camera rays).
CRT_LANE_KERNEL void k_extend_lane(Ctx C, int qsel) {
    const LaneStack stk = lane_stack(C.s);
    uint32_t k;
    if (!queue_entry(C, qsel, 0, LANE_BLOCK, threadIdx.x, k)) return;
    const float4 o = C.st.o[k], d = C.st.d[k];
    const RayPre r = make_ray_pre(o.x, o.y, o.z, d.x, d.y, d.z);
    const HitRec h = trace_lane<false, false, true>(C.s, r, 0.0f, __int_as_float(0x7f800000), stk);
    C.st.m[k].w = (uint32_t)h.slot;
    C.st.h[k] = make_float2(h.u, h.v);
}

// Persistent form: octets claim `grab` consecutive queue slots per atomic
// over the first `span` slots of every shard.
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_persistent(Ctx C, int qsel, uint32_t span) {
    OCTET_SETUP(C)
    (void)local;
    uint32_t* work = C.cnt + CNT_WORK_EXT;
    const uint32_t n = SHARDS * span;
    while (true) {
        const uint32_t base = group_grab(work, sub == 0, lane_base, C.grab);   // octet-uniform
        if (base >= n) break;
        const uint32_t end = min(base + C.grab, n);
        for (uint32_t q = base; q < end; ++q) {
            const uint32_t s = q / span, off = q - s * span;
            if (off < C.qcnt[qsel][s * SHARD_STRIDE])
                extend_one(C, C.queue[qsel][s * C.region + off], sref, sd, sub, lane_base);
        }
    }
}

// ---- shading -----------------------------------------------------------
// `key` = rng_key(camera seed, pixel, sample)
__device__ inline void light_sample(const Ctx& C, uint32_t cam, int l, float3 p, uint64_t key, int bounce,
                                    float3& dir, float& tmax, float& irr) {
    const LightDev& L = C.lights[C.cams[cam].light_base + (uint32_t)l];
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
        const float r1 = (float)rng_u01k(key, rng_dim(bounce, RNG_LIGHT + 2 * l));
        const float r2 = (float)rng_u01k(key, rng_dim(bounce, RNG_LIGHT + 2 * l + 1));
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
// The reference's z = -sqrt(1 - x^2 - y^2) equals -sqrt(1 - r1) (x^2 + y^2 =
// r1); it is taken from omr = 1 - r1 formed in FP64, because in FP32 the
// subtraction cancels for grazing samples (x^2 + y^2 -> 1) and would tilt
// the bounce ray by up to ~1e-4 rad.
__device__ inline float3 cosine_sample(float3 n, float r1, float r2, float omr) {
    const float r = sqrtf(r1);
    const float th = r2 * 2.0f * 3.1415926f;
    float sn, cs;
    sincosf(th, &sn, &cs);
    const float x = -r * cs, y = -r * sn;
    const float z = -sqrtf(omr);
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

// ImageUVMap::operator() (src/crt/materials/material.hpp:50-60): nearest
// texel x = (size_t)(u W + 0.5), y = (size_t)(v H + 0.5), channels / 255.
// Indices are clamped to the image, where the reference reads past the end
// of its pixel array for u or v near 1.
__device__ inline float3 tex_lookup(const SceneView& s, int t, float u, float v) {
    const uint4 ti = s.tex_info[t];
    const double fx = (double)(u * (float)ti.y) + 0.5, fy = (double)(v * (float)ti.z) + 0.5;
    const int64_t x = min(fx > 0.0 ? (int64_t)fx : (int64_t)0, (int64_t)ti.y - 1);
    const int64_t y = min(fy > 0.0 ? (int64_t)fy : (int64_t)0, (int64_t)ti.z - 1);
    const uint32_t p = s.texels[(size_t)ti.x + (size_t)y * ti.y + (size_t)x];
    return make_float3((float)(p & 0xFFu) / 255.f, (float)((p >> 8) & 0xFFu) / 255.f,
                       (float)((p >> 16) & 0xFFu) / 255.f);
}

// Shading of a path that hit triangle `slot` (leaf order) with barycentrics
// (hu, hv) at this bounce, arriving along din: one shadow record per light at
// live slot `ls` (records ls*nl .. ls*nl+nl-1) and, unless this is the last
// bounce, the next ray of path k with the pending bounce factor in w.w.
// The instance's material (src/crt/materials/material.hpp:90-184) computes
// the light term and samples the next direction; without a material table
// every instance is ColoredLambertian with its colour.
// MATS: the scene has a material table (crt_scene_set_materials); without one
// the kernels are instantiated with no material code at all (fewer registers).
template <bool MATS>
__device__ inline void shade_hit(const Ctx& C, int bounce, uint32_t k, uint32_t cam, uint32_t pix, uint32_t smp,
                                 int32_t slot, float hu, float hv, uint32_t ls, float3 din) {
    const SceneView& s = C.s;
    const float4 a = s.tris[3 * slot], b = s.tris[3 * slot + 1], c = s.tris[3 * slot + 2];
    const uint32_t orig = __float_as_uint(a.w);
    const uint32_t inst = s.tri_inst[orig];
    const float u = hu, v = hv, w = 1.0f - u - v;
    // geometric normal (inward): normalize((p0 - p1) x (p2 - p0))
    const float3 e1 = make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
    const float3 e2 = make_float3(c.x - a.x, c.y - a.y, c.z - a.z);
    const float3 gn = norm3(cross3(e1, e2));
    float3 sn = gn;
    const bool inst_ok = inst >= 1 && inst <= s.n_instances;
    if (s.vnormals && inst_ok && s.inst_smooth[inst - 1]) {
        const uint32_t i0 = s.tri_vidx[3 * orig], i1 = s.tri_vidx[3 * orig + 1], i2 = s.tri_vidx[3 * orig + 2];
        sn = norm3(make_float3(u * s.vnormals[3 * i1] + v * s.vnormals[3 * i2] + w * s.vnormals[3 * i0],
                               u * s.vnormals[3 * i1 + 1] + v * s.vnormals[3 * i2 + 1] + w * s.vnormals[3 * i0 + 1],
                               u * s.vnormals[3 * i1 + 2] + v * s.vnormals[3 * i2 + 2] + w * s.vnormals[3 * i0 + 2]));
    }
    float3 p = make_float3(u * b.x + v * c.x + w * a.x, u * b.y + v * c.y + w * a.y, u * b.z + v * c.z + w * a.z);
    p = make_float3(p.x - C.offset * gn.x, p.y - C.offset * gn.y, p.z - C.offset * gn.z);
    // material: colour map (constant or texture at the interpolated UV) and,
    // for Blinn-Phong, the ka / kd / ks map
    uint32_t mtype = 0;
    float3 col, kc = make_float3(0.f, 0.f, 0.f);
    float alpha = 0.f;
    if (!MATS) {
        const float* ic = s.inst_color + 3 * (inst_ok ? inst - 1 : 0);
        col = make_float3(ic[0], ic[1], ic[2]);
    } else {
        const MatDev M = s.mats[inst_ok ? inst - 1 : 0];
        mtype = M.type;
        alpha = M.alpha;
        col = make_float3(M.color[0], M.color[1], M.color[2]);
        kc = make_float3(M.coeffs[0], M.coeffs[1], M.coeffs[2]);
        if ((mtype == 1 || mtype == 2) && (M.tex >= 0 || (mtype == 2 && M.spec_tex >= 0))) {
            // unidirectional.hpp:55: u uv1 + v uv2 first, then + w uv0, in float
            float iu = 0.f, iv = 0.f;
            if (s.vuv) {
                const uint32_t j0 = s.tri_vidx[3 * orig], j1 = s.tri_vidx[3 * orig + 1], j2 = s.tri_vidx[3 * orig + 2];
                const float2 t0 = s.vuv[j0], t1 = s.vuv[j1], t2 = s.vuv[j2];
                iu = u * t1.x + v * t2.x + w * t0.x;
                iv = u * t1.y + v * t2.y + w * t0.y;
            }
            if (M.tex >= 0) col = tex_lookup(s, M.tex, iu, iv);
            if (mtype == 2 && M.spec_tex >= 0) kc = tex_lookup(s, M.spec_tex, iu, iv);
        }
    }
    const uint64_t key = rng_key(C.cams[cam].seed, pix, smp);
    for (int l = 0; l < C.nl; ++l) {
        float3 ld;
        float tmax, irr;
        light_sample(C, cam, l, p, key, bounce, ld, tmax, irr);
        float3 lc;
        if (mtype == 3) {
            // MirrorMaterial::compute is 0 whatever the shadow ray finds: no
            // traversal (tmax < 0 rejects every node and triangle)
            lc = make_float3(0.f, 0.f, 0.f);
            tmax = -1.0f;
        } else if (mtype == 2) {
            // TexturedBlinnPhongMaterial::compute (material.hpp:152-160):
            // clamp(map (ka + kd diffuse) + ks spec), spec = pow(n.h, alpha)
            const float diffuse = -dot3(ld, sn);
            const float3 hv3 = norm3(make_float3(ld.x - din.x, ld.y - din.y, ld.z - din.z));
            const float spec = powf(dot3(norm3(sn), hv3), alpha);
            const float sp = kc.z * spec * 1.0f;
            lc = make_float3(fminf(fmaxf(col.x * (kc.x + kc.y * diffuse) + sp, 0.f), 1.f) * irr,
                             fminf(fmaxf(col.y * (kc.x + kc.y * diffuse) + sp, 0.f), 1.f) * irr,
                             fminf(fmaxf(col.z * (kc.x + kc.y * diffuse) + sp, 0.f), 1.f) * irr);
        } else {
            // (Colored|Textured)LambertianMaterial::compute: colour * (-L.n_s), no max(0, .)
            const float ldn = -dot3(ld, sn);
            lc = make_float3(col.x * ldn * irr, col.y * ldn * irr, col.z * ldn * irr);
        }
        ShadowRec r;
        r.o = make_float4(p.x, p.y, p.z, tmax);
        r.d = make_float4(ld.x, ld.y, ld.z, __uint_as_float((uint32_t)slot));
        r.c = make_float4(lc.x, lc.y, lc.z, 0.f);
        C.shadow[ls * (uint32_t)C.nl + (uint32_t)l] = r;
    }
    if (bounce < C.nb - 1) {
        float3 nd;
        float bw;
        if (mtype == 3) {
            // MirrorMaterial::sample (material.hpp:179-182): d - 2 n (d.n), factor 1
            const float dn = dot3(din, sn);
            nd = make_float3(din.x - sn.x * 2.0f * dn, din.y - sn.y * 2.0f * dn, din.z - sn.z * 2.0f * dn);
            bw = 1.0f;
        } else {
            const double r1 = rng_u01k(key, rng_dim(bounce, RNG_BRDF));
            const float r2 = (float)rng_u01k(key, rng_dim(bounce, RNG_BRDF + 1));
            bw = (float)(1.0 - r1);   // the bounce weight (1 - r1), as the reference's float(1 - r1)
            nd = cosine_sample(sn, (float)r1, r2, bw);
        }
        C.st.o[k] = make_float4(p.x, p.y, p.z, 0.f);
        C.st.d[k] = make_float4(nd.x, nd.y, nd.z, 0.f);
        C.st.w[k].w = bw;
    }
}

template <bool MATS>
__global__ __launch_bounds__(TB) void k_shade(Ctx C, int bounce, int qsel) {
    bool has_hit = false;
    uint32_t k = 0;
    uint4 m = make_uint4(0u, 0u, 0u, 0u);
    if (queue_entry(C, qsel, 0, TB, threadIdx.x, k)) {
        m = C.st.m[k];
        if ((int32_t)m.w < 0) end_path(C, k, m.x, m.y, m.z, C.st.L[k]);   // path ends without a hit
        else has_hit = true;
    }
    // live position = shadow records ls*nl .. ls*nl+nl-1
    const uint32_t ls = shard_append(C, has_hit, C.lcnt, block_shard());
    if (ls == NO_POS) return;
    C.live[ls] = k;
    const float2 hb = C.st.h[k];
    float3 din = make_float3(0.f, 0.f, 0.f);
    if (MATS) {   // Blinn-Phong and mirrors use the incoming direction
        const float4 d4 = C.st.d[k];
        din = make_float3(d4.x, d4.y, d4.z);
    }
    shade_hit<MATS>(C, bounce, k, m.x, m.y, m.z, (int32_t)m.w, hb.x, hb.y, ls, din);
}

// Traced rays of a bounce: the extension rays of queue qsel and nl shadow
// rays per live path.  One wave.
__device__ inline void tally_bounce(const Ctx& C, int qsel) {
    const uint32_t ext = shard_total(C.qcnt[qsel]);
    const uint32_t lv = shard_total(C.lcnt);
    if (threadIdx.x == 0) {
        C.tally[0] += ext;
        C.tally[1] += (unsigned long long)lv * (uint32_t)C.nl;
    }
}

__global__ void k_tally(Ctx C, int qsel) { tally_bounce(C, qsel); }

// Direct light of this bounce (sum over the unoccluded lights, clamped to
// [0,1] from bounce 1 on) times the path weight; then the weight takes the
// pending bounce factor.  Returns true when the path goes on to the next
// bounce; after the last bounce the path's sample is written.
__device__ inline bool accumulate_path(const Ctx& C, int bounce, uint32_t k, float lr[3]) {
    if (bounce >= 1)
        for (int j = 0; j < 3; ++j) lr[j] = fminf(fmaxf(lr[j], 0.0f), 1.0f);
    float4 L = C.st.L[k];
    float4 w = C.st.w[k];
    L.x += lr[0] * w.x;
    L.y += lr[1] * w.y;
    L.z += lr[2] * w.z;
    C.st.L[k] = L;
    if (bounce == C.nb - 1) {
        end_path(C, k, L);
        return false;
    }
    w.x *= w.w; w.y *= w.w; w.z *= w.w;
    C.st.w[k] = w;
    return true;
}

__global__ __launch_bounds__(TB) void k_accumulate(Ctx C, int bounce, int next) {
    bool cont = false;
    uint32_t k = 0, q;
    if (shard_entry(C, C.lcnt, TB, threadIdx.x, q)) {
        k = C.live[q];
        float lr[3] = {0.f, 0.f, 0.f};
        for (int l = 0; l < C.nl; ++l) {
            const float4 c = C.shadow[q * (uint32_t)C.nl + (uint32_t)l].c;
            if (c.w == 0.0f) { lr[0] += c.x; lr[1] += c.y; lr[2] += c.z; }
        }
        cont = accumulate_path(C, bounce, k, lr);
    }
    const uint32_t pos = shard_append(C, cont, C.qcnt[next], block_shard());
    if (pos != NO_POS) C.queue[next][pos] = k;
}

// ---- fused bounce kernels ------------------------------------------------
// k_extend_shade_lane: closest hit of a queued path and, in the same lane,
// either its shading (shadow records, next ray) or the end of a path that
// missed.  With gen != 0 it also makes the camera rays of bounce 0 itself
// (one lane per path of the pass, no queue).  k_extend_shade_octet: the
// same with an octet per path.
// k_shadow_accum: one quad per live path tests the path's nl shadow rays
// and accumulates its direct light, then queues the path for the next
// bounce.  The arithmetic is that of k_raygen, k_extend_lane, k_shade,
// k_shadow_quad and k_accumulate in the same order, so images are
// bit-identical to the one-kernel-per-stage pipeline, with less path-state
// traffic and fewer launches per bounce.
template <bool MATS>
__device__ inline void finish_extend(const Ctx& C, int bounce, uint32_t k, const HitRec& h, uint32_t ls,
                                     float3 din) {
    if (h.slot < 0) {
        end_path(C, k, C.st.L[k]);   // path ends without a hit
        return;
    }
    if (ls == NO_POS) return;   // shard full (flagged)
    C.live[ls] = k;
    const uint4 m = C.st.m[k];   // cam, pix, sample
    shade_hit<MATS>(C, bounce, k, m.x, m.y, m.z, h.slot, h.u, h.v, ls, din);
}

template <bool MATS>
CRT_LANE_KERNEL void k_extend_shade_lane(Ctx C, int bounce, int qsel, int gen, int sorted) {
    const LaneStack stk = lane_stack(C.s);
    const uint32_t q = blockIdx.x * LANE_BLOCK + threadIdx.x;
    uint32_t k = 0;
    bool alive = false;
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f), d = o;
    if (gen) {
        const PassCtl ctl = *C.ctl;
        if (q < ctl.npath) {
            k = q;
            PathState ps;
            alive = gen_camera_path(C, ctl, k, ps);
            if (alive) {
                // the ray itself stays in registers; the next bounce's ray is
                // written by the shading
                C.st.w[k] = ps.w;
                C.st.L[k] = ps.L;
                C.st.m[k] = make_uint4(ps.cam, ps.pix, ps.sample, 0u);
                o = ps.o;
                d = ps.d;
            }
        }
        // count of traced camera rays (tally)
        (void)wave_append(alive, C.qcnt[qsel] + block_shard() * SHARD_STRIDE);
    } else if (queue_entry(C, qsel, sorted, LANE_BLOCK, threadIdx.x, k)) {
        o = C.st.o[k];
        d = C.st.d[k];
        alive = true;
    }
    if (!alive) return;
    const RayPre r = make_ray_pre(o.x, o.y, o.z, d.x, d.y, d.z);
    const HitRec h = trace_lane<false, false, true>(C.s, r, 0.0f, __int_as_float(0x7f800000), stk);
    const uint32_t ls = shard_append(C, h.slot >= 0, C.lcnt, block_shard());   // live position
    finish_extend<MATS>(C, bounce, k, h, ls, make_float3(d.x, d.y, d.z));
}

// Octet form for the bounces after the first.  A block takes a chunk of
// 8R queued paths (R = 1..8, by launch size) and lane p shades path p, so a
// wave carries 8R paths through the shading and one atomic per wave appends
// them.  With STEAL its octets take the chunk's paths one at a time from a
// counter in LDS, so an octet that finishes early traces more of them and
// the wave lasts about the chunk's mean work instead of its slowest octet's
// share, and each traced path leaves its hit (triangle, u, v) in LDS for
// its shading lane; without, octet o traces paths oR .. oR + R - 1 in turn
// and keeps each hit in the lane that shades it.  STEAL costs OCTET_RES_BYTES
// of LDS per wave, so it runs only where that still lets a CU hold its 32
// waves (steal_fits).  Those rays leave the body more often than camera rays
// hit it; camera rays traced by octets use k_trace + k_shade.
constexpr size_t OCTET_RES_BYTES = 16 + 3 * 4 * 64;   // path counter + (triangle, u, v) per chunk path

// gfx950: 160 KB of LDS per CU; the bounce kernels ask for 8 waves per SIMD
constexpr size_t LDS_PER_CU = 160 * 1024;
constexpr size_t WAVES_PER_CU = 32;
inline bool steal_fits(size_t stack_bytes, size_t res_bytes) {
    return (stack_bytes + res_bytes) * WAVES_PER_CU <= LDS_PER_CU;
}

template <bool MATS, bool STEAL>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(8)))
void k_extend_shade_octet(Ctx C, int bounce, int qsel, int R, int sorted) {
    OCTET_SETUP(C)
    const uint32_t items = OCTETS_PER_BLOCK * (uint32_t)R;   // <= 64
    uint32_t k_mine = 0;
    HitRec h_mine;
    h_mine.t = 0.f; h_mine.u = 0.f; h_mine.v = 0.f; h_mine.slot = -1;
    float3 d_mine = make_float3(0.f, 0.f, 0.f);
    bool mine = false;
    if constexpr (STEAL) {
        uint32_t* next_path = smem + (size_t)2 * OCTETS_PER_BLOCK * C.s.stack_cap;
        uint32_t* res = next_path + 4;
        if (threadIdx.x == 0) *next_path = 0u;
        __syncthreads();
#pragma unroll 1
        while (true) {
            const uint32_t j = group_grab(next_path, sub == 0, lane_base);   // octet-uniform
            uint32_t k;
            if (j >= items || !queue_entry(C, qsel, sorted, items, j, k)) break;
            const float4 o = C.st.o[k], d = C.st.d[k];
            const RayPre r = make_ray_pre(o.x, o.y, o.z, d.x, d.y, d.z);
            const HitRec h = trace_octet(C.s, r, 0.0f, __int_as_float(0x7f800000), sref, sd, sub, lane_base);
            if (sub == 0) {
                res[3 * j] = (uint32_t)h.slot;
                res[3 * j + 1] = __float_as_uint(h.u);
                res[3 * j + 2] = __float_as_uint(h.v);
            }
        }
        __syncthreads();
        const uint32_t p = threadIdx.x;
        mine = p < items && queue_entry(C, qsel, sorted, items, p, k_mine);
        if (mine) {
            h_mine.slot = (int32_t)res[3 * p];
            h_mine.u = __uint_as_float(res[3 * p + 1]);
            h_mine.v = __uint_as_float(res[3 * p + 2]);
            const float4 d = C.st.d[k_mine];
            d_mine = make_float3(d.x, d.y, d.z);
        }
    } else {
        const uint32_t t0 = (uint32_t)local * (uint32_t)R;
#pragma unroll 1
        for (int p = 0; p < R; ++p) {
            uint32_t k;
            if (!queue_entry(C, qsel, sorted, items, t0 + (uint32_t)p, k)) break;   // octet-uniform
            const float4 o = C.st.o[k], d = C.st.d[k];
            const RayPre r = make_ray_pre(o.x, o.y, o.z, d.x, d.y, d.z);
            const HitRec h = trace_octet(C.s, r, 0.0f, __int_as_float(0x7f800000), sref, sd, sub, lane_base);
            if (sub == p) {   // the hit is octet-uniform
                k_mine = k;
                h_mine = h;
                d_mine = make_float3(d.x, d.y, d.z);
                mine = true;
            }
        }
    }
    const uint32_t ls = shard_append(C, mine && h_mine.slot >= 0, C.lcnt, block_shard());
    if (mine) finish_extend<MATS>(C, bounce, k_mine, h_mine, ls, d_mine);
}

// A block takes a chunk of 16R live paths (R = 1..4, by launch size), tests
// their shadow rays (origin triangle first) by quads, and lane p then adds
// path p's clamped direct light and queues it on.  With STEAL the quads take
// the chunk's paths one at a time from a counter in LDS (as the octet
// bounce kernel does) and leave each path's unoccluded light in LDS; without,
// quad q tests paths qR .. qR + R - 1 in turn.
constexpr size_t QUAD_RES_BYTES = 16 + 3 * 4 * 64;   // path counter + light rgb per chunk path

// The nl shadow rays of live path q: summed colour of the unoccluded ones.
__device__ inline void path_shadows(const Ctx& C, uint32_t q, uint32_t* qstk, int qj, int qbase, float lr[3]) {
    lr[0] = lr[1] = lr[2] = 0.f;
    for (int l = 0; l < C.nl; ++l) {
        const uint32_t i = q * (uint32_t)C.nl + (uint32_t)l;
        const float4 o = C.shadow[i].o, d = C.shadow[i].d;
        const RayPre r = make_ray_pre(o.x, o.y, o.z, d.x, d.y, d.z);
        const bool occ = hint_blocks(C.s, r, __float_as_uint(d.w), 0.0f, o.w) ||
                         trace_quad_any<false>(C.s, r, 0.0f, o.w, qstk, qj, qbase);
        if (!occ) {
            const float4 c = C.shadow[i].c;
            lr[0] += c.x; lr[1] += c.y; lr[2] += c.z;
        }
    }
}

template <bool STEAL>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(8)))
void k_shadow_accum(Ctx C, int bounce, int qsel, int next, int R) {
    QUAD_SETUP(C)
    const uint32_t items = (TRACE_BLOCK / QUAD) * (uint32_t)R;   // <= 64
    if (blockIdx.x == 0) tally_bounce(C, qsel);   // block 0 is one whole wave
    float lr_mine[3] = {0.f, 0.f, 0.f};
    uint32_t k_mine = 0;
    bool mine = false;
    if constexpr (STEAL) {
        uint32_t* next_path = quad_stacks + (size_t)(TRACE_BLOCK / QUAD) * C.s.stack_cap;
        float* res = reinterpret_cast<float*>(next_path + 4);
        if (threadIdx.x == 0) *next_path = 0u;
        __syncthreads();
#pragma unroll 1
        while (true) {
            const uint32_t j = group_grab(next_path, qj == 0, qbase);   // quad-uniform
            uint32_t q;
            if (j >= items || !shard_entry(C, C.lcnt, items, j, q)) break;
            float lr[3];
            path_shadows(C, q, qstk, qj, qbase, lr);
            if (qj == 0) {
                res[3 * j] = lr[0];
                res[3 * j + 1] = lr[1];
                res[3 * j + 2] = lr[2];
            }
        }
        __syncthreads();
        const uint32_t p = threadIdx.x;
        uint32_t q = 0;
        mine = p < items && shard_entry(C, C.lcnt, items, p, q);
        if (mine) {
            k_mine = C.live[q];
            lr_mine[0] = res[3 * p]; lr_mine[1] = res[3 * p + 1]; lr_mine[2] = res[3 * p + 2];
        }
    } else {
        const uint32_t t0 = (uint32_t)qlocal * (uint32_t)R;
#pragma unroll 1
        for (int p = 0; p < R; ++p) {
            uint32_t q;
            if (!shard_entry(C, C.lcnt, items, t0 + (uint32_t)p, q)) break;   // quad-uniform
            float lr[3];
            path_shadows(C, q, qstk, qj, qbase, lr);
            if (qj == p) {
                k_mine = C.live[q];
                lr_mine[0] = lr[0]; lr_mine[1] = lr[1]; lr_mine[2] = lr[2];
                mine = true;
            }
        }
    }
    const bool cont = mine && accumulate_path(C, bounce, k_mine, lr_mine);
    const uint32_t pos = shard_append(C, cont, C.qcnt[next], block_shard());
    if (pos != NO_POS) C.queue[next][pos] = k_mine;
}

__global__ __launch_bounds__(TB) void k_zero_samples(Ctx C) {
    const uint64_t k = (uint64_t)blockIdx.x * TB + threadIdx.x;
    if (k < C.ctl->npath) C.samples[k] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Fold the pass's samples into the running mean (do_render.hpp:100-108).
__global__ __launch_bounds__(TB) void k_fold(Ctx C) {
    const uint32_t t = blockIdx.x * TB + threadIdx.x;
    const PassCtl ctl = *C.ctl;
    if (t >= ctl.nfold) return;
    const uint32_t si = seg_search_wave(ctl.nseg, t, [&](uint32_t i) { return (uint64_t)C.segs[i].fold_begin; });
    const Segment sg = C.segs[si];
    const uint32_t i = t - sg.fold_begin;
    const CamDev& cd = C.cams[sg.cam];
    const uint32_t pix = sg.plist_off == NO_LIST ? tile_order(i, (uint32_t)cd.v.w, (uint32_t)cd.v.h, C.tile_w)
                                                 : C.plist[sg.plist_off + i];
    const uint32_t g = cd.pix_base + pix;
    float4 m = C.mean[g];
    int used = C.spp[g];
    bool act = true;
    const uint64_t base = sg.path_begin + (uint64_t)i * sg.s_count;
    for (uint32_t j = 0; j < sg.s_count; ++j) {
        const uint32_t smp = sg.s_first + j;
        used = (int)smp;
        if (!fold_step(m, C.samples[base + j], smp, C.min_samples, C.thr)) { act = false; break; }
    }
    C.mean[g] = m;
    C.spp[g] = used;
    C.active[g] = act ? 1 : 0;
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

// Active pixels (global ids) of the cameras listed in `mask` (1 byte per pixel).
__global__ __launch_bounds__(TB) void k_active_list(const uint8_t* __restrict__ active, uint32_t npix,
                                                    uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
    const uint32_t i = blockIdx.x * TB + threadIdx.x;
    const bool a = i < npix && active[i];
    const uint32_t slot = wave_append(a, count);
    if (a) list[slot] = i;
}

__global__ __launch_bounds__(TB) void k_fill_u8(uint8_t* p, uint32_t n, uint8_t v) {
    const uint32_t i = blockIdx.x * TB + threadIdx.x;
    if (i < n) p[i] = v;
}

// Launch the kernels of one pass.  Grids are sized for `cap_paths` /
// `cap_fold` so that the same sequence can be captured once and replayed.
void launch_shade(const Ctx& C, bool mats, uint32_t grid, int b, int cur, hipStream_t st) {
    if (mats)
        hipLaunchKernelGGL(k_shade<true>, dim3(grid), dim3(TB), 0, st, C, b, cur);
    else
        hipLaunchKernelGGL(k_shade<false>, dim3(grid), dim3(TB), 0, st, C, b, cur);
}

void launch_pass(const Ctx& C, uint64_t cap_paths, uint32_t cap_fold, const crt_batch_opts& B, int num_cus,
                 hipStream_t st) {
    const bool mats = C.s.mats != nullptr;
    const size_t lds = trace_lds_bytes(C.s);
    const uint32_t gp = std::max(1u, cdiv(cap_paths, TB));
    const uint32_t go = std::max(1u, cdiv(cap_paths, OCTETS_PER_BLOCK));
    const uint32_t gsq = std::max(1u, cdiv(cap_paths * std::max(1, C.nl), TRACE_BLOCK / QUAD));
    const size_t qlds = quad_lds_bytes(C.s);
    const uint32_t gpers = std::max(1u, B.blocks_per_cu) * (uint32_t)num_cus;
    // entries any shard can hold in this pass, and the grid of a consumer of
    // a sharded buffer with `items` inputs per block
    const uint32_t span = std::min<uint32_t>(C.region, shard_region(cap_paths));
    const auto gsh = [&](uint64_t items) { return SHARDS * cdiv(span, items); };
    const uint32_t nl = (uint32_t)std::max(1, C.nl);
    constexpr uint32_t QPB = TRACE_BLOCK / QUAD;
    constexpr uint32_t ALL_SETS = (1u << N_SETS) - 1u;
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(SHARDS), 0, st, C.cnt,
                       (1u << CNT_WORK_EXT) | (1u << CNT_WORK_SHADOW) | (1u << CNT_SORTED), ALL_SETS);
    // extend_mode 0 (default): lanes for the coherent camera rays of passes
    // large enough to fill the GPU with one lane per ray, octets otherwise
    const uint32_t em = B.extend_mode != 0 ? B.extend_mode : (cap_paths >= LANE_PASS_MIN ? 3u : 1u);
    const auto lane_at = [&](int b) { return em == 2 || (em == 3 && b == 0); };
    const bool fused = B.separate_kernels != 1 && !B.persistent;
    if (fused) {
        const bool gen = C.nb > 0 && lane_at(0);
        if (!gen && C.nb > 0) hipLaunchKernelGGL(k_raygen, dim3(gp), dim3(TB), 0, st, C);
        // paths per octet / quad: enough groups for ~32k waves, at most one path per lane
        const int ro = B.rays_per_octet != 0 ? (int)B.rays_per_octet
                       : cap_paths >= (1u << 21) ? 8 : cap_paths >= (1u << 20) ? 4 : cap_paths >= (1u << 19) ? 2 : 1;
        const int rq = B.rays_per_quad != 0 ? (int)B.rays_per_quad
                       : cap_paths >= (1u << 21) ? 4 : cap_paths >= (1u << 20) ? 2 : 1;
        const bool sort = B.coherence_sort == 1;
        for (int b = 0; b < C.nb; ++b) {
            const int cur = b & 1, nxt = cur ^ 1;
            hipLaunchKernelGGL(k_clear, dim3(1), dim3(SHARDS), 0, st, C.cnt, 0u, (1u << nxt) | (1u << SET_LIVE));
            const bool sorted = sort && b > 0 && (lane_at(b) || B.separate_kernels != 2);
            if (sorted) {
                const uint32_t gs = gsh(SORT_TILE);
                hipLaunchKernelGGL(k_sort_zero, dim3(1), dim3(SORT_TB), 0, st, C.bins);
                hipLaunchKernelGGL(k_sort_hist, dim3(gs), dim3(SORT_TB), 0, st, C, cur);
                hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, st, C.bins, C.cnt + CNT_SORTED);
                hipLaunchKernelGGL(k_sort_scatter, dim3(gs), dim3(SORT_TB), 0, st, C, cur);
            }
            // a sorted queue is dense (as many blocks as its capacity needs); a
            // sharded one takes SHARDS chunks of each shard; bounce 0 of the
            // lane form makes its rays (one lane per path)
            const auto grid = [&](uint32_t items) {
                return std::max(1u, sorted || (gen && b == 0) ? cdiv(cap_paths, items) : gsh(items));
            };
            const dim3 gl(grid(LANE_BLOCK));
            const dim3 goct(grid(OCTETS_PER_BLOCK * ro));
            if (lane_at(b)) {
                if (mats)
                    hipLaunchKernelGGL(k_extend_shade_lane<true>, gl, dim3(LANE_BLOCK), lane_lds_bytes(C.s), st, C, b,
                                       cur, (int)(gen && b == 0), (int)sorted);
                else
                    hipLaunchKernelGGL(k_extend_shade_lane<false>, gl, dim3(LANE_BLOCK), lane_lds_bytes(C.s), st, C,
                                       b, cur, (int)(gen && b == 0), (int)sorted);
            } else if (b > 0 && B.separate_kernels != 2) {
                const bool steal = B.fixed_share == 0 && steal_fits(lds, OCTET_RES_BYTES);
                const size_t olds = steal ? lds + OCTET_RES_BYTES : lds;
                if (mats && steal)
                    hipLaunchKernelGGL((k_extend_shade_octet<true, true>), goct, dim3(TRACE_BLOCK), olds, st, C, b, cur,
                                       ro, (int)sorted);
                else if (mats)
                    hipLaunchKernelGGL((k_extend_shade_octet<true, false>), goct, dim3(TRACE_BLOCK), olds, st, C, b, cur,
                                       ro, (int)sorted);
                else if (steal)
                    hipLaunchKernelGGL((k_extend_shade_octet<false, true>), goct, dim3(TRACE_BLOCK), olds, st, C, b,
                                       cur, ro, (int)sorted);
                else
                    hipLaunchKernelGGL((k_extend_shade_octet<false, false>), goct, dim3(TRACE_BLOCK), olds, st, C, b,
                                       cur, ro, (int)sorted);
            } else {
                hipLaunchKernelGGL(k_trace, dim3(gsh(OCTETS_PER_BLOCK)), dim3(TRACE_BLOCK), lds, st, C, cur);
                launch_shade(C, mats, gsh(TB), b, cur, st);
            }
            if (B.fixed_share == 0 && steal_fits(qlds, QUAD_RES_BYTES)) {
                hipLaunchKernelGGL(k_shadow_accum<true>, dim3(gsh(QPB * rq)), dim3(TRACE_BLOCK), qlds + QUAD_RES_BYTES, st,
                                   C, b, cur, nxt, rq);
            } else {
                // with fixed shares the bounce-0 queue, which holds every
                // camera-ray hit, runs best at half the paths per quad, twice
                // the waves (C4: 5.0 -> 4.55 ms per pass; a 6.3M-triangle
                // geodesic render 9.9 -> 9.5 ms); the later, sparse bounces
                // keep rq
                const int rqb = (b == 0 && B.rays_per_quad == 0) ? std::max(1, rq / 2) : rq;
                hipLaunchKernelGGL(k_shadow_accum<false>, dim3(gsh(QPB * rqb)), dim3(TRACE_BLOCK), qlds, st, C, b, cur,
                                   nxt, rqb);
            }
        }
        if (C.nb == 0) hipLaunchKernelGGL(k_zero_samples, dim3(gp), dim3(TB), 0, st, C);
        if (!C.direct_fold) hipLaunchKernelGGL(k_fold, dim3(std::max(1u, cdiv(cap_fold, TB))), dim3(TB), 0, st, C);
        return;
    }
    hipLaunchKernelGGL(k_raygen, dim3(gp), dim3(TB), 0, st, C);
    for (int b = 0; b < C.nb; ++b) {
        const int cur = b & 1, nxt = cur ^ 1;
        hipLaunchKernelGGL(k_clear, dim3(1), dim3(SHARDS), 0, st, C.cnt,
                           (1u << CNT_WORK_EXT) | (1u << CNT_WORK_SHADOW), (1u << nxt) | (1u << SET_LIVE));
        if (lane_at(b))
            hipLaunchKernelGGL(k_extend_lane, dim3(gsh(LANE_BLOCK)), dim3(LANE_BLOCK), lane_lds_bytes(C.s), st, C,
                               cur);
        else if (B.persistent)
            hipLaunchKernelGGL(k_trace_persistent, dim3(std::min(gpers, go)), dim3(TRACE_BLOCK), lds, st, C, cur,
                               span);
        else
            hipLaunchKernelGGL(k_trace, dim3(gsh(OCTETS_PER_BLOCK)), dim3(TRACE_BLOCK), lds, st, C, cur);
        launch_shade(C, mats, gsh(TB), b, cur, st);
        hipLaunchKernelGGL(k_tally, dim3(1), dim3(64), 0, st, C, cur);
        if (C.nl > 0) {
            if (B.persistent)
                hipLaunchKernelGGL(k_shadow_quad_persistent, dim3(std::min(gpers, gsq)), dim3(TRACE_BLOCK), qlds, st,
                                   C, span);
            else
                hipLaunchKernelGGL(k_shadow_quad, dim3(SHARDS * cdiv((uint64_t)span * nl, QPB)), dim3(TRACE_BLOCK),
                                   qlds, st, C);
        }
        hipLaunchKernelGGL(k_accumulate, dim3(gsh(TB)), dim3(TB), 0, st, C, b, nxt);
    }
    if (C.nb == 0) hipLaunchKernelGGL(k_zero_samples, dim3(gp), dim3(TB), 0, st, C);
    if (!C.direct_fold) hipLaunchKernelGGL(k_fold, dim3(std::max(1u, cdiv(cap_fold, TB))), dim3(TB), 0, st, C);
}

}  // namespace

// Grow-only device buffers of one render caller (a scene's default
// workspace, or one per serving worker so that workers never allocate).
struct RenderWorkspace {
    DevBuf d_cams, d_lights, d_segs, d_ctl, d_plist, mean, spp, active, counters, tally, pstate, q0, q1, qs, bins,
        live, shadow, samples, o_rgba, o_u8;
    size_t bytes() const {
        return d_cams.bytes + d_lights.bytes + d_segs.bytes + d_ctl.bytes + d_plist.bytes + mean.bytes + spp.bytes +
               active.bytes + counters.bytes + tally.bytes + pstate.bytes + q0.bytes + q1.bytes + qs.bytes +
               bins.bytes + live.bytes +
               shadow.bytes + samples.bytes + o_rgba.bytes + o_u8.bytes;
    }
};

RenderWorkspace* new_render_workspace() { return new RenderWorkspace(); }
void delete_render_workspace(RenderWorkspace* w) { delete w; }

size_t render_workspace_bytes(const RenderWorkspace* w) { return w ? w->bytes() : 0; }

void check_batch_opts(const crt_batch_opts& B) {
    CRT_REQUIRE(B.extend_mode <= 3, CRT_ERR_INVALID, "extend_mode must be 0..3");
    CRT_REQUIRE(B.separate_kernels <= 2, CRT_ERR_INVALID, "separate_kernels must be 0..2");
    CRT_REQUIRE(B.coherence_sort <= 1, CRT_ERR_INVALID, "coherence_sort must be 0 or 1");
    CRT_REQUIRE(B.rays_per_octet <= 8, CRT_ERR_INVALID, "rays_per_octet must be 0 (by pass size) or 1..8");
    CRT_REQUIRE(B.rays_per_quad <= 4, CRT_ERR_INVALID, "rays_per_quad must be 0 (by pass size) or 1..4");
    CRT_REQUIRE(B.fixed_share <= 1, CRT_ERR_INVALID, "fixed_share must be 0 or 1");
    CRT_REQUIRE(B.tile_width == 0 || B.tile_width == 4 || B.tile_width == 8 || B.tile_width == 16 ||
                    B.tile_width == 32,
                CRT_ERR_INVALID, "tile_width must be 0 (8), 4, 8, 16 or 32");
}

void check_light(const crt_light& L) {
    CRT_REQUIRE(L.type >= 0 && L.type <= 2, CRT_ERR_TYPE, "unknown light type");
    bool fin = std::isfinite(L.intensity) && std::isfinite(L.size[0]) && std::isfinite(L.size[1]);
    for (int k = 0; k < 3; ++k) fin = fin && std::isfinite(L.position[k]) && std::isfinite(L.direction[k]);
    for (int k = 0; k < 9; ++k) fin = fin && std::isfinite(L.rotation[k]);
    CRT_REQUIRE(fin, CRT_ERR_INVALID, "light parameters must be finite");
    if (L.type == CRT_LIGHT_DIRECTIONAL)
        CRT_REQUIRE(L.direction[0] != 0.0 || L.direction[1] != 0.0 || L.direction[2] != 0.0, CRT_ERR_INVALID,
                    "directional light needs a non-zero direction");
}

// Paths per pass: 2^23, fewer with many lights so that the shadow records
// (48 B per path and light) stay within 4 GiB; a camera's pixels always fit.
uint64_t pass_budget(int n_lights) {
    const uint64_t per = (uint64_t)std::max(n_lights, 1) * sizeof(ShadowRec);
    return std::max<uint64_t>((uint64_t)1 << 16, std::min<uint64_t>((uint64_t)1 << 23, ((uint64_t)4 << 30) / per));
}

size_t render_bytes_estimate(uint64_t pixels, int spp, int n_lights) {
    const uint64_t paths = std::max<uint64_t>(pixels, std::min<uint64_t>(pixels * (uint64_t)std::max(spp, 1),
                                                                         pass_budget(n_lights)));
    const uint64_t slots = (uint64_t)SHARDS * shard_region(paths);   // sharded queues, live list, shadow records
    return (size_t)(paths * (PATH_BYTES + 4 + 16) + slots * (12 + std::max(n_lights, 1) * sizeof(ShadowRec)) +
                    pixels * (16 + 4 + 1 + 4 + 16 + 4));
}

int render_batch(crt_scene* scene, uint32_t n_cams, const crt_camera* cams, const int32_t* max_samples,
                 const uint64_t* seeds, const crt_light* lights, uint32_t n_lights, bool per_camera_lights,
                 const crt_render_opts* opts, const crt_batch_opts* bopts, float* rgba, uint8_t* rgba_u8,
                 int32_t* spp_used, uint32_t flags, void* stream, crt_timing* timing, RenderWorkspace* wsp) {
    std::unique_ptr<RenderWorkspace> own;
    if (!wsp) {
        own.reset(new RenderWorkspace());
        wsp = own.get();
    }
    RenderWorkspace& W = *wsp;
    CRT_REQUIRE(scene != nullptr, CRT_ERR_INVALID, "scene is NULL");
    CRT_REQUIRE(n_cams >= 1 && cams != nullptr, CRT_ERR_INVALID, "no cameras");
    CRT_REQUIRE(opts != nullptr, CRT_ERR_INVALID, "render options are NULL");
    CRT_REQUIRE(n_lights <= (1u << 20), CRT_ERR_INVALID, "too many lights");
    CRT_REQUIRE(n_lights == 0 || lights != nullptr, CRT_ERR_INVALID, "lights is NULL");
    CRT_REQUIRE(opts->num_bounces >= 0, CRT_ERR_INVALID, "num_bounces must be >= 0");
    crt_batch_opts B{};
    if (bopts) B = *bopts;
    if (B.grab == 0) B.grab = 4;
    if (B.blocks_per_cu == 0) B.blocks_per_cu = 8;
    if (B.tile_width == 0) B.tile_width = 8;
    check_batch_opts(B);
    hipStream_t st = (hipStream_t)stream;
    // stream capture is not allowed on the legacy null stream: graph replay
    // on a null stream runs on a private stream (this call synchronises at
    // its end anyway)
    struct OwnStream {
        hipStream_t s = nullptr;
        ~OwnStream() {
            if (s) (void)hipStreamDestroy(s);
        }
    } own_stream;
    if (B.use_graph && st == nullptr) {
        CRT_HIP_CHECK(hipStreamCreateWithFlags(&own_stream.s, hipStreamNonBlocking));
        st = own_stream.s;
    }
    int dev = 0, num_cus = 0;
    CRT_HIP_CHECK(hipGetDevice(&dev));
    CRT_HIP_CHECK(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev));

    // ---- cameras ----
    std::vector<CamDev> hc(n_cams);
    std::vector<int> spp_c(n_cams);
    uint64_t total_pix = 0;
    for (uint32_t c = 0; c < n_cams; ++c) {
        check_camera_args(&cams[c]);
        hc[c].v = make_cam_view(cams[c], scene);
        spp_c[c] = max_samples ? max_samples[c] : opts->max_samples;
        CRT_REQUIRE(spp_c[c] >= 1, CRT_ERR_INVALID, "max_samples must be >= 1");
        hc[c].seed = seeds ? seeds[c] : opts->seed + c;
        hc[c].jitter = spp_c[c] != 1;
        hc[c].pix_base = (uint32_t)total_pix;
        total_pix += (uint64_t)hc[c].v.w * hc[c].v.h;
        CRT_REQUIRE(total_pix < (1ull << 31), CRT_ERR_INVALID, "too many pixels in one batch");
    }
    const uint32_t P = (uint32_t)total_pix;
    const int min_s = std::max(1, opts->min_samples);
    bool adaptive = opts->noise_threshold > 0.0;
    {
        bool any = false;
        for (uint32_t c = 0; c < n_cams; ++c) any |= spp_c[c] > min_s;
        adaptive = adaptive && any;
    }

    Ctx C;
    std::memset(&C, 0, sizeof(C));
    C.s = scene->view;
    // lights: one slice per camera (or one shared slice), in the scene frame
    const uint32_t n_slices = per_camera_lights ? n_cams : 1u;
    std::vector<LightDev> hl((size_t)n_slices * n_lights);
    for (uint32_t c = 0; c < n_cams; ++c) hc[c].light_base = per_camera_lights ? c * n_lights : 0u;
    for (uint32_t c = 0; c < n_slices; ++c) {
        for (uint32_t l = 0; l < n_lights; ++l) {
            const crt_light& L = lights[c * n_lights + l];
            check_light(L);
            LightDev& D = hl[(size_t)c * n_lights + l];
            D.type = L.type;
            D.intensity = (float)L.intensity;
            for (int k = 0; k < 3; ++k) D.pos[k] = (float)(L.position[k] - scene->view.origin[k]);   // scene frame
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
    }
    C.nl = (int)n_lights;
    C.nb = opts->num_bounces;
    C.min_samples = min_s;
    C.thr = (float)opts->noise_threshold;
    C.offset = self_offset(scene);
    C.grab = B.grab;
    C.tile_w = B.tile_width;

    // ---- pass planning ----
    const uint64_t budget = pass_budget(C.nl);   // paths per pass
    int S = (int)B.samples_per_pass;
    if (S <= 0) {
        int smax = 1;
        for (int v : spp_c) smax = std::max(smax, v);
        S = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)smax, budget / std::max<uint64_t>(P, 1)));
    }
    // one sample per pixel per pass (S = 1 makes every segment's s_count 1):
    // paths fold their own samples
    C.direct_fold = (S == 1 && C.nb > 0) ? 1 : 0;
    // capacity: the largest pass we will plan
    const uint64_t cap_paths = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)P * S, std::max(budget, (uint64_t)P)));

    DevBuf &d_cams = W.d_cams, &d_segs = W.d_segs, &d_ctl = W.d_ctl, &d_plist = W.d_plist, &mean = W.mean,
           &spp = W.spp, &active = W.active, &counters = W.counters, &tally = W.tally, &pstate = W.pstate,
           &q0 = W.q0, &q1 = W.q1, &live = W.live, &shadow = W.shadow, &samples = W.samples;
    d_cams.reserve(sizeof(CamDev) * n_cams);
    W.d_lights.reserve(std::max<size_t>(hl.size(), 1) * sizeof(LightDev));
    d_segs.reserve(sizeof(Segment) * (n_cams + 1));
    d_ctl.reserve(sizeof(PassCtl));
    d_plist.reserve((size_t)P * 4);
    mean.reserve((size_t)P * 16);
    spp.reserve((size_t)P * 4);
    active.reserve(P);
    const uint32_t region = shard_region(cap_paths);
    const uint64_t slots = (uint64_t)SHARDS * region;
    // shadow records are indexed in 32 bits
    CRT_REQUIRE(slots * (uint64_t)std::max(1, C.nl) < (1ull << 32), CRT_ERR_INVALID, "too many lights for one pass");
    counters.reserve(counter_words() * 4);
    pstate.reserve(cap_paths * PATH_BYTES);
    q0.reserve(slots * 4);
    q1.reserve(slots * 4);
    W.qs.reserve(cap_paths * 4);
    W.bins.reserve(2 * SORT_BINS * 4);
    live.reserve(slots * 4);
    shadow.reserve(slots * std::max(1, C.nl) * sizeof(ShadowRec));
    tally.reserve(16);
    if (!C.direct_fold) samples.reserve(cap_paths * 16);
    CRT_HIP_CHECK(hipMemcpyAsync(d_cams.p, hc.data(), sizeof(CamDev) * n_cams, hipMemcpyHostToDevice, st));
    if (!hl.empty())
        CRT_HIP_CHECK(hipMemcpyAsync(W.d_lights.p, hl.data(), hl.size() * sizeof(LightDev), hipMemcpyHostToDevice, st));
    CRT_HIP_CHECK(hipMemsetAsync(mean.p, 0, (size_t)P * 16, st));
    CRT_HIP_CHECK(hipMemsetAsync(spp.p, 0, (size_t)P * 4, st));
    CRT_HIP_CHECK(hipMemsetAsync(tally.p, 0, 16, st));
    hipLaunchKernelGGL(k_fill_u8, dim3(std::max(1u, cdiv(P, TB))), dim3(TB), 0, st, active.as<uint8_t>(), P, (uint8_t)1);
    C.cams = d_cams.as<CamDev>();
    C.lights = W.d_lights.as<LightDev>();
    C.segs = d_segs.as<Segment>();
    C.ctl = d_ctl.as<PassCtl>();
    C.plist = d_plist.as<uint32_t>();
    {
        float4* base = pstate.as<float4>();
        C.st.o = base;
        C.st.d = base + cap_paths;
        C.st.w = base + 2 * cap_paths;
        C.st.L = base + 3 * cap_paths;
        C.st.m = reinterpret_cast<uint4*>(base + 4 * cap_paths);
        C.st.h = reinterpret_cast<float2*>(base + 5 * cap_paths);
    }
    C.region = region;
    C.queue[0] = q0.as<uint32_t>();
    C.queue[1] = q1.as<uint32_t>();
    C.sorted = W.qs.as<uint32_t>();
    C.bins = W.bins.as<uint32_t>();
    C.live = live.as<uint32_t>();
    C.shadow = shadow.as<ShadowRec>();
    C.tally = tally.as<unsigned long long>();
    C.samples = samples.as<float4>();
    C.mean = mean.as<float4>();
    C.spp = spp.as<int32_t>();
    C.active = active.as<uint8_t>();
    C.cnt = counters.as<uint32_t>();
    C.qcnt[0] = C.cnt + CNT_PLAIN + (size_t)SET_Q0 * SHARDS * SHARD_STRIDE;
    C.qcnt[1] = C.cnt + CNT_PLAIN + (size_t)SET_Q1 * SHARDS * SHARD_STRIDE;
    C.lcnt = C.cnt + CNT_PLAIN + (size_t)SET_LIVE * SHARDS * SHARD_STRIDE;
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(SHARDS), 0, st, C.cnt, 1u << CNT_OVERFLOW, 0u);

    hipEvent_t e0, e1;
    CRT_HIP_CHECK(hipEventCreate(&e0));
    CRT_HIP_CHECK(hipEventCreate(&e1));
    CRT_HIP_CHECK(hipEventRecord(e0, st));

    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    const bool use_graph = B.use_graph && !adaptive;
    std::vector<int> done(n_cams, 0);
    std::vector<uint32_t> n_act(n_cams);
    std::vector<uint32_t> list_off(n_cams, NO_LIST);
    for (uint32_t c = 0; c < n_cams; ++c) n_act[c] = (uint32_t)hc[c].v.w * hc[c].v.h;
    uint64_t passes = 0;
    while (true) {
        // segments of this pass
        std::vector<Segment> segs;
        uint64_t npath = 0;
        uint32_t nfold = 0;
        for (uint32_t c = 0; c < n_cams; ++c) {
            if (done[c] >= spp_c[c] || n_act[c] == 0) continue;
            Segment sg;
            sg.path_begin = npath;
            sg.fold_begin = nfold;
            sg.cam = c;
            sg.npix = n_act[c];
            sg.s_first = (uint32_t)done[c] + 1;
            sg.s_count = (uint32_t)std::min(S, spp_c[c] - done[c]);
            sg.plist_off = list_off[c];
            if (npath + (uint64_t)sg.npix * sg.s_count > cap_paths) {
                if (!segs.empty()) break;                 // this camera waits for the next pass
                sg.s_count = (uint32_t)std::max<uint64_t>(1, cap_paths / sg.npix);
                CRT_REQUIRE((uint64_t)sg.npix * sg.s_count <= cap_paths, CRT_ERR_INVALID, "camera too large");
            }
            npath += (uint64_t)sg.npix * sg.s_count;
            nfold += sg.npix;
            segs.push_back(sg);
        }
        if (segs.empty()) break;
        PassCtl ctl{npath, nfold, (uint32_t)segs.size()};
        CRT_HIP_CHECK(hipMemcpyAsync(d_segs.p, segs.data(), sizeof(Segment) * segs.size(), hipMemcpyHostToDevice, st));
        CRT_HIP_CHECK(hipMemcpyAsync(d_ctl.p, &ctl, sizeof(ctl), hipMemcpyHostToDevice, st));
        CRT_HIP_CHECK(hipStreamSynchronize(st));   // host tables are reused next pass
        if (use_graph) {
            if (!gexec) {
                CRT_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                launch_pass(C, cap_paths, P, B, num_cus, st);
                CRT_HIP_CHECK(hipStreamEndCapture(st, &graph));
                CRT_HIP_CHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
            }
            CRT_HIP_CHECK(hipGraphLaunch(gexec, st));
        } else {
            launch_pass(C, npath, nfold, B, num_cus, st);
        }
        CRT_HIP_CHECK(hipGetLastError());
        ++passes;
        for (const Segment& sg : segs) done[sg.cam] += (int)sg.s_count;
        if (adaptive) {
            // remaining active pixels, per camera, in increasing pixel order
            CRT_HIP_CHECK(hipMemsetAsync(C.cnt + CNT_ACTIVE, 0, 4, st));
            hipLaunchKernelGGL(k_active_list, dim3(std::max(1u, cdiv(P, TB))), dim3(TB), 0, st, C.active, P,
                               d_plist.as<uint32_t>(), C.cnt + CNT_ACTIVE);
            uint32_t na = 0;
            CRT_HIP_CHECK(hipMemcpyAsync(&na, C.cnt + CNT_ACTIVE, 4, hipMemcpyDeviceToHost, st));
            CRT_HIP_CHECK(hipStreamSynchronize(st));
            std::vector<uint32_t> hl(na);
            if (na) CRT_HIP_CHECK(hipMemcpy(hl.data(), d_plist.p, (size_t)na * 4, hipMemcpyDeviceToHost));
            std::sort(hl.begin(), hl.end());
            std::vector<uint32_t> local(na);
            size_t pos = 0;
            for (uint32_t c = 0; c < n_cams; ++c) {
                const uint32_t lo = hc[c].pix_base, hi = lo + (uint32_t)hc[c].v.w * hc[c].v.h;
                const size_t start = pos;
                while (pos < hl.size() && hl[pos] < hi) { local[pos] = hl[pos] - lo; ++pos; }
                n_act[c] = (uint32_t)(pos - start);
                list_off[c] = (uint32_t)start;
            }
            if (na) CRT_HIP_CHECK(hipMemcpy(d_plist.p, local.data(), (size_t)na * 4, hipMemcpyHostToDevice));
        }
    }
    const bool dev_out = (flags & CRT_OUT_DEVICE) != 0;
    float* d_rgba = rgba;
    uint8_t* d_u8 = rgba_u8;
    if (!dev_out) {
        if (rgba) { W.o_rgba.reserve((size_t)P * 16); d_rgba = W.o_rgba.as<float>(); }
        if (rgba_u8) { W.o_u8.reserve((size_t)P * 4); d_u8 = W.o_u8.as<uint8_t>(); }
    }
    hipLaunchKernelGGL(k_finish, dim3(std::max(1u, cdiv(P, TB))), dim3(TB), 0, st, C.mean, P, d_rgba, d_u8);
    CRT_HIP_CHECK(hipGetLastError());
    CRT_HIP_CHECK(hipEventRecord(e1, st));
    if (!dev_out) {
        if (rgba) CRT_HIP_CHECK(hipMemcpyAsync(rgba, d_rgba, (size_t)P * 16, hipMemcpyDeviceToHost, st));
        if (rgba_u8) CRT_HIP_CHECK(hipMemcpyAsync(rgba_u8, d_u8, (size_t)P * 4, hipMemcpyDeviceToHost, st));
        if (spp_used) CRT_HIP_CHECK(hipMemcpyAsync(spp_used, spp.p, (size_t)P * 4, hipMemcpyDeviceToHost, st));
    } else if (spp_used) {
        CRT_HIP_CHECK(hipMemcpyAsync(spp_used, spp.p, (size_t)P * 4, hipMemcpyDeviceToDevice, st));
    }
    unsigned long long ht[2] = {0, 0};
    uint32_t overflow = 0;
    CRT_HIP_CHECK(hipMemcpyAsync(ht, tally.p, 16, hipMemcpyDeviceToHost, st));
    CRT_HIP_CHECK(hipMemcpyAsync(&overflow, C.cnt + CNT_OVERFLOW, 4, hipMemcpyDeviceToHost, st));
    CRT_HIP_CHECK(hipStreamSynchronize(st));
    CRT_REQUIRE(overflow == 0, CRT_ERR_STATE, "render queue shard overflow");
    float ms = 0.f;
    CRT_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (timing) {
        std::memset(timing, 0, sizeof(*timing));
        timing->total_ms = ms;
        timing->rays_traced = ht[0];
        timing->shadow_rays = ht[1];
        (void)passes;
    }
    return CRT_OK;
}

}  // namespace crt

using namespace crt;

extern "C" int crt_render(crt_scene* scene, const crt_camera* camera, const crt_light* lights, uint32_t n_lights,
                          const crt_render_opts* opts, float* rgba, uint8_t* rgba_u8, int32_t* spp_used,
                          uint32_t flags, void* stream, crt_timing* timing) {
    return guarded([&] {
        CRT_REQUIRE(opts != nullptr, CRT_ERR_INVALID, "render options are NULL");
        CRT_REQUIRE(scene != nullptr, CRT_ERR_INVALID, "scene is NULL");
        const uint64_t seed = opts->seed;
        crt_scene::StreamSlot& slot = scene->slot(stream);
        std::lock_guard<std::mutex> slot_lock(slot.mu);
        if (!slot.rws) slot.rws = new_render_workspace();
        render_batch(scene, 1, camera, nullptr, &seed, lights, n_lights, false, opts, nullptr, rgba, rgba_u8, spp_used,
                     flags, stream, timing, slot.rws);
    });
}

extern "C" int crt_render_batch(crt_scene* scene, uint32_t n_cams, const crt_camera* cams,
                                const int32_t* max_samples, const uint64_t* seeds, const crt_light* lights,
                                uint32_t n_lights, const crt_render_opts* opts, const crt_batch_opts* bopts,
                                float* rgba, uint8_t* rgba_u8, int32_t* spp_used, uint32_t flags, void* stream,
                                crt_timing* timing) {
    return guarded([&] {
        CRT_REQUIRE(scene != nullptr, CRT_ERR_INVALID, "scene is NULL");
        crt_scene::StreamSlot& slot = scene->slot(stream);
        std::lock_guard<std::mutex> slot_lock(slot.mu);
        if (!slot.rws) slot.rws = new_render_workspace();
        render_batch(scene, n_cams, cams, max_samples, seeds, lights, n_lights, false, opts, bopts, rgba, rgba_u8,
                     spp_used, flags, stream, timing, slot.rws);
    });
}
