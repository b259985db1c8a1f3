// bvh_build.hip -- GPU construction of the 8-wide BVH.
//
// Replaces the reference's CPU build pipeline (SweepSahBuilder ->
// ParallelReinsertionOptimizer -> NodeLayoutOptimizer,
// src/crt/rendering_body_fixed/body_fixed_group.hpp:66-80) with:
//   1. k_prep     : per-triangle AABB + centroid, block-reduced centroid and
//                   scene bounds (utilities.hpp:159-189 counterpart)
//   2. k_morton   : 63-bit Morton code of the centroid (21 bits per axis)
//   3. radix sort : 8 passes of an 8-bit LSD sort -- per-tile LDS digit
//                   histograms, a per-digit row scan plus a digit scan, and a
//                   stable scatter whose ranks come from 64-lane ballot matching
//   4. binary tree, one of:
//      LBVH (default): k_karras binary radix tree (Karras 2012) on the
//                   Morton order; node boxes as range unions over a segment
//                   tree of the leaf boxes (k_seg_level + k_range_boxes:
//                   O(log n) blocks per node, exact min/max)
//      PLOC (CRT_BUILD_PLOC): k_ploc_nn / k_ploc_merge / k_ploc_compact rounds of
//                   locally-ordered clustering on the Morton order (merged
//                   surface area as the distance), then leaf counts bottom-up
//                   and depth-first leaf positions top-down (one launch per
//                   round), so every subtree is a contiguous triangle range.
//                   Lower SAH; on the L8 body it traces shadow rays ~4 %
//                   faster and camera rays ~4 % slower, at ~3x the build time
//   5. k_collapse : top-down, level-synchronous collapse into 8-wide nodes;
//                   each wide node greedily opens its largest-area inner
//                   children until 8 slots are used; subtrees with at most
//                   `max_leaf` triangles become leaves (contiguous ranges)
//   6. k_compress : 96-B compressed copy of every wide node
//   7. k_emit     : triangles written in leaf order (48 B each)
#include "crt_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace crt {
namespace {

constexpr int BLK = 256;

__device__ inline uint32_t f2o(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float o2f(uint32_t u) {
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

// bounds[0..5]  : centroid min xyz, max xyz (ordered uint)
// bounds[6..11] : triangle AABB min xyz, max xyz (ordered uint)
__global__ __launch_bounds__(BLK) void k_prep(const float* __restrict__ V, const uint32_t* __restrict__ I,
                                              uint32_t n, float4* __restrict__ lo, float4* __restrict__ hi,
                                              float4* __restrict__ cen, uint32_t* __restrict__ bounds) {
    __shared__ float red[12][BLK];
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    float v[12];
    if (i < n) {
        const uint32_t a = I[3 * i], b = I[3 * i + 1], c = I[3 * i + 2];
        const float ax = V[3 * a], ay = V[3 * a + 1], az = V[3 * a + 2];
        const float bx = V[3 * b], by = V[3 * b + 1], bz = V[3 * b + 2];
        const float cx = V[3 * c], cy = V[3 * c + 1], cz = V[3 * c + 2];
        const float lx = fminf(ax, fminf(bx, cx)), ly = fminf(ay, fminf(by, cy)), lz = fminf(az, fminf(bz, cz));
        const float hx = fmaxf(ax, fmaxf(bx, cx)), hy = fmaxf(ay, fmaxf(by, cy)), hz = fmaxf(az, fmaxf(bz, cz));
        const float mx = (ax + bx + cx) * (1.0f / 3.0f), my = (ay + by + cy) * (1.0f / 3.0f),
                    mz = (az + bz + cz) * (1.0f / 3.0f);
        lo[i] = make_float4(lx, ly, lz, 0.f);
        hi[i] = make_float4(hx, hy, hz, 0.f);
        cen[i] = make_float4(mx, my, mz, 0.f);
        v[0] = mx; v[1] = my; v[2] = mz; v[3] = mx; v[4] = my; v[5] = mz;
        v[6] = lx; v[7] = ly; v[8] = lz; v[9] = hx; v[10] = hy; v[11] = hz;
    } else {
        const float inf = __int_as_float(0x7f800000);
        for (int k = 0; k < 3; ++k) { v[k] = inf; v[3 + k] = -inf; v[6 + k] = inf; v[9 + k] = -inf; }
    }
    for (int k = 0; k < 12; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = BLK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int k = 0; k < 12; ++k) {
                const bool mn = (k < 3) || (k >= 6 && k < 9);
                const float x = red[k][threadIdx.x], y = red[k][threadIdx.x + s];
                red[k][threadIdx.x] = mn ? fminf(x, y) : fmaxf(x, y);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 12) {
        const int k = threadIdx.x;
        const bool mn = (k < 3) || (k >= 6 && k < 9);
        const uint32_t o = f2o(red[k][0]);
        if (mn) atomicMin(&bounds[k], o); else atomicMax(&bounds[k], o);
    }
}

__device__ inline uint64_t split3_21(uint32_t a) {
    uint64_t x = a & 0x1FFFFFull;
    x = (x | x << 32) & 0x1F00000000FFFFull;
    x = (x | x << 16) & 0x1F0000FF0000FFull;
    x = (x | x << 8) & 0x100F00F00F00F00Full;
    x = (x | x << 4) & 0x10C30C30C30C30C3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ __launch_bounds__(BLK) void k_morton(const float4* __restrict__ cen, uint32_t n,
                                                const uint32_t* __restrict__ bounds,
                                                uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const float4 c = cen[i];
    uint32_t q[3];
    const float cc[3] = {c.x, c.y, c.z};
    for (int k = 0; k < 3; ++k) {
        const float lo = o2f(bounds[k]), hi = o2f(bounds[3 + k]);
        const float ext = hi - lo;
        float f = ext > 0.f ? (cc[k] - lo) / ext : 0.f;
        f = fminf(fmaxf(f, 0.f), 1.f);
        uint32_t qi = (uint32_t)(f * 2097152.0f);
        q[k] = qi > 2097151u ? 2097151u : qi;
    }
    keys[i] = (split3_21(q[0]) << 2) | (split3_21(q[1]) << 1) | split3_21(q[2]);
    vals[i] = i;
}

// ---------------- LSD radix sort (8-bit digits) ----------------
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = BLK * RS_ITEMS;

__global__ __launch_bounds__(BLK) void k_rs_hist(const uint64_t* __restrict__ keys, uint32_t n, int shift,
                                                 uint32_t nblocks, uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * RS_TILE;
    for (int it = 0; it < RS_ITEMS; ++it) {
        const uint32_t i = base + it * BLK + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Inclusive scan across a wave (64 lanes).
__device__ inline uint32_t wave_scan_incl(uint32_t v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    return v;
}

// Exclusive scan of each digit's row of per-block counts (one block per
// digit, hist is digit-major); the row totals go to `totals`.
__global__ __launch_bounds__(BLK) void k_rs_scan_rows(uint32_t* __restrict__ hist, uint32_t nb,
                                                      uint32_t* __restrict__ totals) {
    __shared__ uint32_t wsum[BLK / 64];
    __shared__ uint32_t carry;
    uint32_t* row = hist + (size_t)blockIdx.x * nb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0u;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += BLK) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t x = i < nb ? row[i] : 0u;
        const uint32_t v = wave_scan_incl(x, lane);
        if (lane == 63) wsum[wave] = v;
        __syncthreads();
        uint32_t pre = carry;
        for (int w = 0; w < wave; ++w) pre += wsum[w];
        if (i < nb) row[i] = pre + v - x;
        __syncthreads();
        if (threadIdx.x == BLK - 1) carry = pre + v;
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Exclusive scan of the 256 digit totals in place (one block of 256).
__global__ __launch_bounds__(BLK) void k_rs_scan_digits(uint32_t* __restrict__ totals) {
    __shared__ uint32_t wsum[BLK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t x = totals[threadIdx.x];
    const uint32_t v = wave_scan_incl(x, lane);
    if (lane == 63) wsum[wave] = v;
    __syncthreads();
    uint32_t pre = 0u;
    for (int w = 0; w < wave; ++w) pre += wsum[w];
    totals[threadIdx.x] = pre + v - x;
}

__global__ __launch_bounds__(BLK) void k_rs_scatter(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                    uint32_t n, int shift, uint32_t nblocks,
                                                    const uint32_t* __restrict__ hist,
                                                    const uint32_t* __restrict__ digit_base) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[BLK / 64][256];
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    run[tid] = hist[tid * nblocks + blockIdx.x] + digit_base[tid];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t base = blockIdx.x * RS_TILE;
    for (int it = 0; it < RS_ITEMS; ++it) {
        for (int w = 0; w < BLK / 64; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        const uint32_t i = base + it * BLK + tid;
        const bool valid = i < n;
        const uint64_t key = valid ? kin[i] : 0ull;
        const uint32_t val = valid ? vin[i] : 0u;
        const uint32_t dig = (uint32_t)(key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (dig >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[wave][dig] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = run[dig] + rank;
            for (int w = 0; w < wave; ++w) off += wcnt[w][dig];
            kout[off] = key;
            vout[off] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < BLK / 64; ++w) add += wcnt[w][tid];
        run[tid] += add;
        __syncthreads();
    }
}

// ---------------- Karras radix tree ----------------
__device__ inline int delta(const uint64_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(a ^ b);
}

// Unified node ids: internal 0..n-2, leaf k -> n-1+k.
__global__ __launch_bounds__(BLK) void k_karras(const uint64_t* __restrict__ keys, int n,
                                                int* __restrict__ left, int* __restrict__ right,
                                                int* __restrict__ parent, int* __restrict__ rfirst,
                                                int* __restrict__ rlast) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int g = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    const int lc = (lo == g) ? (n - 1 + g) : g;
    const int rc = (hi == g + 1) ? (n - 1 + g + 1) : g + 1;
    left[i] = lc;
    right[i] = rc;
    parent[lc] = i;
    parent[rc] = i;
    rfirst[i] = lo;
    rlast[i] = hi;
}

// Segment tree over the sorted leaf boxes.  Level 0 is the leaves
// themselves (blo/bhi[n-1 .. 2n-2]); entry i of level L >= 1 is the union of
// entries 2i and 2i+1 of level L-1, i.e. of leaves [i 2^L, (i+1) 2^L).
struct SegLevels {
    uint32_t off[40];    // first entry of level L in the seg arrays (L >= 1)
    uint32_t count;      // number of levels including level 0
};

__global__ __launch_bounds__(BLK) void k_seg_level(const float4* __restrict__ lo_in, const float4* __restrict__ hi_in,
                                                   uint32_t n_in, float4* __restrict__ lo_out,
                                                   float4* __restrict__ hi_out) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (2u * i >= n_in) return;
    float4 lo = lo_in[2 * i], hi = hi_in[2 * i];
    if (2u * i + 1u < n_in) {
        const float4 l2 = lo_in[2 * i + 1], h2 = hi_in[2 * i + 1];
        lo.x = fminf(lo.x, l2.x); lo.y = fminf(lo.y, l2.y); lo.z = fminf(lo.z, l2.z);
        hi.x = fmaxf(hi.x, h2.x); hi.y = fmaxf(hi.y, h2.y); hi.z = fmaxf(hi.z, h2.z);
    }
    lo_out[i] = lo;
    hi_out[i] = hi;
}

// Box of every internal node = union of its leaf range [rfirst, rlast],
// gathered from O(log n) aligned segment-tree blocks.
__global__ __launch_bounds__(BLK) void k_range_boxes(int n, const int* __restrict__ rfirst,
                                                     const int* __restrict__ rlast, const float4* __restrict__ seg_lo,
                                                     const float4* __restrict__ seg_hi, SegLevels L, float4* blo,
                                                     float4* bhi) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n - 1) return;
    float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
    auto take = [&](uint32_t level, uint32_t e) {
        const float4 a = level == 0 ? blo[n - 1 + e] : seg_lo[L.off[level] + e];
        const float4 b = level == 0 ? bhi[n - 1 + e] : seg_hi[L.off[level] + e];
        lo.x = fminf(lo.x, a.x); lo.y = fminf(lo.y, a.y); lo.z = fminf(lo.z, a.z);
        hi.x = fmaxf(hi.x, b.x); hi.y = fmaxf(hi.y, b.y); hi.z = fmaxf(hi.z, b.z);
    };
    uint32_t l = (uint32_t)rfirst[i], r = (uint32_t)rlast[i] + 1u, level = 0;
    while (l < r) {
        if (l & 1u) take(level, l++);
        if (r & 1u) take(level, --r);
        l >>= 1;
        r >>= 1;
        ++level;
    }
    blo[i] = lo;
    bhi[i] = hi;
}

// Leaf boxes in sorted order: blo[offset+k] = box of triangle vals[k].
__global__ __launch_bounds__(BLK) void k_leaf_boxes(int n, int offset, const uint32_t* __restrict__ vals,
                                                    const float4* __restrict__ tlo, const float4* __restrict__ thi,
                                                    float4* __restrict__ blo, float4* __restrict__ bhi) {
    const int k = blockIdx.x * BLK + threadIdx.x;
    if (k >= n) return;
    const uint32_t t = vals[k];
    blo[offset + k] = tlo[t];
    bhi[offset + k] = thi[t];
}

// ---- PLOC: parallel locally-ordered clustering (Meister and Bittner,
// IEEE TVCG 2018) over the Morton-sorted triangles.  Every round, each
// cluster finds its nearest neighbour (smallest merged surface area) among
// the PLOC_R clusters on either side; mutual nearest neighbours merge (the
// lower index keeps the new cluster); survivors are compacted in order.
// Cluster refs: PLOC_LEAF | sorted leaf index, or an internal node id
// (allocated in merge order, so children always precede their parent).
constexpr int PLOC_R = 16;
constexpr uint32_t PLOC_LEAF = 0x80000000u;

__device__ inline float union_area(float4 a0, float4 a1, float4 b0, float4 b1) {
    const float dx = fmaxf(a1.x, b1.x) - fminf(a0.x, b0.x);
    const float dy = fmaxf(a1.y, b1.y) - fminf(a0.y, b0.y);
    const float dz = fmaxf(a1.z, b1.z) - fminf(a0.z, b0.z);
    return dx * dy + dy * dz + dz * dx;
}

__global__ __launch_bounds__(BLK) void k_ploc_nn(uint32_t c, const float4* __restrict__ lo,
                                                 const float4* __restrict__ hi, uint32_t* __restrict__ nn) {
    __shared__ float4 slo[BLK + 2 * PLOC_R], shi[BLK + 2 * PLOC_R];
    const int base = (int)(blockIdx.x * BLK) - PLOC_R;
    for (int t = threadIdx.x; t < BLK + 2 * PLOC_R; t += BLK) {
        const int g = base + t;
        if (g >= 0 && g < (int)c) { slo[t] = lo[g]; shi[t] = hi[g]; }
    }
    __syncthreads();
    const int i = (int)(blockIdx.x * BLK + threadIdx.x);
    if (i >= (int)c) return;
    const int li = (int)threadIdx.x + PLOC_R;
    const float4 a0 = slo[li], a1 = shi[li];
    float best = INFINITY;
    int bj = i;
    for (int d = -PLOC_R; d <= PLOC_R; ++d) {   // ascending j: ties keep the lower index
        const int j = i + d;
        if (d == 0 || j < 0 || j >= (int)c) continue;
        const float a = union_area(a0, a1, slo[li + d], shi[li + d]);
        if (a < best) { best = a; bj = j; }
    }
    nn[i] = (uint32_t)bj;
}

__global__ __launch_bounds__(BLK) void k_ploc_merge(uint32_t c, const uint32_t* __restrict__ nn, uint32_t* ref,
                                                    float4* lo, float4* hi, uint32_t* __restrict__ node_counter,
                                                    uint32_t* __restrict__ nleft, uint32_t* __restrict__ nright,
                                                    float4* __restrict__ nlo, float4* __restrict__ nhi,
                                                    uint8_t* __restrict__ keep, uint32_t* __restrict__ bcount) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    bool k = false;
    if (i < c) {
        const uint32_t j = nn[i];
        const bool mutual = j != i && nn[j] == i;
        if (mutual && i < j) {
            // only i writes slot i and only i reads slot j: no other thread touches either
            const float4 a0 = lo[i], a1 = hi[i], b0 = lo[j], b1 = hi[j];
            const float4 u0 = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.f);
            const float4 u1 = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.f);
            const uint32_t m = atomicAdd(node_counter, 1u);
            nleft[m] = ref[i];
            nright[m] = ref[j];
            nlo[m] = u0;
            nhi[m] = u1;
            ref[i] = m;
            lo[i] = u0;
            hi[i] = u1;
            k = true;
        } else {
            k = !mutual;   // a mutual pair's higher index is absorbed
        }
        keep[i] = k ? 1 : 0;
    }
    const int cnt = __syncthreads_count(k);
    if (threadIdx.x == 0) bcount[blockIdx.x] = (uint32_t)cnt;
}

// Exclusive scan of nb block counts in place (one block); total at [nb].
__global__ __launch_bounds__(BLK) void k_ploc_scan(uint32_t* __restrict__ bcount, uint32_t nb) {
    __shared__ uint32_t wsum[BLK / 64];
    __shared__ uint32_t carry;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0u;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += BLK) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t x = i < nb ? bcount[i] : 0u;
        const uint32_t v = wave_scan_incl(x, lane);
        if (lane == 63) wsum[wave] = v;
        __syncthreads();
        uint32_t pre = carry;
        for (int w = 0; w < wave; ++w) pre += wsum[w];
        if (i < nb) bcount[i] = pre + v - x;
        __syncthreads();
        if (threadIdx.x == BLK - 1) carry = pre + v;
        __syncthreads();
    }
    if (threadIdx.x == 0) bcount[nb] = carry;
}

// Ordered compaction of the surviving clusters.
__global__ __launch_bounds__(BLK) void k_ploc_compact(uint32_t c, const uint8_t* __restrict__ keep,
                                                      const uint32_t* __restrict__ boff,
                                                      const uint32_t* __restrict__ ref, const float4* __restrict__ lo,
                                                      const float4* __restrict__ hi, uint32_t* __restrict__ ref2,
                                                      float4* __restrict__ lo2, float4* __restrict__ hi2) {
    __shared__ uint32_t wcnt[BLK / 64];
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool k = i < c && keep[i];
    const uint64_t m = __ballot(k);
    if (lane == 0) wcnt[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (!k) return;
    uint32_t pos = boff[blockIdx.x] + (uint32_t)__popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    for (int w = 0; w < wave; ++w) pos += wcnt[w];
    ref2[pos] = ref[i];
    lo2[pos] = lo[i];
    hi2[pos] = hi[i];
}

__global__ __launch_bounds__(BLK) void k_ploc_init(uint32_t n, const float4* __restrict__ llo,
                                                   const float4* __restrict__ lhi, uint32_t* __restrict__ ref,
                                                   float4* __restrict__ lo, float4* __restrict__ hi) {
    const uint32_t k = blockIdx.x * BLK + threadIdx.x;
    if (k >= n) return;
    ref[k] = PLOC_LEAF | k;
    lo[k] = llo[k];
    hi[k] = lhi[k];
}

// Leaf counts of the nodes made in one round (children come from earlier rounds).
__global__ __launch_bounds__(BLK) void k_ploc_sizes(uint32_t begin, uint32_t end, const uint32_t* __restrict__ nleft,
                                                    const uint32_t* __restrict__ nright, uint32_t* size) {
    const uint32_t m = begin + blockIdx.x * BLK + threadIdx.x;
    if (m >= end) return;
    const uint32_t l = nleft[m], r = nright[m];
    size[m] = ((l & PLOC_LEAF) ? 1u : size[l]) + ((r & PLOC_LEAF) ? 1u : size[r]);
}

// Depth-first leaf positions, top-down one round at a time: every subtree
// gets a contiguous range [off, off + size).
__global__ __launch_bounds__(BLK) void k_ploc_offsets(uint32_t begin, uint32_t end,
                                                      const uint32_t* __restrict__ nleft,
                                                      const uint32_t* __restrict__ nright,
                                                      const uint32_t* __restrict__ size, uint32_t* off,
                                                      uint32_t* __restrict__ leaf_pos) {
    const uint32_t m = begin + blockIdx.x * BLK + threadIdx.x;
    if (m >= end) return;
    const uint32_t l = nleft[m], r = nright[m], o = off[m];
    const uint32_t sl = (l & PLOC_LEAF) ? 1u : size[l];
    if (l & PLOC_LEAF) leaf_pos[l & ~PLOC_LEAF] = o; else off[l] = o;
    if (r & PLOC_LEAF) leaf_pos[r & ~PLOC_LEAF] = o + sl; else off[r] = o + sl;
}

// PLOC tree -> the unified layout the collapse reads: internal node id
// n-2-m (root 0), leaves n-1+position, leaf ranges, boxes, and the triangle
// order for emission.
__global__ __launch_bounds__(BLK) void k_ploc_emit(uint32_t n, const uint32_t* __restrict__ nleft,
                                                   const uint32_t* __restrict__ nright,
                                                   const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                                   const uint32_t* __restrict__ size, const uint32_t* __restrict__ off,
                                                   const uint32_t* __restrict__ leaf_pos,
                                                   const uint32_t* __restrict__ sorted_vals,
                                                   const float4* __restrict__ llo, const float4* __restrict__ lhi,
                                                   int* __restrict__ left, int* __restrict__ right,
                                                   int* __restrict__ rfirst, int* __restrict__ rlast,
                                                   float4* __restrict__ blo, float4* __restrict__ bhi,
                                                   uint32_t* __restrict__ vals) {
    const uint32_t t = blockIdx.x * BLK + threadIdx.x;
    auto map = [&](uint32_t r) -> int {
        return (r & PLOC_LEAF) ? (int)(n - 1 + leaf_pos[r & ~PLOC_LEAF]) : (int)(n - 2 - r);
    };
    if (t < n - 1) {
        const int id = (int)(n - 2 - t);
        left[id] = map(nleft[t]);
        right[id] = map(nright[t]);
        rfirst[id] = (int)off[t];
        rlast[id] = (int)(off[t] + size[t] - 1u);
        blo[id] = nlo[t];
        bhi[id] = nhi[t];
    }
    if (t < n) {
        const uint32_t p = leaf_pos[t];
        blo[n - 1 + p] = llo[t];
        bhi[n - 1 + p] = lhi[t];
        vals[p] = sorted_vals[t];
    }
}

__device__ inline float half_area(float4 lo, float4 hi) {
    const float dx = fmaxf(hi.x - lo.x, 0.f), dy = fmaxf(hi.y - lo.y, 0.f), dz = fmaxf(hi.z - lo.z, 0.f);
    return dx * dy + dy * dz + dz * dx;
}

__device__ inline float pad_down(float x) { return x - fmaxf(fabsf(x) * 4.8e-7f, 1e-30f); }
__device__ inline float pad_up(float x) { return x + fmaxf(fabsf(x) * 4.8e-7f, 1e-30f); }

struct Item { int bnode; int wnode; };

__global__ __launch_bounds__(BLK) void k_collapse(int n, const Item* __restrict__ in, uint32_t n_in,
                                                  Item* __restrict__ out, uint32_t* __restrict__ counters,
                                                  const int* __restrict__ left, const int* __restrict__ right,
                                                  const int* __restrict__ rfirst, const int* __restrict__ rlast,
                                                  const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                                  uint32_t max_leaf, float4* __restrict__ nodes) {
    // counters[0] = wide node count, counters[1] = next frontier size, counters[2] = leaves
    const uint32_t w = blockIdx.x * BLK + threadIdx.x;
    if (w >= n_in) return;
    const Item it = in[w];
    auto count_of = [&](int b) -> uint32_t {
        return b >= n - 1 ? 1u : (uint32_t)(rlast[b] - rfirst[b] + 1);
    };
    int slot[8];
    int cnt = 0;
    if (it.bnode >= n - 1) {
        slot[cnt++] = it.bnode;           // single-leaf root
    } else {
        slot[cnt++] = left[it.bnode];
        slot[cnt++] = right[it.bnode];
    }
    while (cnt < 8) {
        int best = -1;
        float barea = -1.f;
        for (int k = 0; k < cnt; ++k) {
            const int b = slot[k];
            if (b >= n - 1) continue;                 // binary leaf
            if (count_of(b) <= max_leaf) continue;    // will be a leaf
            const float a = half_area(blo[b], bhi[b]);
            if (a > barea) { barea = a; best = k; }
        }
        if (best < 0) break;
        const int b = slot[best];
        slot[best] = left[b];
        slot[cnt++] = right[b];
    }
    float4* nd = nodes + (size_t)it.wnode * 16u;
    uint32_t nleaf = 0;
    for (int k = 0; k < 8; ++k) {
        if (k < cnt) {
            const int b = slot[k];
            const float4 lo = blo[b], hi = bhi[b];
            uint32_t ref;
            const uint32_t c = count_of(b);
            if (b >= n - 1 || c <= max_leaf) {
                const uint32_t first = b >= n - 1 ? (uint32_t)(b - (n - 1)) : (uint32_t)rfirst[b];
                ref = make_leaf_ref(first, c);
                ++nleaf;
            } else {
                const uint32_t wn = atomicAdd(&counters[0], 1u);
                const uint32_t q = atomicAdd(&counters[1], 1u);
                out[q] = Item{b, (int)wn};
                ref = wn;
            }
            nd[2 * k] = make_float4(pad_down(lo.x), pad_down(lo.y), pad_down(lo.z), __uint_as_float(ref));
            nd[2 * k + 1] = make_float4(pad_up(hi.x), pad_up(hi.y), pad_up(hi.z), __uint_as_float(1u));
        } else {
            const float inf = __int_as_float(0x7f800000);
            nd[2 * k] = make_float4(inf, inf, inf, __uint_as_float(REF_EMPTY));
            nd[2 * k + 1] = make_float4(-inf, -inf, -inf, __uint_as_float(0u));
        }
    }
    atomicAdd(&counters[2], nleaf);
}

// ---- SAH-optimal collapse (default; CRT_BUILD_GREEDY_WIDE selects k_collapse) ----
// Dynamic programming over the binary tree (Ylitie, Karras and Laine, HPG
// 2017): C(m, i) = cheapest SAH cost of representing subtree m by at most i
// subtrees in its parent's wide node, with
//   leaf(m)     = A(m) * (C_LEAF + C_TRI * N(m))   (N(m) <= max_leaf)
//   dist(m, j)  = min_k C(a, k) + C(b, j - k)      (a, b children of m)
//   wide(m)     = A(m) * C_NODE + dist(m, 8)
//   C(m, 1)     = min(leaf(m), wide(m)),   C(m, j) = min(C(m, 1), dist(m, j))
// computed level by level from the deepest binary level up; the collapse
// then expands every wide node's 8 slots along the recorded choices.
// C_NODE: a wide-node visit; C_LEAF: the fixed cost of entering a leaf;
// C_TRI: one triangle test.  Chosen by measurement on the 1.31M-triangle
// body (camera rays, shadow rays, flash LiDAR, flux, path tracing): with a
// per-triangle cost of 0.3 the DP makes leaves of 1-2 triangles and camera
// rays get 8-11 % slower than with the greedy collapse; C_TRI 0.05-0.1 with
// C_LEAF 2-4 keeps camera rays level and makes shadow rays ~6 % and path
// tracing ~2 % faster.  Since the lane traversal tests leaves once half of
// its wave holds one, a leaf entry costs less: C_LEAF 0 / 1 / 2 / 3 / 4 with
// C_TRI 0.1 gave the 2048^2 sun pass 8826-8880 / 8770-8778 / 8749-8783 /
// 8681-8719 / 8643-8737 Mrays/s and the 256-camera batch 156.3-156.7 /
// 156.1-156.3 / 155.9-156.7 / 156.7-157.0 / 156.8-157.5 ms (C_TRI 0.15 and
// 0.2 at C_LEAF 0: sun pass 8929-8966 and 8856-8928, batch 157.6-157.7 and
// 158.4-158.6 ms).  C_LEAF 0 merges the 36 large triangles of the Cornell
// box into a few big leaves (its render 0.74 -> 0.84 ms); C_LEAF 1 keeps the
// Cornell render level (0.740-0.761 ms), the sun pass at 8790-8827, the batch
// at 156.1-156.2 ms and renders a 3.2M-triangle geodesic sphere 4 % faster.
constexpr float DP_C_NODE = 1.0f;
constexpr float DP_C_LEAF = 1.0f;
constexpr float DP_C_TRI = 0.1f;

// Internal binary nodes of one level -> their internal children (next level).
// Writes beyond `cap` entries are dropped (the host then rejects the count).
__global__ __launch_bounds__(BLK) void k_bin_level(int n, const int* __restrict__ cur, uint32_t n_cur,
                                                   int* __restrict__ next, uint32_t* __restrict__ n_next,
                                                   uint32_t cap, const int* __restrict__ left,
                                                   const int* __restrict__ right) {
    const uint32_t t = blockIdx.x * BLK + threadIdx.x;
    int kids[2] = {-1, -1};
    if (t < n_cur) {
        const int m = cur[t];
        if (left[m] < n - 1) kids[0] = left[m];
        if (right[m] < n - 1) kids[1] = right[m];
    }
    for (int c = 0; c < 2; ++c) {
        const bool want = kids[c] >= 0;
        const uint64_t mask = __ballot(want);
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((long long)mask) - 1;
        uint32_t base = 0;
        if (mask != 0ull && lane == leader) base = atomicAdd(n_next, (uint32_t)__popcll(mask));
        base = (uint32_t)__shfl((int)base, leader < 0 ? 0 : leader, 64);
        const uint32_t pos = base + (uint32_t)__popcll(mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
        if (want && pos < cap) next[pos] = kids[c];
    }
}

// DP costs of the internal nodes of one level (children already done).
// cost[m*8 + i-1] = C(m, i); split[m*8 + j] = argmin k of dist(m, j);
// flags[m]: bit 0 = leaf(m) <= wide(m), bit j = C(m, j) takes dist(m, j).
__global__ __launch_bounds__(BLK) void k_dp(int n, const int* __restrict__ nodes, uint32_t count,
                                            const int* __restrict__ left, const int* __restrict__ right,
                                            const int* __restrict__ rfirst, const int* __restrict__ rlast,
                                            const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                            uint32_t max_leaf, float* __restrict__ cost, uint8_t* __restrict__ split,
                                            uint16_t* __restrict__ flags) {
    const uint32_t t = blockIdx.x * BLK + threadIdx.x;
    if (t >= count) return;
    const int m = nodes[t];
    float ca[8], cb[8];
    const int ab[2] = {left[m], right[m]};
    for (int c = 0; c < 2; ++c) {
        float* cc = c == 0 ? ca : cb;
        const int ch = ab[c];
        if (ch >= n - 1) {
            const float a = half_area(blo[ch], bhi[ch]) * (DP_C_LEAF + DP_C_TRI);
            for (int i = 0; i < 8; ++i) cc[i] = a;
        } else {
            for (int i = 0; i < 8; ++i) cc[i] = cost[(size_t)ch * 8 + i];
        }
    }
    const float A = half_area(blo[m], bhi[m]);
    const uint32_t N = (uint32_t)(rlast[m] - rfirst[m] + 1);
    float dist[9];
    uint8_t sp[9];
    for (int j = 2; j <= 8; ++j) {
        float best = INFINITY;
        int bk = 1;
        for (int k = 1; k < j; ++k) {
            const float v = ca[k - 1] + cb[j - k - 1];
            if (v < best) { best = v; bk = k; }
        }
        dist[j] = best;
        sp[j] = (uint8_t)bk;
    }
    const float leaf = N <= max_leaf ? A * (DP_C_LEAF + DP_C_TRI * (float)N) : INFINITY;
    const float wide = A * DP_C_NODE + dist[8];
    const float c1 = fminf(leaf, wide);
    uint16_t fl = leaf <= wide ? 1u : 0u;
    cost[(size_t)m * 8] = c1;
    split[(size_t)m * 8] = 0;
    split[(size_t)m * 8 + 1] = 0;
    for (int j = 2; j <= 8; ++j) {
        const bool use = dist[j] < c1;
        if (use) fl |= (uint16_t)(1u << j);
        if (j < 8) cost[(size_t)m * 8 + j - 1] = use ? dist[j] : c1;
        else cost[(size_t)m * 8 + 7] = use ? dist[j] : c1;
        split[(size_t)m * 8 + (j & 7)] = sp[j];   // j = 8 stored in slot 0
    }
    flags[m] = fl;
}

// Collapse along the DP choices: the wide node at internal binary node
// it.bnode gets the slots of expand(left, s) + expand(right, 8 - s), where
// expand(c, i) is c itself if c is a binary leaf, i == 1 or C(c, i) does not
// take dist(c, i), and expand(left(c), s') + expand(right(c), i - s')
// otherwise.  A slot becomes a leaf when it is a binary leaf or leaf(c) won,
// else a new wide node.  Output format as k_collapse.
__global__ __launch_bounds__(BLK) void k_collapse_dp(int n, const Item* __restrict__ in, uint32_t n_in,
                                                     Item* __restrict__ out, uint32_t* __restrict__ counters,
                                                     const int* __restrict__ left, const int* __restrict__ right,
                                                     const int* __restrict__ rfirst, const int* __restrict__ rlast,
                                                     const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                                     const uint8_t* __restrict__ split,
                                                     const uint16_t* __restrict__ flags,
                                                     float4* __restrict__ nodes) {
    const uint32_t w = blockIdx.x * BLK + threadIdx.x;
    if (w >= n_in) return;
    const Item it = in[w];
    int slot[8];
    int cnt = 0;
    if (it.bnode >= n - 1) {
        slot[cnt++] = it.bnode;           // single-leaf root
    } else {
        // explicit stack of (binary node, slot budget); budgets sum to <= 8
        int stn[8], stb[8];
        int sp = 0;
        const int s8 = (int)split[(size_t)it.bnode * 8];   // split of dist(., 8)
        stn[sp] = right[it.bnode]; stb[sp++] = 8 - s8;
        stn[sp] = left[it.bnode]; stb[sp++] = s8;
        while (sp > 0 && cnt < 8) {
            --sp;
            const int c = stn[sp], i = stb[sp];
            if (c >= n - 1 || i <= 1 || !((flags[c] >> i) & 1u) || sp + 2 > 8) {
                slot[cnt++] = c;
            } else {
                const int k = (int)split[(size_t)c * 8 + (i & 7)];
                stn[sp] = right[c]; stb[sp++] = i - k;
                stn[sp] = left[c]; stb[sp++] = k;
            }
        }
    }
    float4* nd = nodes + (size_t)it.wnode * 16u;
    uint32_t nleaf = 0;
    for (int k = 0; k < 8; ++k) {
        if (k < cnt) {
            const int b = slot[k];
            const float4 lo = blo[b], hi = bhi[b];
            uint32_t ref;
            if (b >= n - 1 || (flags[b] & 1u)) {
                const uint32_t first = b >= n - 1 ? (uint32_t)(b - (n - 1)) : (uint32_t)rfirst[b];
                const uint32_t c = b >= n - 1 ? 1u : (uint32_t)(rlast[b] - rfirst[b] + 1);
                ref = make_leaf_ref(first, c);
                ++nleaf;
            } else {
                const uint32_t wn = atomicAdd(&counters[0], 1u);
                const uint32_t q = atomicAdd(&counters[1], 1u);
                out[q] = Item{b, (int)wn};
                ref = wn;
            }
            nd[2 * k] = make_float4(pad_down(lo.x), pad_down(lo.y), pad_down(lo.z), __uint_as_float(ref));
            nd[2 * k + 1] = make_float4(pad_up(hi.x), pad_up(hi.y), pad_up(hi.z), __uint_as_float(1u));
        } else {
            const float inf = __int_as_float(0x7f800000);
            nd[2 * k] = make_float4(inf, inf, inf, __uint_as_float(REF_EMPTY));
            nd[2 * k + 1] = make_float4(-inf, -inf, -inf, __uint_as_float(0u));
        }
    }
    atomicAdd(&counters[2], nleaf);
}

__global__ __launch_bounds__(BLK) void k_emit(const float* __restrict__ V, const uint32_t* __restrict__ I, uint32_t n,
                                              const uint32_t* __restrict__ vals, float4* __restrict__ tris) {
    const uint32_t k = blockIdx.x * BLK + threadIdx.x;
    if (k >= n) return;
    const uint32_t t = vals[k];
    const uint32_t a = I[3 * t], b = I[3 * t + 1], c = I[3 * t + 2];
    tris[3 * k + 0] = make_float4(V[3 * a], V[3 * a + 1], V[3 * a + 2], __uint_as_float(t));
    tris[3 * k + 1] = make_float4(V[3 * b], V[3 * b + 1], V[3 * b + 2], 0.f);
    tris[3 * k + 2] = make_float4(V[3 * c], V[3 * c + 1], V[3 * c + 2], 0.f);
}

// Compressed copy of every wide node (layout in crt_device.hpp): origin =
// min corner of the children's union, per-axis power-of-two step such that
// the union spans at most 255 steps, child bounds rounded outwards to steps.
__global__ __launch_bounds__(BLK) void k_compress(const float4* __restrict__ nodes, uint32_t n,
                                                  float4* __restrict__ cnodes) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const float4* nd = nodes + (size_t)i * 16u;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t refs[8];
    for (int k = 0; k < 8; ++k) {
        const float4 l = nd[2 * k], h = nd[2 * k + 1];
        refs[k] = __float_as_uint(l.w);
        if (refs[k] == REF_EMPTY) continue;
        lo[0] = fminf(lo[0], l.x); lo[1] = fminf(lo[1], l.y); lo[2] = fminf(lo[2], l.z);
        hi[0] = fmaxf(hi[0], h.x); hi[1] = fmaxf(hi[1], h.y); hi[2] = fmaxf(hi[2], h.z);
    }
    int e[3];
    double step[3];
    for (int a = 0; a < 3; ++a) {
        if (!(lo[a] <= hi[a])) { lo[a] = 0.f; hi[a] = 0.f; }
        const double ext = (double)hi[a] - (double)lo[a];
        int ea = -100;
        if (ext > 0.0) {
            ea = (int)ceil(log2(ext / 255.0));
            if (ea < -100) ea = -100;
            while (ldexp(255.0, ea) < ext) ++ea;
            while (ea > -100 && ldexp(255.0, ea - 1) >= ext) --ea;
        }
        e[a] = ea;
        step[a] = ldexp(1.0, ea);
    }
    // lo xyz, hi xyz; empty slots keep the inverted box qlo = 255, qhi = 0
    uint32_t q[6][2] = {{~0u, ~0u}, {~0u, ~0u}, {~0u, ~0u}, {0u, 0u}, {0u, 0u}, {0u, 0u}};
    for (int k = 0; k < 8; ++k) {
        if (refs[k] == REF_EMPTY) continue;
        const float4 l = nd[2 * k], h = nd[2 * k + 1];
        const float cl[3] = {l.x, l.y, l.z}, ch[3] = {h.x, h.y, h.z};
        for (int a = 0; a < 3; ++a) {
            double ql = floor(((double)cl[a] - (double)lo[a]) / step[a]);
            double qh = ceil(((double)ch[a] - (double)lo[a]) / step[a]);
            ql = fmin(fmax(ql, 0.0), 255.0);
            qh = fmin(fmax(qh, 0.0), 255.0);
            const uint32_t sh = 8u * (uint32_t)(k & 3);
            q[a][k >> 2] = (q[a][k >> 2] & ~(0xFFu << sh)) | ((uint32_t)ql << sh);
            q[3 + a][k >> 2] |= (uint32_t)qh << sh;
        }
    }
    // byte 3: mask of the non-empty slots, so traversal never depends on the
    // inverted box of an empty slot failing the slab test (a node of extent
    // below the slab test's relative slack could otherwise let one through)
    uint32_t valid = 0u;
    for (int k = 0; k < 8; ++k) valid |= (refs[k] != REF_EMPTY) ? (1u << k) : 0u;
    const uint32_t ew = ((uint32_t)(uint8_t)(int8_t)e[0]) | ((uint32_t)(uint8_t)(int8_t)e[1] << 8) |
                        ((uint32_t)(uint8_t)(int8_t)e[2] << 16) | (valid << 24);
    float4* o = cnodes + (size_t)i * 6u;
    o[0] = make_float4(lo[0], lo[1], lo[2], __uint_as_float(ew));
    o[1] = make_float4(__uint_as_float(q[0][0]), __uint_as_float(q[0][1]), __uint_as_float(q[1][0]),
                       __uint_as_float(q[1][1]));
    o[2] = make_float4(__uint_as_float(q[2][0]), __uint_as_float(q[2][1]), __uint_as_float(q[3][0]),
                       __uint_as_float(q[3][1]));
    o[3] = make_float4(__uint_as_float(q[4][0]), __uint_as_float(q[4][1]), __uint_as_float(q[5][0]),
                       __uint_as_float(q[5][1]));
    o[4] = make_float4(__uint_as_float(refs[0]), __uint_as_float(refs[1]), __uint_as_float(refs[2]),
                       __uint_as_float(refs[3]));
    o[5] = make_float4(__uint_as_float(refs[4]), __uint_as_float(refs[5]), __uint_as_float(refs[6]),
                       __uint_as_float(refs[7]));
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---- SAH of the wide BVH (node visit cost 1, triangle cost 1) ----------
// Per block, the sum over child records of area x (leaf ? triangles : 1);
// the host adds the block sums in order (deterministic) and the root.
__global__ __launch_bounds__(BLK) void k_sah_blocks(const float4* __restrict__ nodes, uint32_t n_rec,
                                                    double* __restrict__ partial) {
    __shared__ double red[BLK];
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    double a = 0.0;
    if (i < n_rec) {
        const float4 l = nodes[2 * (size_t)i], h = nodes[2 * (size_t)i + 1];
        const uint32_t ref = __float_as_uint(l.w);
        if (ref != REF_EMPTY) {
            const double dx = fmax(0.0, (double)h.x - l.x), dy = fmax(0.0, (double)h.y - l.y),
                         dz = fmax(0.0, (double)h.z - l.z);
            a = (dx * dy + dy * dz + dz * dx) * ((ref & REF_LEAF) ? (double)leaf_count(ref) : 1.0);
        }
    }
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = BLK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ---- refit (crt_scene_refit) -------------------------------------------
// Leaf-ordered triangles rewritten in place from new vertices (the original
// triangle index sits in p0.w).
__global__ __launch_bounds__(BLK) void k_refit_emit(const float* __restrict__ V, const uint32_t* __restrict__ I,
                                                    uint32_t n, float4* __restrict__ tris) {
    const uint32_t k = blockIdx.x * BLK + threadIdx.x;
    if (k >= n) return;
    const uint32_t t = __float_as_uint(tris[3 * k].w);
    const uint32_t a = I[3 * t], b = I[3 * t + 1], c = I[3 * t + 2];
    tris[3 * k + 0] = make_float4(V[3 * a], V[3 * a + 1], V[3 * a + 2], __uint_as_float(t));
    tris[3 * k + 1] = make_float4(V[3 * b], V[3 * b + 1], V[3 * b + 2], 0.f);
    tris[3 * k + 2] = make_float4(V[3 * c], V[3 * c + 1], V[3 * c + 2], 0.f);
}

// One wide level: child k of node i gets the exact union of what it holds
// (its triangles for a leaf, the exact child boxes of its node otherwise,
// already refitted since deeper levels run first), kept unpadded in `exact`
// and written padded into the node -- the build pads each child box once
// from its exact box in the same way, so unmoved vertices refit to the same
// nodes bit for bit.
__global__ __launch_bounds__(BLK) void k_refit_level(float4* __restrict__ nodes, float4* __restrict__ exact,
                                                     const float4* __restrict__ tris, uint32_t first,
                                                     uint32_t count) {
    const uint32_t t = blockIdx.x * BLK + threadIdx.x;
    if (t >= count * 8u) return;
    const uint32_t i = first + t / 8u, k = t % 8u;
    float4* rec = nodes + (size_t)i * 16u + 2u * k;
    const uint32_t ref = __float_as_uint(rec[0].w);
    if (ref == REF_EMPTY) return;
    const float inf = __int_as_float(0x7f800000);
    float lx = inf, ly = inf, lz = inf, hx = -inf, hy = -inf, hz = -inf;
    if (ref & REF_LEAF) {
        const uint32_t f = leaf_first(ref), c = leaf_count(ref);
        for (uint32_t j = 0; j < 3u * c; ++j) {
            const float4 p = tris[3 * (size_t)f + j];
            lx = fminf(lx, p.x); ly = fminf(ly, p.y); lz = fminf(lz, p.z);
            hx = fmaxf(hx, p.x); hy = fmaxf(hy, p.y); hz = fmaxf(hz, p.z);
        }
    } else {
        for (uint32_t m = 0; m < 8u; ++m) {
            if (__float_as_uint(nodes[(size_t)ref * 16u + 2u * m].w) == REF_EMPTY) continue;
            const float4 el = exact[(size_t)ref * 16u + 2u * m], eh = exact[(size_t)ref * 16u + 2u * m + 1u];
            lx = fminf(lx, el.x); ly = fminf(ly, el.y); lz = fminf(lz, el.z);
            hx = fmaxf(hx, eh.x); hy = fmaxf(hy, eh.y); hz = fmaxf(hz, eh.z);
        }
    }
    exact[(size_t)i * 16u + 2u * k] = make_float4(lx, ly, lz, 0.f);
    exact[(size_t)i * 16u + 2u * k + 1u] = make_float4(hx, hy, hz, 0.f);
    rec[0] = make_float4(pad_down(lx), pad_down(ly), pad_down(lz), rec[0].w);
    rec[1] = make_float4(pad_up(hx), pad_up(hy), pad_up(hz), rec[1].w);
}

}  // namespace

double wide_sah(const BuildResult& b, hipStream_t st) {
    const uint32_t n_rec = b.n_nodes * 8u;
    const uint32_t nb = cdiv(n_rec, BLK);
    DevBuf part;
    part.alloc((size_t)nb * 8);
    hipLaunchKernelGGL(k_sah_blocks, dim3(nb), dim3(BLK), 0, st, b.nodes.as<float4>(), n_rec, part.as<double>());
    CRT_HIP_CHECK(hipGetLastError());
    std::vector<double> hp(nb);
    float4 root[16];
    CRT_HIP_CHECK(hipMemcpyAsync(hp.data(), part.p, (size_t)nb * 8, hipMemcpyDeviceToHost, st));
    CRT_HIP_CHECK(hipMemcpyAsync(root, b.nodes.p, sizeof(root), hipMemcpyDeviceToHost, st));
    CRT_HIP_CHECK(hipStreamSynchronize(st));
    double cost = 0.0;
    for (double x : hp) cost += x;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 8; ++k) {
        uint32_t ref;
        std::memcpy(&ref, &root[2 * k].w, 4);
        if (ref == REF_EMPTY) continue;
        const float l[3] = {root[2 * k].x, root[2 * k].y, root[2 * k].z};
        const float h[3] = {root[2 * k + 1].x, root[2 * k + 1].y, root[2 * k + 1].z};
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], (double)l[a]); hi[a] = std::max(hi[a], (double)h[a]); }
    }
    const double dx = std::max(0.0, hi[0] - lo[0]), dy = std::max(0.0, hi[1] - lo[1]), dz = std::max(0.0, hi[2] - lo[2]);
    const double root_area = dx * dy + dy * dz + dz * dx;
    return root_area > 0 ? (root_area + cost) / root_area : 0.0;
}

// Level boundaries of a tree whose nodes were numbered level by level
// (scene blobs do not store them): depth of every node from the refs.
void wide_levels(BuildResult& b) {
    std::vector<float4> hn((size_t)b.n_nodes * 16);
    CRT_HIP_CHECK(hipMemcpy(hn.data(), b.nodes.p, hn.size() * 16, hipMemcpyDeviceToHost));
    std::vector<uint32_t> depth(b.n_nodes, 0);
    if (b.n_nodes) depth[0] = 1;
    b.level_end.clear();
    for (uint32_t i = 0; i < b.n_nodes; ++i) {
        CRT_REQUIRE(i == 0 || depth[i] >= depth[i - 1], CRT_ERR_STATE, "BVH nodes are not numbered by level");
        if (i > 0 && depth[i] != depth[i - 1]) b.level_end.push_back(i);
        for (int k = 0; k < 8; ++k) {
            uint32_t ref;
            std::memcpy(&ref, &hn[(size_t)i * 16 + 2 * k].w, 4);
            if (ref != REF_EMPTY && !(ref & REF_LEAF)) depth[ref] = depth[i] + 1;
        }
    }
    b.level_end.push_back(b.n_nodes);
}

// Boxes of every wide node recomputed bottom-up from the (already uploaded)
// new vertices, triangles re-emitted in leaf order, compressed nodes and the
// SAH recomputed; the tree's topology is kept.
void refit_bvh(const float* d_vertices, const uint32_t* d_indices, uint32_t n_tris, hipStream_t st, BuildResult& b,
               DevBuf& scratch) {
    if (b.level_end.empty()) wide_levels(b);
    hipEvent_t e0, e1;
    CRT_HIP_CHECK(hipEventCreate(&e0));
    CRT_HIP_CHECK(hipEventCreate(&e1));
    CRT_HIP_CHECK(hipEventRecord(e0, st));
    hipLaunchKernelGGL(k_refit_emit, dim3(cdiv(n_tris, BLK)), dim3(BLK), 0, st, d_vertices, d_indices, n_tris,
                       b.tris.as<float4>());
    scratch.reserve((size_t)b.n_nodes * 256);
    for (size_t lv = b.level_end.size(); lv-- > 0;) {
        const uint32_t first = lv == 0 ? 0u : b.level_end[lv - 1];
        const uint32_t count = b.level_end[lv] - first;
        if (count == 0) continue;
        hipLaunchKernelGGL(k_refit_level, dim3(cdiv((uint64_t)count * 8, BLK)), dim3(BLK), 0, st,
                           b.nodes.as<float4>(), scratch.as<float4>(), b.tris.as<float4>(), first, count);
    }
    hipLaunchKernelGGL(k_compress, dim3(cdiv(b.n_nodes, BLK)), dim3(BLK), 0, st, b.nodes.as<float4>(), b.n_nodes,
                       b.cnodes.as<float4>());
    CRT_HIP_CHECK(hipGetLastError());
    CRT_HIP_CHECK(hipEventRecord(e1, st));
    b.sah_cost = wide_sah(b, st);
    float ms = 0.f;
    CRT_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    b.refit_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

// PLOC rounds on the Morton-sorted leaves, then the tree written in the
// unified layout (see k_ploc_emit); `vals_out` receives the triangle order.
void build_ploc(uint32_t n, const uint32_t* sorted_vals, DevBuf& tlo, DevBuf& thi, hipStream_t st, DevBuf& left,
                DevBuf& right, DevBuf& rfirst, DevBuf& rlast, DevBuf& blo, DevBuf& bhi, DevBuf& vals_out,
                uint32_t& rounds) {
    DevBuf llo, lhi, cref[2], clo[2], chi[2], nnb, keep, bcount, nleft, nright, nlo, nhi, size, off, leaf_pos, ctr;
    llo.alloc((size_t)n * 16);
    lhi.alloc((size_t)n * 16);
    for (int k = 0; k < 2; ++k) {
        cref[k].alloc((size_t)n * 4);
        clo[k].alloc((size_t)n * 16);
        chi[k].alloc((size_t)n * 16);
    }
    nnb.alloc((size_t)n * 4);
    keep.alloc((size_t)n);
    bcount.alloc((size_t)(cdiv(n, BLK) + 1) * 4);
    nleft.alloc((size_t)(n - 1) * 4);
    nright.alloc((size_t)(n - 1) * 4);
    nlo.alloc((size_t)(n - 1) * 16);
    nhi.alloc((size_t)(n - 1) * 16);
    size.alloc((size_t)(n - 1) * 4);
    off.alloc((size_t)(n - 1) * 4);
    leaf_pos.alloc((size_t)n * 4);
    ctr.alloc(16);
    vals_out.alloc((size_t)n * 4);
    CRT_HIP_CHECK(hipMemsetAsync(ctr.p, 0, 16, st));
    hipLaunchKernelGGL(k_leaf_boxes, dim3(cdiv(n, BLK)), dim3(BLK), 0, st, (int)n, 0, sorted_vals, tlo.as<float4>(),
                       thi.as<float4>(), llo.as<float4>(), lhi.as<float4>());
    hipLaunchKernelGGL(k_ploc_init, dim3(cdiv(n, BLK)), dim3(BLK), 0, st, n, llo.as<float4>(), lhi.as<float4>(),
                       cref[0].as<uint32_t>(), clo[0].as<float4>(), chi[0].as<float4>());
    std::vector<uint32_t> starts{0u};
    uint32_t c = n;
    int cur = 0;
    while (c > 1) {
        const uint32_t nb = cdiv(c, BLK);
        hipLaunchKernelGGL(k_ploc_nn, dim3(nb), dim3(BLK), 0, st, c, clo[cur].as<float4>(), chi[cur].as<float4>(),
                           nnb.as<uint32_t>());
        hipLaunchKernelGGL(k_ploc_merge, dim3(nb), dim3(BLK), 0, st, c, nnb.as<uint32_t>(), cref[cur].as<uint32_t>(),
                           clo[cur].as<float4>(), chi[cur].as<float4>(), ctr.as<uint32_t>(), nleft.as<uint32_t>(),
                           nright.as<uint32_t>(), nlo.as<float4>(), nhi.as<float4>(), keep.as<uint8_t>(),
                           bcount.as<uint32_t>());
        hipLaunchKernelGGL(k_ploc_scan, dim3(1), dim3(BLK), 0, st, bcount.as<uint32_t>(), nb);
        hipLaunchKernelGGL(k_ploc_compact, dim3(nb), dim3(BLK), 0, st, c, keep.as<uint8_t>(), bcount.as<uint32_t>(),
                           cref[cur].as<uint32_t>(), clo[cur].as<float4>(), chi[cur].as<float4>(),
                           cref[cur ^ 1].as<uint32_t>(), clo[cur ^ 1].as<float4>(), chi[cur ^ 1].as<float4>());
        uint32_t h[2] = {0u, 0u};
        CRT_HIP_CHECK(hipMemcpyAsync(&h[0], bcount.as<uint32_t>() + nb, 4, hipMemcpyDeviceToHost, st));
        CRT_HIP_CHECK(hipMemcpyAsync(&h[1], ctr.p, 4, hipMemcpyDeviceToHost, st));
        CRT_HIP_CHECK(hipStreamSynchronize(st));
        // the closest pair of the round is always mutual, so every round merges
        CRT_REQUIRE(h[0] < c && h[1] == starts.back() + (c - h[0]), CRT_ERR_HIP, "PLOC round made no progress");
        starts.push_back(h[1]);
        c = h[0];
        cur ^= 1;
    }
    CRT_REQUIRE(starts.back() == n - 1, CRT_ERR_HIP, "PLOC node count mismatch");
    const uint32_t T = (uint32_t)starts.size() - 1;
    rounds = T;
    for (uint32_t t = 0; t < T; ++t)
        hipLaunchKernelGGL(k_ploc_sizes, dim3(cdiv(starts[t + 1] - starts[t], BLK)), dim3(BLK), 0, st, starts[t],
                           starts[t + 1], nleft.as<uint32_t>(), nright.as<uint32_t>(), size.as<uint32_t>());
    CRT_HIP_CHECK(hipMemsetAsync(off.as<uint32_t>() + (n - 2), 0, 4, st));   // the root (made last) starts at 0
    for (uint32_t t = T; t-- > 0;)
        hipLaunchKernelGGL(k_ploc_offsets, dim3(cdiv(starts[t + 1] - starts[t], BLK)), dim3(BLK), 0, st, starts[t],
                           starts[t + 1], nleft.as<uint32_t>(), nright.as<uint32_t>(), size.as<uint32_t>(),
                           off.as<uint32_t>(), leaf_pos.as<uint32_t>());
    hipLaunchKernelGGL(k_ploc_emit, dim3(cdiv(n, BLK)), dim3(BLK), 0, st, n, nleft.as<uint32_t>(),
                       nright.as<uint32_t>(), nlo.as<float4>(), nhi.as<float4>(), size.as<uint32_t>(),
                       off.as<uint32_t>(), leaf_pos.as<uint32_t>(), sorted_vals, llo.as<float4>(), lhi.as<float4>(),
                       left.as<int>(), right.as<int>(), rfirst.as<int>(), rlast.as<int>(), blo.as<float4>(),
                       bhi.as<float4>(), vals_out.as<uint32_t>());
    CRT_HIP_CHECK(hipGetLastError());
}

void build_bvh(const float* d_vertices, const uint32_t* d_indices, uint32_t n_tris, uint32_t max_leaf,
               uint32_t build_flags, hipStream_t st, BuildResult& out) {
    // PLOC clusters by surface area where the LBVH only sorts centroids; on
    // small scenes its tree is as fast or faster for every ray kind (the
    // 36-triangle Cornell box renders 25 % faster, meshes of 1k-82k
    // triangles trace sun passes 0-8 % faster), from ~200k triangles the
    // LBVH tree traces faster (1.3M-triangle body: sun pass 3.6 % faster) and
    // builds ~3x faster
    const bool use_ploc = (build_flags & CRT_BUILD_PLOC) != 0 ||
                          ((build_flags & CRT_BUILD_LBVH) == 0 && n_tris <= CRT_BUILD_PLOC_AUTO_MAX);
    const bool sah_wide = (build_flags & CRT_BUILD_GREEDY_WIDE) == 0;
    CRT_REQUIRE(n_tris > 0, CRT_ERR_INVALID, "scene has no triangles");
    CRT_REQUIRE(n_tris < (1u << 27), CRT_ERR_INVALID, "scene exceeds 2^27 triangles");
    CRT_REQUIRE(max_leaf >= 1 && max_leaf <= 16, CRT_ERR_INVALID, "max_leaf_size must be in 1..16");
    const uint32_t n = n_tris;
    hipEvent_t e0, e1;
    CRT_HIP_CHECK(hipEventCreate(&e0));
    CRT_HIP_CHECK(hipEventCreate(&e1));
    CRT_HIP_CHECK(hipEventRecord(e0, st));

    DevBuf tlo, thi, cen, bounds, keys[2], vals[2], hist;
    tlo.alloc((size_t)n * 16);
    thi.alloc((size_t)n * 16);
    cen.alloc((size_t)n * 16);
    bounds.alloc(12 * 4);
    {
        uint32_t init[12];
        for (int k = 0; k < 12; ++k) {
            const bool mn = (k < 3) || (k >= 6 && k < 9);
            init[k] = mn ? 0xFFFFFFFFu : 0u;
        }
        CRT_HIP_CHECK(hipMemcpyAsync(bounds.p, init, sizeof(init), hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(k_prep, dim3(cdiv(n, BLK)), dim3(BLK), 0, st, d_vertices, d_indices, n,
                       tlo.as<float4>(), thi.as<float4>(), cen.as<float4>(), bounds.as<uint32_t>());
    keys[0].alloc((size_t)n * 8);
    keys[1].alloc((size_t)n * 8);
    vals[0].alloc((size_t)n * 4);
    vals[1].alloc((size_t)n * 4);
    hipLaunchKernelGGL(k_morton, dim3(cdiv(n, BLK)), dim3(BLK), 0, st, cen.as<float4>(), n,
                       bounds.as<uint32_t>(), keys[0].as<uint64_t>(), vals[0].as<uint32_t>());
    const uint32_t nb = cdiv(n, RS_TILE);
    hist.alloc((size_t)256 * (nb + 1) * 4);   // per-block counts, then the 256 digit totals
    int cur = 0;
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = pass * 8;
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(BLK), 0, st, keys[cur].as<uint64_t>(), n, shift, nb,
                           hist.as<uint32_t>());
        hipLaunchKernelGGL(k_rs_scan_rows, dim3(256), dim3(BLK), 0, st, hist.as<uint32_t>(), nb,
                           hist.as<uint32_t>() + 256u * nb);
        hipLaunchKernelGGL(k_rs_scan_digits, dim3(1), dim3(BLK), 0, st, hist.as<uint32_t>() + 256u * nb);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(BLK), 0, st, keys[cur].as<uint64_t>(),
                           vals[cur].as<uint32_t>(), keys[cur ^ 1].as<uint64_t>(), vals[cur ^ 1].as<uint32_t>(),
                           n, shift, nb, hist.as<uint32_t>(), hist.as<uint32_t>() + 256u * nb);
        cur ^= 1;
    }
    const uint64_t* sk = keys[cur].as<uint64_t>();
    const uint32_t* sv = vals[cur].as<uint32_t>();

    // Binary tree in the unified layout the collapse reads: internal nodes
    // 0..n-2 (root 0) with children, leaf ranges and boxes; leaves n-1+k.
    const uint32_t nn = 2 * n - 1;
    DevBuf left, right, parent, rfirst, rlast, blo, bhi, seg_lo, seg_hi, ploc_vals;
    left.alloc((size_t)std::max(n - 1, 1u) * 4);
    right.alloc((size_t)std::max(n - 1, 1u) * 4);
    rfirst.alloc((size_t)std::max(n - 1, 1u) * 4);
    rlast.alloc((size_t)std::max(n - 1, 1u) * 4);
    blo.alloc((size_t)nn * 16);
    bhi.alloc((size_t)nn * 16);
    if (use_ploc && n > 1) {
        build_ploc(n, sv, tlo, thi, st, left, right, rfirst, rlast, blo, bhi, ploc_vals, out.ploc_rounds);
        sv = ploc_vals.as<uint32_t>();   // triangle order = depth-first leaf order of the PLOC tree
    } else {
        parent.alloc((size_t)nn * 4);
        CRT_HIP_CHECK(hipMemsetAsync(parent.p, 0xFF, (size_t)nn * 4, st));   // root parent = -1
        if (n > 1)
            hipLaunchKernelGGL(k_karras, dim3(cdiv(n - 1, BLK)), dim3(BLK), 0, st, sk, (int)n, left.as<int>(),
                               right.as<int>(), parent.as<int>(), rfirst.as<int>(), rlast.as<int>());
        hipLaunchKernelGGL(k_leaf_boxes, dim3(cdiv(n, BLK)), dim3(BLK), 0, st, (int)n, (int)(n - 1), sv,
                           tlo.as<float4>(), thi.as<float4>(), blo.as<float4>(), bhi.as<float4>());
    }
    if (n > 1 && !use_ploc) {
        SegLevels L{};
        uint32_t total = 0, sz = n;
        L.count = 1;
        while (sz > 1u) {
            sz = (sz + 1u) / 2u;
            CRT_REQUIRE(L.count < 40u, CRT_ERR_HIP, "segment tree too deep");
            L.off[L.count++] = total;
            total += sz;
        }
        seg_lo.alloc((size_t)total * 16);
        seg_hi.alloc((size_t)total * 16);
        sz = n;
        for (uint32_t lv = 1; lv < L.count; ++lv) {
            const uint32_t out_n = (sz + 1u) / 2u;
            const float4* in_lo = lv == 1 ? blo.as<float4>() + (n - 1) : seg_lo.as<float4>() + L.off[lv - 1];
            const float4* in_hi = lv == 1 ? bhi.as<float4>() + (n - 1) : seg_hi.as<float4>() + L.off[lv - 1];
            hipLaunchKernelGGL(k_seg_level, dim3(cdiv(out_n, BLK)), dim3(BLK), 0, st, in_lo, in_hi, sz,
                               seg_lo.as<float4>() + L.off[lv], seg_hi.as<float4>() + L.off[lv]);
            sz = out_n;
        }
        hipLaunchKernelGGL(k_range_boxes, dim3(cdiv(n - 1, BLK)), dim3(BLK), 0, st, (int)n, rfirst.as<int>(),
                           rlast.as<int>(), seg_lo.as<float4>(), seg_hi.as<float4>(), L, blo.as<float4>(),
                           bhi.as<float4>());
    }

    // ---- SAH-optimal collapse: binary levels top-down, DP costs bottom-up ----
    DevBuf dp_cost, dp_split, dp_flags, lvl_nodes, lvl_count;
    if (sah_wide && n > 1) {
        lvl_nodes.alloc((size_t)(n - 1) * 4);
        lvl_count.alloc(4 * 2);
        dp_cost.alloc((size_t)(n - 1) * 8 * 4);
        dp_split.alloc((size_t)(n - 1) * 8);
        dp_flags.alloc((size_t)(n - 1) * 2);
        std::vector<uint32_t> loff{0u, 1u};   // level k = lvl_nodes[loff[k] .. loff[k+1])
        const int root = 0;
        CRT_HIP_CHECK(hipMemcpyAsync(lvl_nodes.p, &root, 4, hipMemcpyHostToDevice, st));
        while (true) {
            const uint32_t a = loff[loff.size() - 2], b = loff.back();
            CRT_REQUIRE(b <= n - 1, CRT_ERR_HIP, "binary levels exceed the node count");
            CRT_HIP_CHECK(hipMemsetAsync(lvl_count.p, 0, 4, st));
            hipLaunchKernelGGL(k_bin_level, dim3(cdiv(b - a, BLK)), dim3(BLK), 0, st, (int)n,
                               lvl_nodes.as<int>() + a, b - a, lvl_nodes.as<int>() + b, lvl_count.as<uint32_t>(),
                               (n - 1) - b, left.as<int>(), right.as<int>());
            uint32_t next = 0;
            CRT_HIP_CHECK(hipMemcpyAsync(&next, lvl_count.p, 4, hipMemcpyDeviceToHost, st));
            CRT_HIP_CHECK(hipStreamSynchronize(st));
            if (next == 0) break;
            CRT_REQUIRE(b + next <= n - 1, CRT_ERR_HIP, "binary levels exceed the node count");
            loff.push_back(b + next);
        }
        for (size_t lv = loff.size() - 1; lv-- > 0;) {
            const uint32_t a = loff[lv], b = loff[lv + 1];
            hipLaunchKernelGGL(k_dp, dim3(cdiv(b - a, BLK)), dim3(BLK), 0, st, (int)n, lvl_nodes.as<int>() + a, b - a,
                               left.as<int>(), right.as<int>(), rfirst.as<int>(), rlast.as<int>(), blo.as<float4>(),
                               bhi.as<float4>(), max_leaf, dp_cost.as<float>(), dp_split.as<uint8_t>(),
                               dp_flags.as<uint16_t>());
        }
        CRT_HIP_CHECK(hipGetLastError());
    }

    // ---- collapse to 8-wide, level by level ----
    const uint32_t node_cap = std::max(n, 2u);
    DevBuf nodes_tmp, items[2], counters;
    nodes_tmp.alloc((size_t)node_cap * 256);
    items[0].alloc((size_t)node_cap * sizeof(Item));
    items[1].alloc((size_t)node_cap * sizeof(Item));
    counters.alloc(4 * 4);
    {
        const uint32_t c0[4] = {1u, 0u, 0u, 0u};
        CRT_HIP_CHECK(hipMemcpyAsync(counters.p, c0, sizeof(c0), hipMemcpyHostToDevice, st));
        const Item root{n > 1 ? 0 : (int)(n - 1), 0};
        CRT_HIP_CHECK(hipMemcpyAsync(items[0].p, &root, sizeof(Item), hipMemcpyHostToDevice, st));
    }
    uint32_t n_in = 1, depth = 0, fr = 0;
    uint32_t host_counters[4] = {0, 0, 0, 0};
    out.level_end.assign(1, 1u);   // wide node ids are assigned level by level: level 1 = the root
    while (n_in > 0) {
        ++depth;
        CRT_HIP_CHECK(hipMemsetAsync(counters.as<uint32_t>() + 1, 0, 4, st));
        if (sah_wide && n > 1)
            hipLaunchKernelGGL(k_collapse_dp, dim3(cdiv(n_in, BLK)), dim3(BLK), 0, st, (int)n, items[fr].as<Item>(),
                               n_in, items[fr ^ 1].as<Item>(), counters.as<uint32_t>(), left.as<int>(),
                               right.as<int>(), rfirst.as<int>(), rlast.as<int>(), blo.as<float4>(), bhi.as<float4>(),
                               dp_split.as<uint8_t>(), dp_flags.as<uint16_t>(), nodes_tmp.as<float4>());
        else
            hipLaunchKernelGGL(k_collapse, dim3(cdiv(n_in, BLK)), dim3(BLK), 0, st, (int)n, items[fr].as<Item>(),
                               n_in, items[fr ^ 1].as<Item>(), counters.as<uint32_t>(), left.as<int>(),
                               right.as<int>(), rfirst.as<int>(), rlast.as<int>(), blo.as<float4>(), bhi.as<float4>(),
                               max_leaf, nodes_tmp.as<float4>());
        CRT_HIP_CHECK(hipMemcpyAsync(host_counters, counters.p, 16, hipMemcpyDeviceToHost, st));
        CRT_HIP_CHECK(hipStreamSynchronize(st));
        n_in = host_counters[1];
        if (n_in > 0) out.level_end.push_back(host_counters[0]);
        fr ^= 1;
        CRT_REQUIRE(host_counters[0] <= node_cap, CRT_ERR_HIP, "wide node capacity exceeded");
    }
    out.n_nodes = host_counters[0];
    out.n_leaves = host_counters[2];
    out.max_depth = depth;
    out.nodes.alloc((size_t)out.n_nodes * 256);
    CRT_HIP_CHECK(hipMemcpyAsync(out.nodes.p, nodes_tmp.p, (size_t)out.n_nodes * 256, hipMemcpyDeviceToDevice, st));
    out.cnodes.alloc((size_t)out.n_nodes * 96);
    hipLaunchKernelGGL(k_compress, dim3(cdiv(out.n_nodes, BLK)), dim3(BLK), 0, st, out.nodes.as<float4>(), out.n_nodes,
                       out.cnodes.as<float4>());
    out.tris.alloc((size_t)n * 48);
    hipLaunchKernelGGL(k_emit, dim3(cdiv(n, BLK)), dim3(BLK), 0, st, d_vertices, d_indices, n, sv,
                       out.tris.as<float4>());
    CRT_HIP_CHECK(hipGetLastError());
    CRT_HIP_CHECK(hipEventRecord(e1, st));
    uint32_t hb[12];
    CRT_HIP_CHECK(hipMemcpyAsync(hb, bounds.p, sizeof(hb), hipMemcpyDeviceToHost, st));
    CRT_HIP_CHECK(hipStreamSynchronize(st));
    float ms = 0.f;
    CRT_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    out.build_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int k = 0; k < 6; ++k) out.bounds[k] = o2f(hb[6 + k]);

    out.sah_cost = wide_sah(out, st);
    out.sah_at_build = out.sah_cost;
}

}  // namespace crt
