// sensors.hip -- LiDAR (simple + batch + flash) and radiation pressure.
//
// crt_lidar_batch: the reference LiDAR casts ONE ray per pose along +/-z
//   rotated by R^T (src/crt/lidars/simple_lidar.hpp:42-96) and returns
//   |hit - position| with the hit point from barycentrics, 0 on a miss
//   (src/crt/do_lidar.hpp:43-70, 111-143).  One lane per pose.
// crt_flash_lidar (new): a beam grid laid out like a pinhole camera; every
//   beam casts its boresight ray (the "single" return, equal to the reference
//   LiDAR pointed along that beam) plus `sub_rays` rays whose directions are
//   perturbed by two independent N(0, sigma^2) angles in the beam's tangent
//   plane (Box-Muller on the counter RNG, key = beam, sub-ray).  first / last
//   = min / max range over the sub-rays that hit; n_returns counts them.
// crt_radiation_pressure (new): parallel rays on an N x N grid over the
//   square circumscribing the bounding disk perpendicular to the sun; per hit
//   on a sun-facing triangle (flat outward normal n, cos = n.s > 0):
//     dF = P A_ray [ (1 - rs) d - 2 (rs cos + rd / 3) n ],  d = -s,
//     dT = (x - ref) x dF.
//   Per-ray force terms in FP32, torque arms in FP64; 64-ray wave sums in
//   plain FP64 (a 6-level tree: exact for the FP32 force terms, a few ulp
//   for torques), then block partials and the final sum in FP64 with TwoSum
//   compensation (error-free transformation, Knuth), so the result does not
//   depend on the summation order to well below 1e-12 relative.
#include "crt_internal.hpp"
#include "crt_rng.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace crt {
namespace {

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Closest hit of a unit FP64 ray (one ray per lane); returns range
// |hit - o| (FP64 from barycentrics) or 0 on a miss.
__device__ inline double range_f64(const SceneView& s, const double o[3], const double d[3], const LaneStack& stk) {
    double t0;
    if (!rebase_t0(s, o, d, t0)) return 0.0;
    const RayPre r = make_ray_pre((float)(o[0] + t0 * d[0]), (float)(o[1] + t0 * d[1]), (float)(o[2] + t0 * d[2]),
                                  (float)d[0], (float)d[1], (float)d[2]);
    const HitRec h = trace_lane<false>(s, r, 0.0f, __int_as_float(0x7f800000), stk);
    if (h.slot < 0) return 0.0;
    double t = t0 + (double)h.t, u = h.u, v = h.v;
    refine_hit(s, h.slot, o, d, t, u, v);   // FP64 barycentrics of the chosen triangle
    const float4 a = s.tris[3 * h.slot], b = s.tris[3 * h.slot + 1], c = s.tris[3 * h.slot + 2];
    const double w = 1.0 - u - v;
    const double px = u * b.x + v * c.x + w * a.x - o[0];
    const double py = u * b.y + v * c.y + w * a.y - o[1];
    const double pz = u * b.z + v * c.z + w * a.z - o[2];
    return sqrt(px * px + py * py + pz * pz);
}

CRT_LANE_KERNEL void k_lidar_batch(SceneView s, const double* __restrict__ pos,
                                                             const double* __restrict__ rot, double zsign,
                                                             double* __restrict__ range, uint64_t n) {
    const LaneStack stk = lane_stack(s);
    const uint64_t i = (uint64_t)blockIdx.x * LANE_BLOCK + threadIdx.x;
    if (i >= n) return;
    const double* R = rot + 9 * i;
    // R^T (0, 0, +/-1) = +/- third row of R
    const double d[3] = {R[6] * zsign, R[7] * zsign, R[8] * zsign};
    const double ow[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    double o[3];
    to_scene(s, ow, o);
    range[i] = range_f64(s, o, d, stk);
}

struct FlashParams {
    CamView cam;
    uint32_t sub_rays;
    double sigma;
    uint64_t seed;
};

// Direction of ray k of a beam (k = 0: the boresight; 1..sub_rays: the
// boresight tilted by two N(0, sigma^2) angles in the beam's tangent plane).
// The Gaussian angles (sigma ~ mrad) are formed in FP32: their rounding moves
// a sub-ray by ~1e-10 rad; the beam direction and the tilted direction stay FP64.
__device__ inline void flash_dir(const FlashParams& F, uint32_t beam, uint32_t k, double d[3]) {
    const int px = (int)(beam % (uint32_t)F.cam.w), py = (int)(beam / (uint32_t)F.cam.w);
    camera_dir(F.cam, (double)px, (double)py, d);
    if (k == 0) return;
    // tangent frame of the beam
    double ax[3] = {1.0, 0.0, 0.0};
    if (fabs(d[0]) > 0.9) { ax[0] = 0.0; ax[1] = 1.0; }
    double e1[3] = {d[1] * ax[2] - d[2] * ax[1], d[2] * ax[0] - d[0] * ax[2], d[0] * ax[1] - d[1] * ax[0]};
    // one FP64 division per normalisation (a division is ~10 FP64 instructions)
    const double il1 = 1.0 / sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    e1[0] *= il1; e1[1] *= il1; e1[2] *= il1;
    const double e2[3] = {d[1] * e1[2] - d[2] * e1[1], d[2] * e1[0] - d[0] * e1[2], d[0] * e1[1] - d[1] * e1[0]};
    const uint64_t key = rng_key(F.seed, beam, k);
    const double r1 = rng_u01k(key, 0), r2 = rng_u01k(key, 1);
    // Box-Muller (1 - r1 in (0, 1]: its FP32 rounding stays > 0)
    const float rad = sqrtf(-2.0f * logf((float)(1.0 - r1)));
    float sn, cs;
    sincosf(6.2831853071795865f * (float)r2, &sn, &cs);
    const float sg = (float)F.sigma;
    const double a1 = (double)tanf(sg * rad * cs), a2 = (double)tanf(sg * rad * sn);
    double nd[3] = {d[0] + a1 * e1[0] + a2 * e2[0], d[1] + a1 * e1[1] + a2 * e2[1], d[2] + a1 * e1[2] + a2 * e2[2]};
    const double iln = 1.0 / sqrt(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]);
    d[0] = nd[0] * iln; d[1] = nd[1] * iln; d[2] = nd[2] * iln;
}

// One lane per (beam, ray).  The per = sub_rays + 1 <= 64 rays of a beam sit
// in consecutive lanes of one wave (64 / per beams per wave, the remaining
// lanes idle), so the returns are reduced across lanes and the per-ray ranges
// never leave registers.  No early exit: every lane takes part in the shuffles.
// `block` is the scan's own block index (one wave per block).
__device__ inline void flash_block(const SceneView& s, const FlashParams& F, uint32_t n_beams, uint32_t block,
                                   double* __restrict__ single, double* __restrict__ first,
                                   double* __restrict__ last, uint8_t* __restrict__ n_returns,
                                   const LaneStack& stk) {
    const uint32_t per = F.sub_rays + 1, groups = LANE_BLOCK / per;
    const uint32_t lane = threadIdx.x, g = lane / per, k = lane - g * per;
    const uint64_t beam64 = (uint64_t)block * groups + g;
    const bool live = g < groups && beam64 < n_beams;
    const uint32_t beam = (uint32_t)beam64;
    double rg = 0.0;
    if (live) {
        double d[3];
        flash_dir(F, beam, k, d);
        rg = range_f64(s, F.cam.pos, d, stk);
    }
    // lane k = 0 of each group gathers lanes +1 .. +sub_rays (all inside the
    // wave: g * per + per - 1 <= 63)
    double mn = 0.0, mx = 0.0;
    uint32_t cnt = 0;
    for (uint32_t j = 1; j < per; ++j) {
        const double x = __shfl(rg, (int)min(lane + j, (uint32_t)LANE_BLOCK - 1u), LANE_BLOCK);
        if (x > 0.0) {
            mn = cnt == 0 ? x : fmin(mn, x);
            mx = cnt == 0 ? x : fmax(mx, x);
            ++cnt;
        }
    }
    if (live && k == 0) {
        if (single) single[beam] = rg;
        if (first) first[beam] = mn;
        if (last) last[beam] = mx;
        if (n_returns) n_returns[beam] = (uint8_t)cnt;
    }
}

CRT_LANE_KERNEL void k_flash(SceneView s, FlashParams F, uint32_t n_beams, double* __restrict__ single,
                             double* __restrict__ first, double* __restrict__ last, uint8_t* __restrict__ n_returns) {
    flash_block(s, F, n_beams, blockIdx.x, single, first, last, n_returns, lane_stack(s));
}

// A batch of scans in one launch: scan i owns blocks [block_begin, +blocks)
// and computes exactly what its own k_flash launch would.
struct FlashSeg {
    FlashParams F;
    uint32_t block_begin, n_beams;
    double *single, *first, *last;
    uint8_t* n_returns;
};

// largest i < n with key(i) <= x, wave-uniform (scalar loads)
template <typename F>
__device__ inline uint32_t uniform_search(uint32_t n, uint32_t x, F key) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (key(mid) <= x) lo = mid; else hi = mid;
    }
    return lo;
}

CRT_LANE_KERNEL void k_flash_multi(SceneView s, const FlashSeg* __restrict__ segs, uint32_t nseg) {
    const uint32_t i = uniform_search(nseg, blockIdx.x, [&](uint32_t j) { return segs[j].block_begin; });
    const FlashSeg& g = segs[i];
    flash_block(s, g.F, g.n_beams, blockIdx.x - g.block_begin, g.single, g.first, g.last, g.n_returns,
                lane_stack(s));
}

// ---- radiation pressure ----
struct Comp {   // compensated FP64 accumulator (value, error)
    double s, c;
};
__device__ inline Comp two_sum(double a, double b) {
    const double s = a + b;
    const double bp = s - a;
    const double e = (a - (s - bp)) + (b - bp);
    return Comp{s, e};
}
__device__ inline Comp comp_add(Comp x, Comp y) {
    const Comp t = two_sum(x.s, y.s);
    return Comp{t.s, t.c + x.c + y.c};
}

struct FluxParams {
    double sun[3], e1[3], e2[3], center[3], ref[3];
    double radius, h, pressure, rho_s, rho_d;
    double origin_back;    // distance of the grid plane from the centre, sunward
    uint32_t grid;
};

// Block partials: one wave of 64 rays per block; 7 sums each (Fx,Fy,Fz,
// Tx,Ty,Tz, hits), reduced across the wave with shuffles, then combined
// across blocks with TwoSum compensation.  `block` is the request's own
// block index; partials are component-major over its nblocks.
__device__ inline void flux_block(const SceneView& s, const FluxParams& F, double* __restrict__ partial,
                                  uint32_t nblocks, uint32_t block, const LaneStack& stk) {
    // one 8 x 8 tile of the grid per wave (tiles row-major): a wave's
    // parallel rays cover a compact patch and share BVH nodes
    const uint32_t tiles = (F.grid + 7u) / 8u;
    const uint32_t i = (block % tiles) * 8u + (threadIdx.x & 7u);
    const uint32_t j = (block / tiles) * 8u + (threadIdx.x >> 3);
    double term[7] = {0, 0, 0, 0, 0, 0, 0};
    if (i < F.grid && j < F.grid) {
        const double x = -F.radius + (i + 0.5) * F.h, y = -F.radius + (j + 0.5) * F.h;
        double o[3], d[3];
        for (int k = 0; k < 3; ++k) {
            o[k] = F.center[k] + F.origin_back * F.sun[k] + x * F.e1[k] + y * F.e2[k];
            d[k] = -F.sun[k];
        }
        double t0;
        if (rebase_t0(s, o, d, t0)) {
            const RayPre r = make_ray_pre((float)(o[0] + t0 * d[0]), (float)(o[1] + t0 * d[1]),
                                          (float)(o[2] + t0 * d[2]), (float)d[0], (float)d[1], (float)d[2]);
            const HitRec h = trace_lane<false>(s, r, 0.0f, __int_as_float(0x7f800000), stk);
            if (h.slot >= 0) {
                const float4 a = s.tris[3 * h.slot], b = s.tris[3 * h.slot + 1], c = s.tris[3 * h.slot + 2];
                // outward normal = -normalize((p0 - p1) x (p2 - p0))  (reference normal is inward)
                const float e1x = a.x - b.x, e1y = a.y - b.y, e1z = a.z - b.z;
                const float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
                float nx = -(e1y * e2z - e1z * e2y), ny = -(e1z * e2x - e1x * e2z), nz = -(e1x * e2y - e1y * e2x);
                const float inv = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz);
                nx *= inv; ny *= inv; nz *= inv;
                const float sx = (float)F.sun[0], sy = (float)F.sun[1], sz = (float)F.sun[2];
                const float cs = nx * sx + ny * sy + nz * sz;
                if (cs > 0.0f) {
                    const float pa = (float)(F.pressure * F.h * F.h);
                    const float kd = pa * (1.0f - (float)F.rho_s);
                    const float kn = pa * 2.0f * ((float)F.rho_s * cs + (float)F.rho_d / 3.0f);
                    const float fx = -sx * kd - kn * nx, fy = -sy * kd - kn * ny, fz = -sz * kd - kn * nz;
                    const double u = h.u, v = h.v, w = 1.0 - u - v;
                    const double lx = u * b.x + v * c.x + w * a.x - F.ref[0];
                    const double ly = u * b.y + v * c.y + w * a.y - F.ref[1];
                    const double lz = u * b.z + v * c.z + w * a.z - F.ref[2];
                    term[0] = fx; term[1] = fy; term[2] = fz;
                    term[3] = ly * fz - lz * fy;
                    term[4] = lz * fx - lx * fz;
                    term[5] = lx * fy - ly * fx;
                    term[6] = 1.0;
                }
            }
        }
    }
    // 64 terms in a 6-level FP64 tree: exact for the FP32 force terms, a
    // few ulp for the FP64 torques; compensation starts at the block partials.
    double acc[7];
    for (int k = 0; k < 7; ++k) acc[k] = term[k];
    for (int w = 32; w > 0; w >>= 1)
        for (int k = 0; k < 7; ++k) acc[k] += __shfl_xor(acc[k], w, 64);
    if (threadIdx.x == 0)
        for (int k = 0; k < 7; ++k) partial[(size_t)k * nblocks + block] = acc[k];   // component-major
}

CRT_LANE_KERNEL void k_flux(SceneView s, FluxParams F, double* __restrict__ partial, uint32_t nblocks) {
    flux_block(s, F, partial, nblocks, blockIdx.x, lane_stack(s));
}

// A batch of flux grids in one launch: request i owns blocks
// [block_begin, +nblocks) and partials at part_off (7 x nblocks doubles).
struct FluxSeg {
    FluxParams F;
    uint32_t block_begin, nblocks;
    uint64_t part_off;
};

// Cross-block reduction in two compensated (TwoSum) stages: FLUX_SLICES
// blocks per component each sum a contiguous slice of the block partials,
// then one block per component combines the slices.  Both stages run
// one-wave blocks: while another stream's traversal kernel holds the CUs, a
// one-wave block takes any free wave slot, where a multi-wave block waits for
// several slots to free up on one CU at once.
constexpr int FLUX_SLICES = 64;
constexpr int FLUX_SLICE_THREADS = 64;

__device__ inline Comp block_comp_sum(Comp acc, Comp* red) {
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = (int)blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = comp_add(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    return red[0];
}

__device__ inline void flux_slice(const double* __restrict__ partial, uint32_t nblocks, int k, uint32_t slice,
                                  Comp* __restrict__ slices, Comp* red) {
    const uint32_t per = (nblocks + FLUX_SLICES - 1) / FLUX_SLICES;
    const uint32_t lo = slice * per, hi = min(lo + per, nblocks);
    Comp acc{0.0, 0.0};
    for (uint32_t b = lo + threadIdx.x; b < hi; b += FLUX_SLICE_THREADS)
        acc = comp_add(acc, Comp{partial[(size_t)k * nblocks + b], 0.0});
    const Comp r = block_comp_sum(acc, red);
    if (threadIdx.x == 0) slices[k * FLUX_SLICES + slice] = r;
}

__global__ __launch_bounds__(FLUX_SLICE_THREADS) void k_flux_slices(const double* __restrict__ partial,
                                                                    uint32_t nblocks, Comp* __restrict__ slices) {
    __shared__ Comp red[FLUX_SLICE_THREADS];
    flux_slice(partial, nblocks, (int)blockIdx.y, blockIdx.x, slices, red);   // component blockIdx.y
}

__global__ __launch_bounds__(FLUX_SLICES) void k_flux_final(const Comp* __restrict__ slices,
                                                            double* __restrict__ out) {
    __shared__ Comp red[FLUX_SLICES];
    const int k = blockIdx.x;   // component 0..6
    const Comp r = block_comp_sum(slices[k * FLUX_SLICES + threadIdx.x], red);
    if (threadIdx.x == 0) out[k] = r.s + r.c;
}

// Batched forms: the same per-request reduction tree (FLUX_SLICES slices per
// component, then one combine), so each request's result is bit-identical
// to its own k_flux / k_flux_slices / k_flux_final launches.
CRT_LANE_KERNEL void k_flux_multi(SceneView s, const FluxSeg* __restrict__ segs, uint32_t nseg,
                                  double* __restrict__ partial) {
    const uint32_t i = uniform_search(nseg, blockIdx.x, [&](uint32_t j) { return segs[j].block_begin; });
    const FluxSeg& g = segs[i];
    flux_block(s, g.F, partial + g.part_off, g.nblocks, blockIdx.x - g.block_begin, lane_stack(s));
}

__global__ __launch_bounds__(FLUX_SLICE_THREADS) void k_flux_slices_multi(const FluxSeg* __restrict__ segs,
                                                                          const double* __restrict__ partial,
                                                                          Comp* __restrict__ slices) {
    __shared__ Comp red[FLUX_SLICE_THREADS];
    const uint32_t r = blockIdx.x / FLUX_SLICES;
    flux_slice(partial + segs[r].part_off, segs[r].nblocks, (int)blockIdx.y, blockIdx.x % FLUX_SLICES,
               slices + (size_t)r * 7 * FLUX_SLICES, red);
}

__global__ __launch_bounds__(FLUX_SLICES) void k_flux_final_multi(const Comp* __restrict__ slices,
                                                                  double* __restrict__ out) {
    __shared__ Comp red[FLUX_SLICES];
    const uint32_t r = blockIdx.x / 7, k = blockIdx.x % 7;
    const Comp c = block_comp_sum(slices[((size_t)r * 7 + k) * FLUX_SLICES + threadIdx.x], red);
    if (threadIdx.x == 0) out[7 * r + k] = c.s + c.c;
}

// Events only when the caller asked for timing (see EventTimer in trace.hip).
struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    explicit Timer(bool on) {
        if (!on) return;
        CRT_HIP_CHECK(hipEventCreate(&a));
        CRT_HIP_CHECK(hipEventCreate(&b));
    }
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    static void rec(hipEvent_t e, hipStream_t st) {
        if (e) CRT_HIP_CHECK(hipEventRecord(e, st));
    }
    double ms() const {
        float f = 0.f;
        CRT_HIP_CHECK(hipEventElapsedTime(&f, a, b));
        return f;
    }
};

// ---- allocation-free radiation pressure (crt_radiation_pressure_noalloc) ----
// Its scratch (the block partials, the 7 x 64 TwoSum slices) lives in one
// file-scope device array instead of a workspace.  In host code `scratch`
// names a host-side stand-in, so kernels receive the device address that
// hipGetSymbolAddress returns for the current device, kept per device.
constexpr uint32_t NOALLOC_MAX_GRID = 2048;
constexpr size_t NOALLOC_PARTIAL_BYTES = (size_t)(NOALLOC_MAX_GRID / 8) * (NOALLOC_MAX_GRID / 8) * 7 * sizeof(double);
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_PARTIAL_BYTES + (size_t)7 * FLUX_SLICES * sizeof(Comp);
static_assert(NOALLOC_PARTIAL_BYTES % 256 == 0, "slices stay 256-B aligned");

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up once per
// device, with that device current).
void* scratch_ptr() {
    int dev = 0;
    CRT_HIP_CHECK(hipGetDevice(&dev));
    CRT_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, CRT_ERR_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        CRT_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        CRT_REQUIRE(p != nullptr, CRT_ERR_HIP, "hipGetSymbolAddress returned NULL for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

// Called by crt_device_init with the device current.
void noalloc_setup_current_device() { (void)scratch_ptr(); }

}  // namespace crt

using namespace crt;

extern "C" {

int crt_lidar_batch(crt_scene* scene, uint64_t n, const double* positions, const double* rotations,
                    int32_t z_positive, double* ranges, uint32_t flags, void* stream, crt_timing* timing) {
    return guarded([&] {
        CRT_REQUIRE(scene != nullptr, CRT_ERR_INVALID, "scene is NULL");
        if (n == 0) return;
        CRT_REQUIRE(positions && rotations && ranges, CRT_ERR_INVALID, "NULL argument");
        CRT_REQUIRE(n <= (1ull << 31), CRT_ERR_INVALID, "too many poses in one call");
        hipStream_t st = (hipStream_t)stream;
        const bool dev = (flags & CRT_OUT_DEVICE) != 0;
        crt_scene::StreamSlot& slot = scene->slot(stream);
        std::lock_guard<std::mutex> slot_lock(slot.mu);
        Workspace& ws = slot.ws;
        const double* dp = positions;
        const double* dr = rotations;
        double* dout = ranges;
        if (!dev) {
            ws.f.reserve(n * 12 * 8);
            double* base = ws.f.as<double>();
            CRT_HIP_CHECK(hipMemcpyAsync(base, positions, n * 24, hipMemcpyHostToDevice, st));
            CRT_HIP_CHECK(hipMemcpyAsync(base + 3 * n, rotations, n * 72, hipMemcpyHostToDevice, st));
            dp = base;
            dr = base + 3 * n;
            ws.g.reserve(n * 8);
            dout = ws.g.as<double>();
        }
        Timer tm(timing != nullptr);
        Timer::rec(tm.a, st);
        hipLaunchKernelGGL(k_lidar_batch, dim3(cdiv(n, LANE_BLOCK)), dim3(LANE_BLOCK),
                           lane_lds_bytes(scene->view), st, scene->view, dp, dr, z_positive ? 1.0 : -1.0, dout, n);
        CRT_HIP_CHECK(hipGetLastError());
        Timer::rec(tm.b, st);
        if (!dev) CRT_HIP_CHECK(hipMemcpyAsync(ranges, dout, n * 8, hipMemcpyDeviceToHost, st));
        if (!dev || timing) CRT_HIP_CHECK(hipStreamSynchronize(st));
        if (timing) {
            std::memset(timing, 0, sizeof(*timing));
            timing->trace_ms = timing->total_ms = tm.ms();
            timing->rays_traced = n;
        }
    });
}

}  // extern "C"

namespace crt {
void flash_lidar_impl(crt_scene* scene, const crt_flash_opts* opts, double* single, double* first, double* last,
                      uint8_t* n_returns, uint32_t flags, void* stream, crt_timing* timing, Workspace& ws) {
    {
        CRT_REQUIRE(scene != nullptr, CRT_ERR_INVALID, "NULL argument");
        check_flash_opts(opts);
        hipStream_t st = (hipStream_t)stream;
        FlashParams F;
        F.cam = make_cam_view(opts->beams, scene);
        F.sub_rays = opts->sub_rays;
        F.sigma = opts->sigma;
        F.seed = opts->seed;
        const uint64_t nb = (uint64_t)F.cam.w * F.cam.h;
        const uint64_t nr = nb * (F.sub_rays + 1);
        CRT_REQUIRE(nr < (1ull << 32), CRT_ERR_INVALID, "flash LiDAR too large");
        const bool dev = (flags & CRT_OUT_DEVICE) != 0;
        double* d_single = single;
        double* d_first = first;
        double* d_last = last;
        uint8_t* d_n = n_returns;
        if (!dev) {
            ws.g.reserve(nb * 25);
            char* base = ws.g.as<char>();
            d_single = single ? reinterpret_cast<double*>(base) : nullptr;
            d_first = first ? reinterpret_cast<double*>(base + nb * 8) : nullptr;
            d_last = last ? reinterpret_cast<double*>(base + nb * 16) : nullptr;
            d_n = n_returns ? reinterpret_cast<uint8_t*>(base + nb * 24) : nullptr;
        }
        Timer tm(timing != nullptr);
        Timer::rec(tm.a, st);
        const uint32_t groups = LANE_BLOCK / (F.sub_rays + 1);   // beams per wave
        hipLaunchKernelGGL(k_flash, dim3(cdiv(nb, groups)), dim3(LANE_BLOCK), lane_lds_bytes(scene->view), st,
                           scene->view, F, (uint32_t)nb, d_single, d_first, d_last, d_n);
        CRT_HIP_CHECK(hipGetLastError());
        Timer::rec(tm.b, st);
        if (!dev) {
            if (single) CRT_HIP_CHECK(hipMemcpyAsync(single, d_single, nb * 8, hipMemcpyDeviceToHost, st));
            if (first) CRT_HIP_CHECK(hipMemcpyAsync(first, d_first, nb * 8, hipMemcpyDeviceToHost, st));
            if (last) CRT_HIP_CHECK(hipMemcpyAsync(last, d_last, nb * 8, hipMemcpyDeviceToHost, st));
            if (n_returns) CRT_HIP_CHECK(hipMemcpyAsync(n_returns, d_n, nb, hipMemcpyDeviceToHost, st));
        }
        if (!dev || timing) CRT_HIP_CHECK(hipStreamSynchronize(st));
        if (timing) {
            std::memset(timing, 0, sizeof(*timing));
            timing->trace_ms = timing->total_ms = tm.ms();
            timing->rays_traced = nr;
        }
    }
}

namespace {

// Ray lattice of a flux request in the scene frame (see crt_flux_opts).
FluxParams flux_params(const crt_scene* scene, const crt_flux_opts* opts) {
    CRT_REQUIRE(scene != nullptr && opts != nullptr, CRT_ERR_INVALID, "NULL argument");
    CRT_REQUIRE(opts->grid >= 1 && opts->grid <= 65536, CRT_ERR_INVALID, "grid must be in 1..65536");
    const double sl = std::sqrt(opts->sun_dir[0] * opts->sun_dir[0] + opts->sun_dir[1] * opts->sun_dir[1] +
                                opts->sun_dir[2] * opts->sun_dir[2]);
    CRT_REQUIRE(sl > 0 && std::isfinite(sl), CRT_ERR_INVALID, "sun_dir must be a finite non-zero vector");
    FluxParams F;
    for (int k = 0; k < 3; ++k) F.sun[k] = opts->sun_dir[k] / sl;
    const double ax[3] = {std::fabs(F.sun[0]) < 0.9 ? 1.0 : 0.0, std::fabs(F.sun[0]) < 0.9 ? 0.0 : 1.0, 0.0};
    double e1[3] = {F.sun[1] * ax[2] - F.sun[2] * ax[1], F.sun[2] * ax[0] - F.sun[0] * ax[2],
                    F.sun[0] * ax[1] - F.sun[1] * ax[0]};
    const double l1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    for (int k = 0; k < 3; ++k) F.e1[k] = e1[k] / l1;
    F.e2[0] = F.sun[1] * F.e1[2] - F.sun[2] * F.e1[1];
    F.e2[1] = F.sun[2] * F.e1[0] - F.sun[0] * F.e1[2];
    F.e2[2] = F.sun[0] * F.e1[1] - F.sun[1] * F.e1[0];
    // grid centre and torque reference point in the scene frame
    if (opts->radius > 0) {
        for (int k = 0; k < 3; ++k) F.center[k] = opts->center[k] - scene->view.origin[k];
        F.radius = opts->radius;
    } else {
        for (int k = 0; k < 3; ++k) F.center[k] = scene->view.center[k];
        F.radius = scene->view.radius;
    }
    for (int k = 0; k < 3; ++k) F.ref[k] = opts->ref_point[k] - scene->view.origin[k];
    F.grid = opts->grid;
    F.h = 2.0 * F.radius / (double)F.grid;
    F.pressure = opts->pressure;
    F.rho_s = opts->rho_s;
    F.rho_d = opts->rho_d;
    F.origin_back = 2.0 * F.radius;
    return F;
}

inline uint32_t flux_blocks(const FluxParams& F) { return cdiv(F.grid, 8) * cdiv(F.grid, 8); }   // 8 x 8 tiles

// Enqueue the flux kernels: `part` holds 7 x flux_blocks doubles, `slices`
// 7 x FLUX_SLICES Comp; out[0..6] = force xyz, torque xyz, hits.
void flux_launch(const crt_scene* scene, const FluxParams& F, double* part, Comp* slices, double* out,
                 hipStream_t st) {
    const uint32_t nblocks = flux_blocks(F);
    hipLaunchKernelGGL(k_flux, dim3(nblocks), dim3(LANE_BLOCK), lane_lds_bytes(scene->view), st, scene->view, F,
                       part, nblocks);
    hipLaunchKernelGGL(k_flux_slices, dim3(FLUX_SLICES, 7), dim3(FLUX_SLICE_THREADS), 0, st, part, nblocks,
                       slices);
    hipLaunchKernelGGL(k_flux_final, dim3(7), dim3(FLUX_SLICES), 0, st, slices, out);
    CRT_HIP_CHECK(hipGetLastError());
}

}  // namespace

void radiation_pressure_impl(crt_scene* scene, const crt_flux_opts* opts, double force[3], double torque[3],
                             uint64_t* n_hits, void* stream, crt_timing* timing, Workspace& ws) {
    {
        CRT_REQUIRE(force && torque, CRT_ERR_INVALID, "NULL argument");
        const FluxParams F = flux_params(scene, opts);
        hipStream_t st = (hipStream_t)stream;
        const uint64_t n = (uint64_t)F.grid * F.grid;
        const uint32_t nblocks = flux_blocks(F);
        ws.h.reserve((size_t)nblocks * 7 * sizeof(double) + (size_t)7 * FLUX_SLICES * sizeof(Comp));
        ws.counters.reserve(256);
        double* out = ws.counters.as<double>();
        Timer tm(timing != nullptr);
        Timer::rec(tm.a, st);
        double* part = ws.h.as<double>();
        flux_launch(scene, F, part, reinterpret_cast<Comp*>(part + (size_t)nblocks * 7), out, st);
        Timer::rec(tm.b, st);
        double h[7];
        CRT_HIP_CHECK(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, st));
        CRT_HIP_CHECK(hipStreamSynchronize(st));
        for (int k = 0; k < 3; ++k) { force[k] = h[k]; torque[k] = h[3 + k]; }
        if (n_hits) *n_hits = (uint64_t)std::llround(h[6]);
        if (timing) {
            std::memset(timing, 0, sizeof(*timing));
            timing->trace_ms = timing->total_ms = tm.ms();
            timing->rays_traced = n;
        }
    }
}

void check_flash_opts(const crt_flash_opts* opts) {
    CRT_REQUIRE(opts != nullptr, CRT_ERR_INVALID, "NULL argument");
    check_camera_args(&opts->beams);
    CRT_REQUIRE(opts->sub_rays < (uint32_t)LANE_BLOCK, CRT_ERR_INVALID, "at most 63 sub-rays per beam");
    CRT_REQUIRE(opts->sigma >= 0.0 && std::isfinite(opts->sigma), CRT_ERR_INVALID, "sigma must be >= 0");
    const uint64_t nb = (uint64_t)std::floor(opts->beams.resolution[0]) * (uint64_t)std::floor(opts->beams.resolution[1]);
    CRT_REQUIRE(nb * (opts->sub_rays + 1) < (1ull << 32), CRT_ERR_INVALID, "flash LiDAR too large");
}

void check_flux_opts(const crt_flux_opts* opts) {
    CRT_REQUIRE(opts != nullptr, CRT_ERR_INVALID, "NULL argument");
    CRT_REQUIRE(opts->grid >= 1 && opts->grid <= 65536, CRT_ERR_INVALID, "grid must be in 1..65536");
    const double sl = std::sqrt(opts->sun_dir[0] * opts->sun_dir[0] + opts->sun_dir[1] * opts->sun_dir[1] +
                                opts->sun_dir[2] * opts->sun_dir[2]);
    CRT_REQUIRE(sl > 0 && std::isfinite(sl), CRT_ERR_INVALID, "sun_dir must be a finite non-zero vector");
    bool fin = std::isfinite(opts->pressure) && std::isfinite(opts->rho_s) && std::isfinite(opts->rho_d) &&
               std::isfinite(opts->radius);
    for (int k = 0; k < 3; ++k) fin = fin && std::isfinite(opts->center[k]) && std::isfinite(opts->ref_point[k]);
    CRT_REQUIRE(fin, CRT_ERR_INVALID, "flux options must be finite");
}

// Several flash-LiDAR scans of one scene in one launch; outputs (single /
// first / last per scan, each beams long) are device pointers.  The device
// segment table lives in ws.e; `table` keeps its host copy alive until the
// caller synchronises the stream.
void flash_lidar_multi(crt_scene* scene, uint32_t n, const crt_flash_opts* opts, double* const* outs,
                       hipStream_t st, Workspace& ws, std::vector<uint8_t>& table_bytes) {
    table_bytes.resize((size_t)n * sizeof(FlashSeg));
    FlashSeg* table = reinterpret_cast<FlashSeg*>(table_bytes.data());
    uint32_t blocks = 0;
    for (uint32_t i = 0; i < n; ++i) {
        check_flash_opts(&opts[i]);
        FlashSeg& g = table[i];
        g.F.cam = make_cam_view(opts[i].beams, scene);
        g.F.sub_rays = opts[i].sub_rays;
        g.F.sigma = opts[i].sigma;
        g.F.seed = opts[i].seed;
        g.n_beams = (uint32_t)((uint64_t)g.F.cam.w * g.F.cam.h);
        g.block_begin = blocks;
        g.single = outs[3 * i];
        g.first = outs[3 * i + 1];
        g.last = outs[3 * i + 2];
        g.n_returns = nullptr;
        const uint64_t nb = cdiv(g.n_beams, LANE_BLOCK / (g.F.sub_rays + 1));
        CRT_REQUIRE((uint64_t)blocks + nb < (1ull << 31), CRT_ERR_INVALID, "LiDAR batch too large");
        blocks += (uint32_t)nb;
    }
    if (blocks == 0) return;
    ws.e.reserve(table_bytes.size());
    CRT_HIP_CHECK(hipMemcpyAsync(ws.e.p, table, table_bytes.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_flash_multi, dim3(blocks), dim3(LANE_BLOCK), lane_lds_bytes(scene->view), st, scene->view,
                       ws.e.as<FlashSeg>(), n);
    CRT_HIP_CHECK(hipGetLastError());
}

// Several flux grids of one scene in one launch sequence; out7 (device, 7 x n
// doubles) receives force xyz, torque xyz and hits per request.
void radiation_pressure_multi(crt_scene* scene, uint32_t n, const crt_flux_opts* opts, double* out7,
                              hipStream_t st, Workspace& ws, std::vector<uint8_t>& table_bytes) {
    table_bytes.resize((size_t)n * sizeof(FluxSeg));
    FluxSeg* table = reinterpret_cast<FluxSeg*>(table_bytes.data());
    uint64_t blocks = 0, parts = 0;
    for (uint32_t i = 0; i < n; ++i) {
        check_flux_opts(&opts[i]);
        FluxSeg& g = table[i];
        g.F = flux_params(scene, &opts[i]);
        g.nblocks = flux_blocks(g.F);
        g.block_begin = (uint32_t)blocks;
        g.part_off = parts;
        blocks += g.nblocks;
        parts += 7ull * g.nblocks;
        CRT_REQUIRE(blocks < (1ull << 31), CRT_ERR_INVALID, "flux batch too large");
    }
    if (n == 0) return;
    ws.e.reserve(table_bytes.size());
    ws.h.reserve((size_t)parts * sizeof(double) + (size_t)n * 7 * FLUX_SLICES * sizeof(Comp));
    double* part = ws.h.as<double>();
    Comp* slices = reinterpret_cast<Comp*>(part + parts);
    CRT_HIP_CHECK(hipMemcpyAsync(ws.e.p, table, table_bytes.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_flux_multi, dim3((uint32_t)blocks), dim3(LANE_BLOCK), lane_lds_bytes(scene->view), st,
                       scene->view, ws.e.as<FluxSeg>(), n, part);
    hipLaunchKernelGGL(k_flux_slices_multi, dim3(FLUX_SLICES * n, 7), dim3(FLUX_SLICE_THREADS), 0, st,
                       ws.e.as<FluxSeg>(), part, slices);
    hipLaunchKernelGGL(k_flux_final_multi, dim3(7 * n), dim3(FLUX_SLICES), 0, st, slices, out7);
    CRT_HIP_CHECK(hipGetLastError());
}

}  // namespace crt

extern "C" {

int crt_flash_lidar(crt_scene* scene, const crt_flash_opts* opts, double* single, double* first, double* last,
                    uint8_t* n_returns, uint32_t flags, void* stream, crt_timing* timing) {
    return guarded([&] {
        CRT_REQUIRE(scene != nullptr, CRT_ERR_INVALID, "scene is NULL");
        crt_scene::StreamSlot& slot = scene->slot(stream);
        std::lock_guard<std::mutex> slot_lock(slot.mu);
        flash_lidar_impl(scene, opts, single, first, last, n_returns, flags, stream, timing, slot.ws);
    });
}

int crt_radiation_pressure(crt_scene* scene, const crt_flux_opts* opts, double force[3], double torque[3],
                           uint64_t* n_hits, void* stream, crt_timing* timing) {
    return guarded([&] {
        CRT_REQUIRE(scene != nullptr, CRT_ERR_INVALID, "scene is NULL");
        crt_scene::StreamSlot& slot = scene->slot(stream);
        std::lock_guard<std::mutex> slot_lock(slot.mu);
        radiation_pressure_impl(scene, opts, force, torque, n_hits, stream, timing, slot.ws);
    });
}

int crt_radiation_pressure_noalloc(crt_scene* scene, const crt_flux_opts* opts, double* out7, void* stream) {
    return guarded([&] {
        CRT_REQUIRE(scene != nullptr && opts != nullptr && out7 != nullptr, CRT_ERR_INVALID, "NULL argument");
        CRT_REQUIRE(opts->grid <= NOALLOC_MAX_GRID, CRT_ERR_INVALID,
                    "grid exceeds the allocation-free maximum of 2048 (use crt_radiation_pressure)");
        int dev = 0;
        CRT_HIP_CHECK(hipGetDevice(&dev));
        CRT_REQUIRE(dev == scene->device, CRT_ERR_INVALID, "the scene's device must be current");
        const FluxParams F = flux_params(scene, opts);
        unsigned char* base = static_cast<unsigned char*>(scratch_ptr());
        flux_launch(scene, F, reinterpret_cast<double*>(base), reinterpret_cast<Comp*>(base + NOALLOC_PARTIAL_BYTES),
                    out7, (hipStream_t)stream);
    });
}

}  // extern "C"
