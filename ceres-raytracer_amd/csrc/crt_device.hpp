// crt_device.hpp -- device-side data layout and the octet traversal core.
//
// BVH layout in HBM ("wide node"): 8 child records of 32 B each, 256 B per
// node, 256-B aligned.  Child record k = { lo.xyz, ref | hi.xyz, meta }.
// A ray is traced by one *octet*: 8 consecutive lanes of a 64-wide wave.
// Lane k of the octet reads child record k (two 16-B loads), so the octet
// reads the node as one contiguous 256-B block (two 128-B cache lines) and
// tests all 8 child boxes in one step.  In leaves, lane k tests triangle k.
//
// ref encoding (also the traversal stack entry):
//   inner child : wide node index                      (bit 31 clear)
//   leaf child  : 0x80000000 | first << 4 | (count-1)  (first < 2^27, count <= 16)
//   empty slot  : 0xFFFFFFFF (its box is empty too)
//
// Triangles are stored in leaf order as 3 float4 (48 B): {p0, orig index
// bits}, {p1, 0}, {p2, 0}, so a leaf of 8 triangles is 384 contiguous bytes.
//
// The triangle test is the watertight test of Woop, Benthin and Wald (JCGT
// 2013): shear/scale into the ray frame, exact-sign 2D edge functions with an
// FP64 fall-back when an edge function is exactly zero.  Acceptance is
// inclusive (edge functions of one sign or zero, tmin <= t <= tmax) like
// lib/bvh/triangle.hpp:150-156.  Barycentrics follow the reference convention
// hit = u*p1 + v*p2 + (1-u-v)*p0 (src/crt/passes.hpp:57).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace crt {

constexpr uint32_t REF_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t REF_LEAF = 0x80000000u;
constexpr int OCTET = 8;
// Traversal kernels of every form (lane, octet, quad) run one wave per
// block: a block's LDS stacks are released as soon as its wave ends, so
// short waves (rays that miss) never hold a long-running neighbour's slot,
// and a one-wave block takes any free wave slot while another stream's
// kernels fill the CUs (octet/quad render kernels of a 256-camera batch
// +8.5 % over 4-wave blocks; the sun pass's quad shadow kernel on 1.8M rays
// 0.35 vs 0.39 ms).
constexpr int TRACE_BLOCK = 64;                  // threads per octet / quad traversal block
constexpr int OCTETS_PER_BLOCK = TRACE_BLOCK / OCTET;
constexpr int LANE_BLOCK = 64;
// Lane kernels ask for 8 waves per SIMD (at most 64 VGPRs): the traversal is
// latency-bound and needs every resident wave it can get.
#define CRT_LANE_KERNEL __global__ __launch_bounds__(LANE_BLOCK) __attribute__((amdgpu_waves_per_eu(8)))

struct __align__(16) WideChild {
    float lo[3];
    uint32_t ref;
    float hi[3];
    uint32_t meta;
};
static_assert(sizeof(WideChild) == 32, "child record is 32 B");

__host__ __device__ inline uint32_t make_leaf_ref(uint32_t first, uint32_t count) {
    return REF_LEAF | (first << 4) | (count - 1u);
}
__host__ __device__ inline uint32_t leaf_first(uint32_t ref) { return (ref >> 4) & 0x07FFFFFFu; }
__host__ __device__ inline uint32_t leaf_count(uint32_t ref) { return (ref & 15u) + 1u; }

// Per-instance material (crt_scene_set_materials), 48 B:
// type 0 ColoredLambertianMaterial, 1 TexturedLambertianMaterial, 2
// TexturedBlinnPhongMaterial, 3 MirrorMaterial (src/crt/materials/
// material.hpp:90-184); tex / spec_tex index a texture or are -1 for the
// constant colour / (ka, kd, ks).
struct MatDev {
    uint32_t type;
    int32_t tex;
    int32_t spec_tex;
    float alpha;
    float color[4];
    float coeffs[4];
};
static_assert(sizeof(MatDev) == 48, "material record is 48 B");

// Device view of a scene (passed by value to kernels).
struct SceneView {
    const float4* nodes;        // n_wide_nodes * 8 * 2 float4
    const float4* cnodes;       // n_wide_nodes * 6 float4 (compressed, lane traversal)
    const float4* tris;         // n_tris * 3 float4, leaf order
    const uint32_t* tri_vidx;   // n_tris * 3, original order
    const float* vnormals;      // n_vertices * 3 (may be null)
    const uint32_t* tri_inst;   // n_tris, original order
    const float* inst_color;    // n_instances * 3
    const uint8_t* inst_smooth; // n_instances
    const MatDev* mats;         // n_instances, or null: ColoredLambertian with inst_color
    const float2* vuv;          // n_vertices (u, v), or null: zero (the reference OBJ loader sets none)
    const uint32_t* texels;     // RGBA8 texels of every texture, row-major, row 0 first
    const uint4* tex_info;      // per texture: first texel, width, height
    uint32_t n_tris;
    uint32_t n_instances;
    uint32_t stack_cap;         // per-octet stack entries
    uint32_t lane_cap;          // per-lane group stack entries (max_depth + 1)
    double center[3];           // bounding sphere for FP64 rebasing (scene frame)
    double radius;
    double origin[3];           // world position of the scene frame's origin
};

// Scene frame.  Triangles and nodes are stored relative to `origin` (the
// scene's box centre rounded to the FP32 ulp of its largest coordinate),
// so FP32 traversal resolves the scene at its own size however far it sits
// from the world origin.  FP64 world points are moved in before rebasing and
// hit points moved out after.
__device__ inline void to_scene(const SceneView& s, const double w[3], double l[3]) {
    l[0] = w[0] - s.origin[0]; l[1] = w[1] - s.origin[1]; l[2] = w[2] - s.origin[2];
}

// Closest-hit result of one traversal (local FP32 frame of the rebased ray).
struct HitRec {
    float t;
    float u, v;
    int32_t slot;   // triangle index in leaf order, -1 = miss
};

// ------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------
__device__ inline float min3f(float a, float b, float c) { return fminf(a, fminf(b, c)); }
__device__ inline float max3f(float a, float b, float c) { return fmaxf(a, fmaxf(b, c)); }

__device__ inline float safe_rcp(float x) {
    // Same intent as bvh::Vector::safe_inverse (lib/bvh/vector.hpp:75-80):
    // tiny components are clamped so 1/x stays finite; the FP32 clamp keeps
    // products with scene coordinates far below FLT_MAX.  v_rcp_f32 (1 ulp):
    // child boxes are padded by several ulp, so slab tests stay conservative.
    const float eps = 1e-20f;
    return __builtin_amdgcn_rcpf(fabsf(x) < eps ? copysignf(eps, x) : x);
}

__device__ inline float sel3(float a, float b, float c, int k) {
    return k == 0 ? a : (k == 1 ? b : c);
}

// ------------------------------------------------------------------------
// Octet lane exchange through DPP (ALU-path data movement, no LDS trip).
// quad_perm permutes within each group of 4 lanes; row_half_mirror reverses
// each group of 8 (lane i <- lane 7 - i, i.e. i ^ 7 inside an octet).
// Every xor distance 1..7 inside an octet is one or two of these.  The
// lanes of an octet are always active together, so no source is disabled.
// ------------------------------------------------------------------------
constexpr int DPP_XOR1 = 0xB1;          // quad_perm [1,0,3,2]
constexpr int DPP_XOR2 = 0x4E;          // quad_perm [2,3,0,1]
constexpr int DPP_XOR3 = 0x1B;          // quad_perm [3,2,1,0]
constexpr int DPP_HALF_MIRROR = 0x141;  // row_half_mirror

template <int CTRL>
__device__ inline uint32_t dpp(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ inline float dppf(float x) {
    return __uint_as_float(dpp<CTRL>(__float_as_uint(x)));
}

// Per-ray constants for the watertight test and the slab test.
struct RayPre {
    float ox, oy, oz;          // origin
    float ix, iy, iz;          // 1/d (clamped)
    float nx, ny, nz;          // -o/d
    int kx, ky, kz;            // axis permutation, kz = dominant axis
    float sx, sy, sz;          // shear constants
};

__device__ inline RayPre make_ray_pre(float ox, float oy, float oz, float dx, float dy, float dz) {
    RayPre r;
    r.ox = ox; r.oy = oy; r.oz = oz;
    r.ix = safe_rcp(dx); r.iy = safe_rcp(dy); r.iz = safe_rcp(dz);
    r.nx = -ox * r.ix; r.ny = -oy * r.iy; r.nz = -oz * r.iz;
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    const float dkz = sel3(dx, dy, dz, kz);
    if (dkz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    // shear constants: any fixed values keep the test watertight (every
    // triangle sees the same transform); 1-ulp reciprocals are enough
    r.sz = __builtin_amdgcn_rcpf(dkz);
    r.sx = sel3(dx, dy, dz, kx) * r.sz;
    r.sy = sel3(dx, dy, dz, ky) * r.sz;
    return r;
}

// Component K (0..2) of (a, b, c), or component k chosen per lane (K = 3).
template <int K>
__device__ inline float pick3(float a, float b, float c, int k) {
    if constexpr (K == 0) return a;
    else if constexpr (K == 1) return b;
    else if constexpr (K == 2) return c;
    else return sel3(a, b, c, k);
}

// Watertight ray/triangle test.  Returns true on hit with tmin <= t <= tmax;
// u = weight of p1, v = weight of p2 (reference convention).  KX, KY, KZ:
// the ray's axis permutation when the caller knows it for the whole wave
// (then no per-lane selects), or 3 to take it from the ray; either way the
// arithmetic, and so the result, is the same.
template <int KX = 3, int KY = 3, int KZ = 3>
__device__ inline bool tri_test(const RayPre& r, float4 a, float4 b, float4 c,
                                float tmin, float tmax, float& t_out, float& u_out, float& v_out) {
    // vertices relative to the origin, permuted into (kx, ky, kz)
    const float Ax0 = a.x - r.ox, Ay0 = a.y - r.oy, Az0 = a.z - r.oz;
    const float Bx0 = b.x - r.ox, By0 = b.y - r.oy, Bz0 = b.z - r.oz;
    const float Cx0 = c.x - r.ox, Cy0 = c.y - r.oy, Cz0 = c.z - r.oz;
    const float Akx = pick3<KX>(Ax0, Ay0, Az0, r.kx), Aky = pick3<KY>(Ax0, Ay0, Az0, r.ky),
                Akz = pick3<KZ>(Ax0, Ay0, Az0, r.kz);
    const float Bkx = pick3<KX>(Bx0, By0, Bz0, r.kx), Bky = pick3<KY>(Bx0, By0, Bz0, r.ky),
                Bkz = pick3<KZ>(Bx0, By0, Bz0, r.kz);
    const float Ckx = pick3<KX>(Cx0, Cy0, Cz0, r.kx), Cky = pick3<KY>(Cx0, Cy0, Cz0, r.ky),
                Ckz = pick3<KZ>(Cx0, Cy0, Cz0, r.kz);
    // shear: one explicit FMA per coordinate, so a vertex shared by two
    // triangles gets identical bits in both
    const float Ax = fmaf(-r.sx, Akz, Akx);
    const float Ay = fmaf(-r.sy, Akz, Aky);
    const float Bx = fmaf(-r.sx, Bkz, Bkx);
    const float By = fmaf(-r.sy, Bkz, Bky);
    const float Cx = fmaf(-r.sx, Ckz, Ckx);
    const float Cy = fmaf(-r.sy, Ckz, Cky);
    // Edge functions: both products rounded, then subtracted.  HIP's
    // __fmul_rn/__fsub_rn are plain operators that hipcc's default
    // -ffp-contract=fast-honor-pragmas fuses into fma(a, b, -round(c d)),
    // which breaks the exact antisymmetry on shared edges (cracks) and gives
    // a zero-area triangle nonzero edge functions of one sign (a false hit);
    // the pragma keeps this block unfused.
    float U, V, W;
    {
#pragma clang fp contract(off)
        U = Cx * By - Cy * Bx;
        V = Ax * Cy - Ay * Cx;
        W = Bx * Ay - By * Ax;
    }
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        const double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        const double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        const double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        U = (float)(CxBy - CyBx);
        V = (float)(AxCy - AyCx);
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    const float rcp = __builtin_amdgcn_rcpf(det);   // v_rcp_f32, 1 ulp
    const float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t;
    u_out = V * rcp;
    v_out = W * rcp;
    return true;
}

// The ray's axis permutation as a code 0..5 (2 kz + 1 if kx and ky are
// swapped), and a call of f with the permutation as template arguments
// (PermTag<KX, KY, KZ>) for a code that is the same on the whole wave.
template <int X, int Y, int Z>
struct PermTag {
    static constexpr int kx = X, ky = Y, kz = Z;
};

__device__ inline int perm_code(const RayPre& r) {
    const int kx0 = r.kz == 2 ? 0 : r.kz + 1;
    return 2 * r.kz + (r.kx != kx0 ? 1 : 0);
}

// Wave-uniform permutation code of the active lanes, or -1 when they differ.
__device__ inline int wave_perm(const RayPre& r) {
    const int pc = perm_code(r);
    const int pc0 = __builtin_amdgcn_readfirstlane(pc);
    return __ballot(pc != pc0) == 0ull ? pc0 : -1;
}

template <bool ON, typename F>
__device__ inline bool perm_dispatch(int pc, F&& f) {
    if constexpr (!ON) {
        return f(PermTag<3, 3, 3>{});
    } else {
        switch (pc) {
            case 0: return f(PermTag<1, 2, 0>{});
            case 1: return f(PermTag<2, 1, 0>{});
            case 2: return f(PermTag<2, 0, 1>{});
            case 3: return f(PermTag<0, 2, 1>{});
            case 4: return f(PermTag<0, 1, 2>{});
            case 5: return f(PermTag<1, 0, 2>{});
            default: return f(PermTag<3, 3, 3>{});
        }
    }
}

// ------------------------------------------------------------------------
// Octet traversal (closest hit).  All 8 lanes of an octet call this with the
// same ray and stay converged inside (every branch below is octet-uniform).
// `stk_ref` / `stk_t` point at this octet's LDS stack of `cap` entries.
// (Any-hit rays use trace_quad_any.)
// ------------------------------------------------------------------------
// Dynamic LDS of a traversal block: 32 stacks of refs, then 32 stacks of
// tnear (stack_cap entries each).
__device__ inline uint32_t* traversal_lds(const SceneView&) {
    extern __shared__ uint32_t crt_dyn_lds[];
    return crt_dyn_lds;
}

__device__ inline HitRec trace_octet(const SceneView& s, const RayPre& r, float tmin, float tmax,
                                     uint32_t* stk_ref, float* stk_t, int sub, int lane_base) {
    HitRec hit;
    hit.t = tmax; hit.u = 0.f; hit.v = 0.f; hit.slot = -1;
    const float4* __restrict__ nodes = s.nodes;
    const float4* __restrict__ tris = s.tris;
    const uint32_t cap = s.stack_cap;
    float best = tmax;
    uint32_t cur = 0u;           // root wide node
    uint32_t sp = 0u;
    while (true) {
        if (!(cur & REF_LEAF)) {
            // ---- inner node: lane `sub` tests child record `sub` ----
            const float4 q0 = nodes[(size_t)cur * 16u + 2u * sub];
            const float4 q1 = nodes[(size_t)cur * 16u + 2u * sub + 1u];
            const uint32_t ref = __float_as_uint(q0.w);
            const float tx0 = fmaf(q0.x, r.ix, r.nx), tx1 = fmaf(q1.x, r.ix, r.nx);
            const float ty0 = fmaf(q0.y, r.iy, r.ny), ty1 = fmaf(q1.y, r.iy, r.ny);
            const float tz0 = fmaf(q0.z, r.iz, r.nz), tz1 = fmaf(q1.z, r.iz, r.nz);
            const float tnear = max3f(fminf(tx0, tx1), fminf(ty0, ty1), fmaxf(fminf(tz0, tz1), tmin));
            const float tfar = min3f(fmaxf(tx0, tx1), fmaxf(ty0, ty1), fminf(fmaxf(tz0, tz1), best));
            const bool h = (ref != REF_EMPTY) && (tnear <= tfar * 1.00000024f);
            const uint32_t m = (uint32_t)((__ballot(h) >> lane_base) & 0xFFull);
            if (m != 0u) {
                const int nh = __popc(m);
                // nearest hit child by tnear (>= 0, so its bits without the
                // sign order like the value; the low 3 bits carry the lane),
                // min-reduced over the octet in three DPP steps
                const uint32_t key = h ? ((__float_as_uint(tnear) & 0x7FFFFFF8u) | (uint32_t)sub) : 0xFFFFFFFFu;
                uint32_t kmin = min(key, dpp<DPP_XOR1>(key));
                kmin = min(kmin, dpp<DPP_XOR2>(kmin));
                kmin = min(kmin, dpp<DPP_HALF_MIRROR>(kmin));
                const int first_lane = (int)(kmin & 7u);
                const uint32_t next = (uint32_t)__shfl((int)ref, lane_base + first_lane, 64);
                // the other hit children are pushed in child order: most
                // bounce rays leave the body and visit every child they
                // cross whatever the order; pops still skip entries beyond
                // the best hit
                if (h && sub != first_lane) {
                    const uint32_t below = m & ((1u << sub) - 1u) & ~(1u << first_lane);
                    const uint32_t pos = sp + (uint32_t)__popc(below);
                    if (pos < cap) {
                        stk_ref[pos] = ref;
                        stk_t[pos] = tnear;
                    }
                }
                sp += (uint32_t)(nh - 1);
                if (sp > cap) sp = cap;   // never happens: cap >= 7 * depth (checked at build)
                cur = next;
                continue;
            }
        } else {
            // ---- leaf (at most 8 triangles, checked at scene creation):
            //      lane `sub` tests triangle first + sub ----
            const uint32_t first = leaf_first(cur);
            const uint32_t count = leaf_count(cur);
            float t = __int_as_float(0x7f800000);
            float u = 0.f, v = 0.f;
            bool ok = false;
            if ((uint32_t)sub < count) {
                const size_t ti = (size_t)(first + (uint32_t)sub) * 3u;
                ok = tri_test(r, tris[ti], tris[ti + 1], tris[ti + 2], tmin, best, t, u, v);
                if (!ok) t = __int_as_float(0x7f800000);
            }
            const uint32_t mh = (uint32_t)((__ballot(ok) >> lane_base) & 0xFFull);
            if (mh != 0u) {
                // octet min-reduce of (t, lane); exact ties go to the later
                // triangle, like the "last tested wins" rule of
                // lib/bvh/single_ray_traverser.hpp:55-60
                // (t >= 0 or +inf: its bits, -0 folded to +0, order like the value)
                uint32_t bk = __float_as_uint(t) & 0x7FFFFFFFu;
                uint32_t bl = (uint32_t)sub;
                {
                    uint32_t ok_ = dpp<DPP_XOR1>(bk), ol = dpp<DPP_XOR1>(bl);
                    if (ok_ < bk || (ok_ == bk && ol > bl)) { bk = ok_; bl = ol; }
                    ok_ = dpp<DPP_XOR2>(bk); ol = dpp<DPP_XOR2>(bl);
                    if (ok_ < bk || (ok_ == bk && ol > bl)) { bk = ok_; bl = ol; }
                    ok_ = dpp<DPP_HALF_MIRROR>(bk); ol = dpp<DPP_HALF_MIRROR>(bl);
                    if (ok_ < bk || (ok_ == bk && ol > bl)) { bk = ok_; bl = ol; }
                }
                const float bt = __uint_as_float(bk);
                if (bt <= best) {
                    best = bt;
                    hit.t = bt;
                    hit.u = __shfl(u, lane_base + (int)bl, 64);
                    hit.v = __shfl(v, lane_base + (int)bl, 64);
                    hit.slot = (int)first + (int)bl;
                }
            }
        }
        // ---- pop the next entry that can still hold a closer hit ----
        bool found = false;
        while (sp > 0u) {
            --sp;
            const uint32_t e = stk_ref[sp];
            const float te = stk_t[sp];
            if (te <= best) { cur = e; found = true; break; }
        }
        if (!found) break;
    }
    return hit;
}

// ------------------------------------------------------------------------
// Quad any-hit traversal: one shadow ray per *quad* (4 lanes, 16 rays per
// wave) over the 256-B nodes.  Lane j of the quad tests children j and j+4
// and, in a leaf, triangles j and j+4, so the per-step bookkeeping (ballot,
// next-node shuffle, stack pushes) is shared by twice as many rays as with
// an octet.  Any hit never tightens the bound, so the stack holds refs only
// and pops are unconditional.  Returns true when some triangle blocks
// [tmin, tmax].  `stk` is the quad's stack of s.stack_cap refs.
// ------------------------------------------------------------------------
constexpr int QUAD = 4;

__device__ inline bool child_hit(const float4 q0, const float4 q1, const RayPre& r, float tmin, float tmax) {
    const float tx0 = fmaf(q0.x, r.ix, r.nx), tx1 = fmaf(q1.x, r.ix, r.nx);
    const float ty0 = fmaf(q0.y, r.iy, r.ny), ty1 = fmaf(q1.y, r.iy, r.ny);
    const float tz0 = fmaf(q0.z, r.iz, r.nz), tz1 = fmaf(q1.z, r.iz, r.nz);
    const float tnear = max3f(fminf(tx0, tx1), fminf(ty0, ty1), fmaxf(fminf(tz0, tz1), tmin));
    const float tfar = min3f(fmaxf(tx0, tx1), fmaxf(ty0, ty1), fminf(fmaxf(tz0, tz1), tmax));
    return __float_as_uint(q0.w) != REF_EMPTY && tnear <= tfar * 1.00000024f;
}

// Any hit is an OR over all triangles, so testing one triangle before the
// traversal cannot change the answer.  Shadow rays test the triangle their
// origin lies on first: when the light is behind that surface the ray
// re-enters it within a few offsets and ends after one test instead of a
// traversal through the body.  Every lane of the quad/octet tests the same
// triangle, so the result is uniform across the group.
__device__ inline bool hint_blocks(const SceneView& s, const RayPre& r, uint32_t slot, float tmin, float tmax) {
    if (slot >= s.n_tris) return false;
    const size_t ti = (size_t)slot * 3u;
    float t, u, v;
    return tri_test(r, s.tris[ti], s.tris[ti + 1], s.tris[ti + 2], tmin, tmax, t, u, v);
}

template <bool FIXED_PERM = true>
__device__ inline bool trace_quad_any(const SceneView& s, const RayPre& r, float tmin, float tmax, uint32_t* stk,
                                      int j, int lane_base) {
    const float4* __restrict__ nodes = s.nodes;
    const float4* __restrict__ tris = s.tris;
    const uint32_t cap = s.stack_cap;
    // leaves test with a fixed axis permutation when the wave shares one
    // (FIXED_PERM: off where the extra code would cost occupancy)
    const int pc = FIXED_PERM ? wave_perm(r) : -1;
    uint32_t cur = 0u;
    uint32_t sp = 0u;
    while (true) {
        if (!(cur & REF_LEAF)) {
            const float4* nd = nodes + (size_t)cur * 16u;
            const float4 a0 = nd[2 * j], a1 = nd[2 * j + 1];
            const float4 b0 = nd[2 * j + 8], b1 = nd[2 * j + 9];
            const bool ha = child_hit(a0, a1, r, tmin, tmax);
            const bool hb = child_hit(b0, b1, r, tmin, tmax);
            // child mask of the quad: bit c for child c (lane c & 3, slot c >> 2)
            const uint32_t m = (uint32_t)((__ballot(ha) >> lane_base) & 0xFull) |
                               ((uint32_t)((__ballot(hb) >> lane_base) & 0xFull) << 4);
            if (m != 0u) {
                const int fc = __ffs(m) - 1;
                const uint32_t sel = (fc >= 4) ? __float_as_uint(b0.w) : __float_as_uint(a0.w);
                const uint32_t next = (uint32_t)__shfl((int)sel, lane_base + (fc & 3), 64);
                const uint32_t rest = m & ~(1u << fc);
                if (ha && ((rest >> j) & 1u)) {
                    const uint32_t pos = sp + (uint32_t)__popc(rest & ((1u << j) - 1u));
                    if (pos < cap) stk[pos] = __float_as_uint(a0.w);
                }
                if (hb && ((rest >> (j + 4)) & 1u)) {
                    const uint32_t pos = sp + (uint32_t)__popc(rest & ((1u << (j + 4)) - 1u));
                    if (pos < cap) stk[pos] = __float_as_uint(b0.w);
                }
                sp = min(sp + (uint32_t)__popc(rest), cap);   // cap >= 7 * depth + 1: never clamps
                cur = next;
                continue;
            }
        } else {
            // leaf of at most 8 triangles: lane j tests j and j + 4
            const uint32_t first = leaf_first(cur);
            const uint32_t count = leaf_count(cur);
            const bool hit = perm_dispatch<FIXED_PERM>(pc, [&](auto P) {
                constexpr int KX = decltype(P)::kx, KY = decltype(P)::ky, KZ = decltype(P)::kz;
                float t, u, v;
                bool h = false;
                if ((uint32_t)j < count) {
                    const size_t ti = (size_t)(first + (uint32_t)j) * 3u;
                    h = tri_test<KX, KY, KZ>(r, tris[ti], tris[ti + 1], tris[ti + 2], tmin, tmax, t, u, v);
                }
                if (!h && (uint32_t)j + 4u < count) {
                    const size_t ti = (size_t)(first + (uint32_t)j + 4u) * 3u;
                    h = tri_test<KX, KY, KZ>(r, tris[ti], tris[ti + 1], tris[ti + 2], tmin, tmax, t, u, v);
                }
                return h;
            });
            if (((__ballot(hit) >> lane_base) & 0xFull) != 0ull) return true;
        }
        if (sp == 0u) return false;
        cur = stk[--sp];
    }
}

// ------------------------------------------------------------------------
// Lane traversal: one ray per lane, 64 independent rays per wave.
//
// A lane evaluates all 8 children of a (compressed) wide node itself.  The stack
// holds node *groups*: (node << 8 | mask of children still to visit, nearest
// tnear among them).  At a node the nearest hit child is entered and the
// other hit children are pushed as one group; a popped group whose nearest
// child is beyond the current closest hit is dropped without a load,
// otherwise its children are re-tested against the closer bound.  The stack
// therefore holds at most one entry per tree level (max_depth entries).
//
// LDS layout per block: stack word i of thread t at [i * blockDim + t]
// (group) and [(cap + i) * blockDim + t] (tnear): lanes of a wave always hit
// 64 consecutive words whatever their stack depths.
// ------------------------------------------------------------------------
struct LaneStack {
    uint32_t* g;       // this lane's group words, stride `stride`
    float* t;          // this lane's tnear words, stride `stride`
    uint32_t stride;
};

__device__ inline LaneStack lane_stack(const SceneView& s) {
    extern __shared__ uint32_t crt_lane_lds[];
    LaneStack k;
    k.stride = blockDim.x;
    k.g = crt_lane_lds + threadIdx.x;
    k.t = reinterpret_cast<float*>(crt_lane_lds + (size_t)s.lane_cap * blockDim.x) + threadIdx.x;
    return k;
}

// Compressed wide node (lane traversal), 96 B = 6 x 16 B, written by the
// build from the 256-B node with the same child order:
//   [0] origin p.xyz (float), w = exponents ex | ey << 8 | ez << 16 (int8)
//       | valid << 24 (bit k set: slot k holds a child)
//   [1] qlo_x[0..7] | qlo_y[0..7]      (u8, child k in byte k)
//   [2] qlo_z[0..7] | qhi_x[0..7]
//   [3] qhi_y[0..7] | qhi_z[0..7]
//   [4] ref[0..3]   [5] ref[4..7]      (same refs as the 256-B node)
// Child k's box on axis a is [p_a + 2^e_a qlo_a[k], p_a + 2^e_a qhi_a[k]],
// rounded outwards from the padded FP32 box, so it always contains it; an
// empty slot has qlo = 255, qhi = 0 on every axis.
// A lane reads the whole node with 6 loads instead of 16; the slab test per
// plane stays one FMA, t = q * (2^e / d) + (p - o) / d, and the ray octant
// picks near and far planes per axis for 4 children at a time.
__device__ inline float cvt_u8(uint32_t w, uint32_t byte) {
    return (float)((w >> (8u * byte)) & 0xFFu);
}

// Median of three unsigned words (one VALU instruction).
__device__ inline uint32_t med3_u32(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t d;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

// Test the children of compressed node `cn` in `mask` against [tmin, best]
// (tmin >= 0).  Returns the hit mask; k0/ref0 = nearest hit child, tn1 <=
// the entry distance of every other hit child (their nearest, rounded down
// by at most 7 ulp).
__device__ inline uint32_t cnode_test(const float4* __restrict__ cn, const RayPre& r, float tmin, float best,
                                      uint32_t mask, uint32_t& k0, uint32_t& ref0, float& tn1) {
    const float4 h = cn[0];
    const uint4 a = __builtin_bit_cast(uint4, cn[1]);
    const uint4 b = __builtin_bit_cast(uint4, cn[2]);
    const uint4 c = __builtin_bit_cast(uint4, cn[3]);
    const uint4 r0 = __builtin_bit_cast(uint4, cn[4]);
    const uint4 r1 = __builtin_bit_cast(uint4, cn[5]);
    const uint32_t ew = __float_as_uint(h.w);
    mask &= ew >> 24;   // empty slots are never tested
    const float sx = __int_as_float(((int)(int8_t)(ew & 0xFFu) + 127) << 23);
    const float sy = __int_as_float(((int)(int8_t)((ew >> 8) & 0xFFu) + 127) << 23);
    const float sz = __int_as_float(((int)(int8_t)((ew >> 16) & 0xFFu) + 127) << 23);
    const float Ax = sx * r.ix, Ay = sy * r.iy, Az = sz * r.iz;
    const float Bx = (h.x - r.ox) * r.ix, By = (h.y - r.oy) * r.iy, Bz = (h.z - r.oz) * r.iz;
    // near / far planes: lo for a positive direction component, hi for a negative one
    const bool nx = r.ix < 0.f, ny = r.iy < 0.f, nz = r.iz < 0.f;
    const uint32_t nx0 = nx ? b.z : a.x, nx1 = nx ? b.w : a.y, fx0 = nx ? a.x : b.z, fx1 = nx ? a.y : b.w;
    const uint32_t ny0 = ny ? c.x : a.z, ny1 = ny ? c.y : a.w, fy0 = ny ? a.z : c.x, fy1 = ny ? a.w : c.y;
    const uint32_t nz0 = nz ? c.z : b.x, nz1 = nz ? c.w : b.y, fz0 = nz ? b.x : c.z, fz1 = nz ? b.y : c.w;
    // Nearest and second nearest as sort keys: tnear's bits (tnear >= +-0,
    // so the unsigned order of the bits is the order of the distances) with
    // the slot in the low 3 bits; an untested or missed child is ~0.  Per
    // child one min and one median keep the two smallest keys (ties go to
    // the lower slot); the nearest child's ref is selected once at the end.
    uint32_t n0 = 0xFFFFFFFFu, n1 = 0xFFFFFFFFu;
    uint32_t hits = 0u;
#pragma unroll
    for (uint32_t k = 0; k < 8u; ++k) {
        const uint32_t byte = k & 3u;
        const bool hi4 = k >= 4u;
        const float tnx = fmaf(cvt_u8(hi4 ? nx1 : nx0, byte), Ax, Bx);
        const float tny = fmaf(cvt_u8(hi4 ? ny1 : ny0, byte), Ay, By);
        const float tnz = fmaf(cvt_u8(hi4 ? nz1 : nz0, byte), Az, Bz);
        const float tfx = fmaf(cvt_u8(hi4 ? fx1 : fx0, byte), Ax, Bx);
        const float tfy = fmaf(cvt_u8(hi4 ? fy1 : fy0, byte), Ay, By);
        const float tfz = fmaf(cvt_u8(hi4 ? fz1 : fz0, byte), Az, Bz);
        const float tnear = fmaxf(max3f(tnx, tny, tnz), tmin);
        const float tfar = fminf(min3f(tfx, tfy, tfz), best);
        // empty slots are out of `mask` (and stored inverted, qlo 255, qhi 0)
        const bool hk = ((mask >> k) & 1u) && (tnear <= tfar * 1.00000024f);
        hits |= hk ? (1u << k) : 0u;
        const uint32_t key = hk ? ((__float_as_uint(tnear) & 0x7FFFFFF8u) | k) : 0xFFFFFFFFu;
        n1 = med3_u32(n0, n1, key);   // n0 <= n1: the second smallest of the three
        n0 = min(n0, key);
    }
    k0 = n0 & 7u;
    const uint4 rq = (k0 & 4u) ? r1 : r0;
    const uint32_t rp0 = (k0 & 2u) ? rq.z : rq.x, rp1 = (k0 & 2u) ? rq.w : rq.y;
    ref0 = (k0 & 1u) ? rp1 : rp0;
    tn1 = __uint_as_float(n1 & ~7u);
    return hits;
}

// Traversal counters of one ray, the counterpart of the reference
// traverser's Statistics (lib/bvh/single_ray_traverser.hpp:130-160): wide
// nodes visited, child boxes tested (the valid children of the node's mask
// at each visit; the reference tests two per traversal step) and triangles
// tested (a leaf counts all its triangles, as intersect_leaf does).
struct TraversalStats {
    uint32_t nodes, boxes, tris;
};

// Lane traversal in "while-while" form (Aila and Laine, HPG 2009): a lane
// that reaches a leaf parks it and keeps descending until it meets a second
// leaf or at least half of the lanes still descending hold one; then the
// wave tests triangles together, so inner-node and leaf work of different
// lanes are not interleaved step by step.  (Waiting for every lane, as the
// original form does, leaves about half of the lanes idle in the node loop;
// leaving at half tests leaves sooner, whose hits then cut node visits:
// the 2048^2 sun pass on the 1.3M-triangle body 8280-8408 -> 8662-8684
// Mrays/s, against 8515 when leaving after every step and 8591-8661 at
// three quarters.)
template <bool ANY_HIT, bool STATS = false, bool FIXED_PERM = true>
__device__ inline HitRec trace_lane(const SceneView& s, const RayPre& r, float tmin, float tmax,
                                    const LaneStack& stk, TraversalStats* ts = nullptr) {
    HitRec hit;
    hit.t = tmax; hit.u = 0.f; hit.v = 0.f; hit.slot = -1;
    float best = tmax;
    // leaves test with a fixed axis permutation when the wave shares one
    // (FIXED_PERM: off where the extra code would cost more than it saves)
    const int pc = FIXED_PERM ? wave_perm(r) : -1;
    uint32_t cur = 0xFFu;      // group (node << 8 | mask), leaf ref, or 0 = nothing
    uint32_t leaf = 0u;        // parked leaf, 0 = none
    uint32_t sp = 0u;
    while (cur != 0u || leaf != 0u) {
        // ---- inner nodes ----
        while (cur != 0u && !(cur & REF_LEAF)) {
            uint32_t k0, ref0;
            float tn1;
            if (STATS) {
                ++ts->nodes;
                ts->boxes += (uint32_t)__popc((cur & 0xFFu) & (__float_as_uint(s.cnodes[(size_t)(cur >> 8) * 6u].w) >> 24));
            }
            const uint32_t hits = cnode_test(s.cnodes + (size_t)(cur >> 8) * 6u, r, tmin, best, cur & 0xFFu, k0,
                                             ref0, tn1);
            uint32_t next = 0u;
            if (hits != 0u) {
                const uint32_t rest = hits & ~(1u << k0);
                if (rest != 0u && sp < s.lane_cap) {   // sp < lane_cap always: one entry per level
                    stk.g[sp * stk.stride] = (cur & ~0xFFu) | rest;
                    stk.t[sp * stk.stride] = tn1;
                    ++sp;
                }
                next = (ref0 & REF_LEAF) ? ref0 : ((ref0 << 8) | 0xFFu);
            }
            if (next == 0u || ((next & REF_LEAF) && leaf == 0u)) {
                if (next != 0u) leaf = next;
                next = 0u;
                while (sp > 0u) {
                    --sp;
                    if (stk.t[sp * stk.stride] <= best) { next = stk.g[sp * stk.stride]; break; }
                }
            }
            cur = next;
            // at least half of the lanes still here hold a parked leaf
            if (2u * (uint32_t)__popcll(__ballot(leaf == 0u)) <= (uint32_t)__popcll(__ballot(true))) break;
        }
        // ---- leaves: the parked one, then cur if it is a leaf ----
        while (leaf != 0u || (cur & REF_LEAF)) {
            const bool parked = leaf != 0u;
            const uint32_t lr = parked ? leaf : cur;
            const uint32_t first = leaf_first(lr);
            const uint32_t count = leaf_count(lr);
            const float4* __restrict__ tp = s.tris + (size_t)first * 3u;
            if (STATS) ts->tris += count;
            const bool any = perm_dispatch<FIXED_PERM>(pc, [&](auto P) {
                constexpr int KX = decltype(P)::kx, KY = decltype(P)::ky, KZ = decltype(P)::kz;
                for (uint32_t k = 0; k < count; ++k) {
                    float t, u, v;
                    if (tri_test<KX, KY, KZ>(r, tp[3u * k], tp[3u * k + 1u], tp[3u * k + 2u], tmin, best, t, u, v)) {
                        // t <= best: exact ties go to the later triangle, like the
                        // "last tested wins" rule of lib/bvh/single_ray_traverser.hpp:55-60
                        hit.t = t; hit.u = u; hit.v = v; hit.slot = (int)(first + k);
                        if (ANY_HIT) return true;
                        best = t;
                    }
                }
                return false;
            });
            if (ANY_HIT && any) return hit;
            if (parked) {
                leaf = 0u;
            } else {
                cur = 0u;
                while (sp > 0u) {
                    --sp;
                    if (stk.t[sp * stk.stride] <= best) { cur = stk.g[sp * stk.stride]; break; }
                }
            }
        }
    }
    return hit;
}

// ------------------------------------------------------------------------
// Pixel order for coherent waves: the i-th pixel of a w x h image in tiles
// of 64 pixels, `tw` wide (8: 8x8, Z-order inside; 4, 16, 32: 4x16, 16x4,
// 32x2, row-major inside), row-major over the tiles, then the pixels that no
// whole tile covers (the right strip row by row, then the bottom rows).
// A bijection of [0, w*h): consecutive i stay in compact blocks, so a wave's
// rays share BVH nodes.
// ------------------------------------------------------------------------
__host__ __device__ inline uint32_t tile_order(uint32_t i, uint32_t w, uint32_t h, uint32_t tw = 8u) {
    const uint32_t th = 64u / tw;
    const uint32_t ntx = w / tw, nty = h / th;
    const uint32_t nfull = ntx * nty * 64u;
    if (i < nfull) {
        const uint32_t tile = i >> 6, z = i & 63u;
        const uint32_t tx = tile % ntx, ty = tile / ntx;
        uint32_t x, y;
        if (tw == 8u) {
            x = (z & 1u) | ((z >> 1) & 2u) | ((z >> 2) & 4u);
            y = ((z >> 1) & 1u) | ((z >> 2) & 2u) | ((z >> 3) & 4u);
        } else {
            x = z % tw;
            y = z / tw;
        }
        return (ty * th + y) * w + tx * tw + x;
    }
    i -= nfull;
    const uint32_t rw = w - ntx * tw;
    const uint32_t nright = nty * th * rw;
    if (i < nright) return (i / rw) * w + ntx * tw + i % rw;
    i -= nright;
    return (nty * th + i / w) * w + i % w;
}

// ------------------------------------------------------------------------
// FP64 helpers: camera rays and rebasing
// ------------------------------------------------------------------------
struct CamView {
    double pos[3];
    double rot[9];       // row-major R; world dir = R^T * cam dir
    double cx, cy, sx, sy, f;   // centre, scale (res/sensor), +/-focal
    int w, h;
};

// src/crt/cameras/simple_camera.hpp:61-81 -- pixel (u,v) -> unit direction.
__device__ inline void camera_dir(const CamView& c, double u, double v, double d[3]) {
    double x = (-c.cx + u) / c.sx;
    double y = (c.cy - v) / c.sy;
    double z = c.f;
    const double inv = 1.0 / sqrt(x * x + y * y + z * z);
    x *= inv; y *= inv; z *= inv;
    d[0] = c.rot[0] * x + c.rot[3] * y + c.rot[6] * z;
    d[1] = c.rot[1] * x + c.rot[4] * y + c.rot[7] * z;
    d[2] = c.rot[2] * x + c.rot[5] * y + c.rot[8] * z;
}

// FP64 re-evaluation of a hit on triangle `slot` (Moller-Trumbore on the FP32
// vertices): t along the FP64 ray from o, u = weight of p1, v = weight of p2.
// The FP32 watertight test decides *which* triangle is hit; this sharpens the
// numbers, which FP32 resolves poorly on needle-thin triangles.  Returns
// false (outputs untouched) if the triangle is degenerate in FP64 as well.
__device__ inline bool refine_hit(const SceneView& s, int slot, const double o[3], const double d[3], double& t,
                                  double& u, double& v) {
    const float4 a = s.tris[3 * slot], b = s.tris[3 * slot + 1], c = s.tris[3 * slot + 2];
    const double e1x = (double)b.x - a.x, e1y = (double)b.y - a.y, e1z = (double)b.z - a.z;
    const double e2x = (double)c.x - a.x, e2y = (double)c.y - a.y, e2z = (double)c.z - a.z;
    const double px = d[1] * e2z - d[2] * e2y, py = d[2] * e2x - d[0] * e2z, pz = d[0] * e2y - d[1] * e2x;
    const double det = e1x * px + e1y * py + e1z * pz;
    if (det == 0.0) return false;
    const double inv = 1.0 / det;
    const double tx = o[0] - a.x, ty = o[1] - a.y, tz = o[2] - a.z;
    const double qx = ty * e1z - tz * e1y, qy = tz * e1x - tx * e1z, qz = tx * e1y - ty * e1x;
    u = (tx * px + ty * py + tz * pz) * inv;
    v = (d[0] * qx + d[1] * qy + d[2] * qz) * inv;
    t = (e2x * qx + e2y * qy + e2z * qz) * inv;
    return true;
}

// Distance along a unit FP64 ray to the entry of the scene's bounding sphere;
// returns false when the ray misses the sphere (no triangle can be hit).
__device__ inline bool rebase_t0(const SceneView& s, const double o[3], const double d[3], double& t0) {
    const double px = o[0] - s.center[0], py = o[1] - s.center[1], pz = o[2] - s.center[2];
    const double b = px * d[0] + py * d[1] + pz * d[2];
    const double c = px * px + py * py + pz * pz - s.radius * s.radius;
    const double disc = b * b - c;
    if (disc < 0.0) return false;
    const double sq = sqrt(disc);
    if (-b + sq < 0.0) return false;     // sphere entirely behind the origin
    // stop a little before the sphere so FP32 rounding of the new origin
    // cannot land inside geometry that touches the sphere
    const double te = -b - sq - 1e-6 * s.radius;
    t0 = te > 0.0 ? te : 0.0;
    return true;
}

}  // namespace crt
