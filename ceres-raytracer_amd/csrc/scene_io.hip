// scene_io.hip -- resident scene <-> byte blob (crt_scene_serialize /
// crt_scene_deserialize): a built scene (mesh in the scene frame, shading
// data, both BVH layouts) saved once and loaded without a rebuild, e.g. to
// warm-start a server's resident meshes.
//
// Blob = Header, then the sections in Header order, each padded to 16 B
// (mesh, shading data, both BVH layouts, materials, UVs, textures).
// Everything is validated before any device memory is touched: magic,
// version, sizes against the counts, an FNV-1a 64 checksum of the sections,
// and the structure itself -- every node reference in range, inner children
// numbered above their parent (so traversal always terminates), leaves of
// 1..8 triangles inside the triangle array, empty slots encoded so that they
// can never be hit, tree depth within the stored maximum, triangle and vertex
// indices in range -- so a damaged blob is rejected instead of being traced.
#include "crt_internal.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace crt {
void scene_bind_view(crt_scene* s);
uint64_t scene_device_bytes(const crt_scene* s);
}

namespace {

using namespace crt;

constexpr char MAGIC[8] = {'C', 'R', 'T', 'S', 'C', 'N', '\0', '\1'};
constexpr uint32_t VERSION = 3;   // 3: valid-slot mask in the compressed node header; material sections

enum Section { S_VERTICES, S_INDICES, S_VNORMALS, S_TRI_INST, S_INST_COLOR, S_INST_SMOOTH, S_NODES, S_CNODES,
               S_TRIS, S_MATS, S_VUV, S_TEX_INFO, S_TEXELS, N_SECTIONS };

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    uint32_t n_tris, n_vertices, n_instances, has_vnormals;
    uint32_t n_nodes, n_leaves, max_depth, pad_;
    uint32_t has_mats, has_vuv, n_textures, pad2_;   // materials (crt_scene_set_materials)
    uint64_t n_texels;
    double origin[3];
    double center[3];
    double radius;
    float bvh_bounds[6];
    crt_bvh_info info;
    uint64_t section_bytes[N_SECTIONS];
    uint64_t checksum;
};

inline uint64_t pad16(uint64_t b) { return (b + 15u) & ~(uint64_t)15u; }

uint64_t fnv1a(const uint8_t* p, uint64_t n, uint64_t h) {
    for (uint64_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

const DevBuf* section_buf(const crt_scene* s, int k) {
    switch (k) {
        case S_VERTICES: return &s->vertices;
        case S_INDICES: return &s->indices;
        case S_VNORMALS: return &s->vnormals;
        case S_TRI_INST: return &s->tri_inst;
        case S_INST_COLOR: return &s->inst_color;
        case S_INST_SMOOTH: return &s->inst_smooth;
        case S_NODES: return &s->bvh.nodes;
        case S_CNODES: return &s->bvh.cnodes;
        case S_MATS: return &s->mats;
        case S_VUV: return &s->vuv;
        case S_TEX_INFO: return &s->tex_info;
        case S_TEXELS: return &s->texels;
        default: return &s->bvh.tris;
    }
}

void expected_sizes(const Header& h, uint64_t out[N_SECTIONS]) {
    out[S_VERTICES] = (uint64_t)h.n_vertices * 12;
    out[S_INDICES] = (uint64_t)h.n_tris * 12;
    out[S_VNORMALS] = h.has_vnormals ? (uint64_t)h.n_vertices * 12 : 0;
    out[S_TRI_INST] = (uint64_t)h.n_tris * 4;
    out[S_INST_COLOR] = (uint64_t)h.n_instances * 12;
    out[S_INST_SMOOTH] = (uint64_t)h.n_instances;
    out[S_NODES] = (uint64_t)h.n_nodes * 256;
    out[S_CNODES] = (uint64_t)h.n_nodes * 96;
    out[S_TRIS] = (uint64_t)h.n_tris * 48;
    out[S_MATS] = h.has_mats ? (uint64_t)h.n_instances * sizeof(MatDev) : 0;
    out[S_VUV] = h.has_vuv ? (uint64_t)h.n_vertices * 8 : 0;
    out[S_TEX_INFO] = (uint64_t)h.n_textures * 16;
    out[S_TEXELS] = h.n_textures ? std::max<uint64_t>(h.n_texels, 1) * 4 : 0;
}

// Structure checks on the host copies (see the file comment).
void validate(const Header& h, const uint8_t* const sec[N_SECTIONS]) {
    const uint32_t nt = h.n_tris, nn = h.n_nodes;
    const uint32_t* idx = reinterpret_cast<const uint32_t*>(sec[S_INDICES]);
    for (uint64_t i = 0; i < 3ull * nt; ++i)
        CRT_REQUIRE(idx[i] < h.n_vertices, CRT_ERR_INVALID, "scene blob: vertex index out of range");
    const uint32_t* inst = reinterpret_cast<const uint32_t*>(sec[S_TRI_INST]);
    for (uint32_t i = 0; i < nt; ++i)
        CRT_REQUIRE(inst[i] >= 1 && inst[i] <= h.n_instances, CRT_ERR_INVALID, "scene blob: instance id out of range");
    const float* tris = reinterpret_cast<const float*>(sec[S_TRIS]);
    for (uint32_t i = 0; i < nt; ++i) {
        uint32_t orig;
        std::memcpy(&orig, tris + 12 * i + 3, 4);
        CRT_REQUIRE(orig < nt, CRT_ERR_INVALID, "scene blob: triangle id out of range");
    }
    // both layouts carry the same refs: 256-B records hold ref in word 3 of
    // each 32-B child, 96-B nodes hold the 8 refs in words 16..23
    const uint32_t* n256 = reinterpret_cast<const uint32_t*>(sec[S_NODES]);
    const uint32_t* n96 = reinterpret_cast<const uint32_t*>(sec[S_CNODES]);
    std::vector<uint32_t> depth(nn, 0);
    if (nn) depth[0] = 1;
    uint32_t max_depth = nn ? 1u : 0u;
    for (uint32_t i = 0; i < nn; ++i) {
        CRT_REQUIRE(i == 0 || depth[i] != 0, CRT_ERR_INVALID, "scene blob: unreachable BVH node");
        const uint32_t valid = n96[(uint64_t)i * 24 + 3] >> 24;
        for (uint32_t k = 0; k < 8; ++k) {
            CRT_REQUIRE(((valid >> k) & 1u) == (n256[(uint64_t)i * 64 + 8 * k + 3] != REF_EMPTY ? 1u : 0u),
                        CRT_ERR_INVALID, "scene blob: valid-slot mask disagrees with the node's refs");
            const uint32_t ref = n256[(uint64_t)i * 64 + 8 * k + 3];
            CRT_REQUIRE(ref == n96[(uint64_t)i * 24 + 16 + k], CRT_ERR_INVALID, "scene blob: node layouts differ");
            if (ref == REF_EMPTY) {
                // traversal relies on empty slots never being hit: their
                // quantized box must be the inverted one (qlo 255, qhi 0)
                const uint8_t* q = reinterpret_cast<const uint8_t*>(n96 + (uint64_t)i * 24 + 4);
                for (int a = 0; a < 3; ++a)
                    CRT_REQUIRE(q[8 * a + k] == 255u && q[24 + 8 * a + k] == 0u, CRT_ERR_INVALID,
                                "scene blob: empty BVH slot not encoded empty");
                continue;
            }
            if (ref & REF_LEAF) {
                const uint64_t first = leaf_first(ref), count = leaf_count(ref);
                CRT_REQUIRE(count <= 8 && first + count <= nt, CRT_ERR_INVALID, "scene blob: leaf out of range");
            } else {
                CRT_REQUIRE(ref > i && ref < nn, CRT_ERR_INVALID, "scene blob: child node out of order");
                CRT_REQUIRE(depth[ref] == 0, CRT_ERR_INVALID, "scene blob: BVH node with two parents");
                depth[ref] = depth[i] + 1;
                max_depth = std::max(max_depth, depth[ref]);
            }
        }
    }
    CRT_REQUIRE(max_depth <= h.max_depth && h.max_depth <= 64, CRT_ERR_INVALID, "scene blob: BVH deeper than stated");
    // materials: types, texture indices and texel ranges
    const uint4* ti = reinterpret_cast<const uint4*>(sec[S_TEX_INFO]);
    for (uint32_t t = 0; t < h.n_textures; ++t)
        CRT_REQUIRE(ti[t].y >= 1 && ti[t].z >= 1 && (uint64_t)ti[t].x + (uint64_t)ti[t].y * ti[t].z <= h.n_texels,
                    CRT_ERR_INVALID, "scene blob: texture outside the texel section");
    if (h.has_mats) {
        const MatDev* m = reinterpret_cast<const MatDev*>(sec[S_MATS]);
        for (uint32_t i = 0; i < h.n_instances; ++i)
            CRT_REQUIRE(m[i].type <= CRT_MAT_MIRROR && m[i].tex >= -1 && m[i].tex < (int64_t)h.n_textures &&
                            m[i].spec_tex >= -1 && m[i].spec_tex < (int64_t)h.n_textures,
                        CRT_ERR_INVALID, "scene blob: bad material");
    }
}

}  // namespace

extern "C" {

int crt_scene_serialize(const crt_scene* scene, void* buffer, uint64_t* size) {
    return guarded([&] {
        CRT_REQUIRE(scene != nullptr && size != nullptr, CRT_ERR_INVALID, "NULL argument");
        Header h;
        std::memset(&h, 0, sizeof(h));
        std::memcpy(h.magic, MAGIC, 8);
        h.version = VERSION;
        h.header_bytes = (uint32_t)sizeof(Header);
        h.n_tris = scene->n_tris;
        h.n_vertices = scene->n_vertices;
        h.n_instances = scene->n_instances;
        h.has_vnormals = scene->vnormals.p ? 1u : 0u;
        h.has_mats = scene->mats.p ? 1u : 0u;
        h.has_vuv = scene->vuv.p ? 1u : 0u;
        h.n_textures = scene->n_textures;
        h.n_texels = scene->n_textures ? scene->texels.bytes / 4 : 0;
        h.n_nodes = scene->bvh.n_nodes;
        h.n_leaves = scene->bvh.n_leaves;
        h.max_depth = scene->bvh.max_depth;
        for (int k = 0; k < 3; ++k) {
            h.origin[k] = scene->view.origin[k];
            h.center[k] = scene->view.center[k];
        }
        h.radius = scene->view.radius;
        for (int k = 0; k < 6; ++k) h.bvh_bounds[k] = scene->bvh.bounds[k];
        h.info = scene->info;
        expected_sizes(h, h.section_bytes);
        uint64_t total = pad16(sizeof(Header));
        for (int k = 0; k < N_SECTIONS; ++k) total += pad16(h.section_bytes[k]);
        if (buffer == nullptr || *size < total) {
            CRT_REQUIRE(buffer == nullptr, CRT_ERR_INVALID, "serialize buffer too small");
            *size = total;
            return;
        }
        uint8_t* out = static_cast<uint8_t*>(buffer);
        std::memset(out, 0, total);
        uint64_t off = pad16(sizeof(Header));
        uint64_t sum = 1469598103934665603ull;
        for (int k = 0; k < N_SECTIONS; ++k) {
            const uint64_t b = h.section_bytes[k];
            if (b) {
                const DevBuf* d = section_buf(scene, k);
                CRT_REQUIRE(d->p != nullptr && d->bytes >= b, CRT_ERR_HIP, "scene buffer missing");
                CRT_HIP_CHECK(hipMemcpy(out + off, d->p, b, hipMemcpyDeviceToHost));
                sum = fnv1a(out + off, b, sum);
            }
            off += pad16(b);
        }
        h.checksum = sum;
        std::memcpy(out, &h, sizeof(Header));
        *size = total;
    });
}

int crt_scene_deserialize(const void* buffer, uint64_t size, crt_scene** out_scene) {
    return guarded([&] {
        CRT_REQUIRE(out_scene != nullptr && buffer != nullptr, CRT_ERR_INVALID, "NULL argument");
        *out_scene = nullptr;
        const uint8_t* in = static_cast<const uint8_t*>(buffer);
        CRT_REQUIRE(size >= sizeof(Header), CRT_ERR_INVALID, "scene blob: truncated header");
        Header h;
        std::memcpy(&h, in, sizeof(Header));
        CRT_REQUIRE(std::memcmp(h.magic, MAGIC, 8) == 0, CRT_ERR_INVALID, "scene blob: bad magic");
        CRT_REQUIRE(h.version == VERSION && h.header_bytes == sizeof(Header), CRT_ERR_INVALID,
                    "scene blob: unsupported version");
        CRT_REQUIRE(h.n_tris > 0 && h.n_tris < (1u << 27) && h.n_vertices > 0 && h.n_instances > 0 &&
                        h.n_nodes > 0 && h.n_nodes < (1u << 23),
                    CRT_ERR_INVALID, "scene blob: bad counts");
        CRT_REQUIRE(h.has_mats <= 1 && h.has_vuv <= 1 && h.n_texels < (1ull << 32) &&
                        (h.n_textures == 0 || h.has_mats),
                    CRT_ERR_INVALID, "scene blob: bad material counts");
        uint64_t expect[N_SECTIONS];
        expected_sizes(h, expect);
        uint64_t total = pad16(sizeof(Header));
        for (int k = 0; k < N_SECTIONS; ++k) {
            CRT_REQUIRE(h.section_bytes[k] == expect[k], CRT_ERR_INVALID, "scene blob: section size mismatch");
            total += pad16(expect[k]);
        }
        CRT_REQUIRE(size >= total, CRT_ERR_INVALID, "scene blob: truncated");
        const uint8_t* sec[N_SECTIONS];
        uint64_t off = pad16(sizeof(Header));
        uint64_t sum = 1469598103934665603ull;
        for (int k = 0; k < N_SECTIONS; ++k) {
            sec[k] = in + off;
            sum = fnv1a(in + off, expect[k], sum);
            off += pad16(expect[k]);
        }
        CRT_REQUIRE(sum == h.checksum, CRT_ERR_INVALID, "scene blob: checksum mismatch");
        for (int k = 0; k < 3; ++k)
            CRT_REQUIRE(std::isfinite(h.origin[k]) && std::isfinite(h.center[k]), CRT_ERR_INVALID,
                        "scene blob: bad frame");
        CRT_REQUIRE(std::isfinite(h.radius) && h.radius > 0.0, CRT_ERR_INVALID, "scene blob: bad bounding sphere");
        validate(h, sec);

        std::unique_ptr<crt_scene> s(new crt_scene());
        CRT_HIP_CHECK(hipGetDevice(&s->device));
        s->n_tris = h.n_tris;
        s->n_vertices = h.n_vertices;
        s->n_instances = h.n_instances;
        s->n_textures = h.n_textures;
        for (int k = 0; k < N_SECTIONS; ++k) {
            if (!expect[k]) continue;
            DevBuf* d = const_cast<DevBuf*>(section_buf(s.get(), k));
            d->alloc(expect[k]);
            CRT_HIP_CHECK(hipMemcpy(d->p, sec[k], expect[k], hipMemcpyHostToDevice));
        }
        s->tri_vidx.alloc(expect[S_INDICES]);
        CRT_HIP_CHECK(hipMemcpy(s->tri_vidx.p, sec[S_INDICES], expect[S_INDICES], hipMemcpyHostToDevice));
        s->bvh.n_nodes = h.n_nodes;
        s->bvh.n_leaves = h.n_leaves;
        s->bvh.max_depth = h.max_depth;
        s->bvh.build_ms = h.info.build_ms;
        s->bvh.sah_cost = h.info.sah_cost;
        s->bvh.sah_at_build = h.info.sah_cost;
        for (int k = 0; k < 6; ++k) s->bvh.bounds[k] = h.bvh_bounds[k];
        s->info = h.info;
        for (int k = 0; k < 3; ++k) {
            s->view.origin[k] = h.origin[k];
            s->view.center[k] = h.center[k];
        }
        s->view.radius = h.radius;
        scene_bind_view(s.get());
        s->info.device_bytes = scene_device_bytes(s.get());
        CRT_REQUIRE(trace_lds_bytes(s->view) <= 64u * 1024u && lane_lds_bytes(s->view) <= 64u * 1024u &&
                        quad_lds_bytes(s->view) <= 64u * 1024u,
                    CRT_ERR_INVALID, "scene blob: BVH too deep for the LDS traversal stack");
        *out_scene = s.release();
    });
}

}  // extern "C"
