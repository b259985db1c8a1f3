// pycrt.cpp -- the `_crt` extension module: C++ host over the HIP C ABI.
//
// Same module name, classes, constructors, methods, argument order, dtypes
// and shapes as the reference binding (src/pybind.cpp:130-718):
//   SimpleCamera, SimpleLidar, PointLight, AreaLight, Entity,
//   BodyFixedEntity, BodyFixedGroup, and module functions render,
//   simulate_lidar, intersection_pass, instance_pass, normal_pass.
// Behind it, BodyFixedGroup owns a crt_scene (mesh + GPU-built BVH resident
// in HBM); every pass / render / LiDAR call is one C-ABI call with the GIL
// released.  Additive API: DirectionalLight, BodyFixedGroup.render_radiance,
// depth_pass, triangle_pass, sun_pass, flash_lidar, radiation_pressure,
// bvh_info.
//
// Error behaviour: where the reference crashes or silently drops input
// (unknown camera/lidar -> null dereference, src/pybind.cpp:83-105; unknown
// light skipped, :476-483; unreadable mesh -> empty triangle list,
// src/crt/rendering_dynamic/entity.hpp:38-40) this module raises TypeError /
// ValueError / RuntimeError instead.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "crt_hip.h"
#include "obj_loader.hpp"

namespace py = pybind11;

namespace {

using darr = py::array_t<double, py::array::c_style | py::array::forcecast>;

void check(int code) {
    if (code == CRT_OK) return;
    const std::string msg = crt_last_error();
    switch (code) {
        case CRT_ERR_INVALID: throw py::value_error(msg);
        case CRT_ERR_TYPE: throw py::type_error(msg);
        case CRT_ERR_OOM: throw std::bad_alloc();
        default: throw std::runtime_error("crt: " + msg);
    }
}

void ensure_device() {
    static bool done = false;
    if (!done) {
        check(crt_device_init(0));
        done = true;
    }
}

struct Pose {
    double position[3] = {0, 0, 0};
    double rotation[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

    void set_position(darr p) {
        if (p.size() < 3) throw py::value_error("position needs 3 values");
        const double* d = p.data();
        for (int k = 0; k < 3; ++k) position[k] = d[k];
    }
    void set_rotation(darr r) {
        if (r.size() < 9) throw py::value_error("rotation needs 3x3 values");
        const double* d = r.data();
        for (int k = 0; k < 9; ++k) rotation[k] = d[k];
    }
    void set_pose(darr p, darr r) {
        set_position(p);
        set_rotation(r);
    }
};

struct SimpleCamera : Pose {
    double focal_length, resolution[2], sensor_size[2];
    bool z_positive;
    crt_camera to_c() const {
        crt_camera c{};
        std::memcpy(c.position, position, sizeof(position));
        std::memcpy(c.rotation, rotation, sizeof(rotation));
        c.focal_length = focal_length;
        c.resolution[0] = resolution[0];
        c.resolution[1] = resolution[1];
        c.sensor_size[0] = sensor_size[0];
        c.sensor_size[1] = sensor_size[1];
        c.z_positive = z_positive ? 1 : 0;
        return c;
    }
    int width() const { return (int)std::floor(resolution[0]); }
    int height() const { return (int)std::floor(resolution[1]); }
};

SimpleCamera make_camera(double f, py::list res, py::list sensor, bool z_positive) {
    // src/pybind.cpp:35-46
    SimpleCamera c;
    c.focal_length = f;
    c.resolution[0] = res[0].cast<double>();
    c.resolution[1] = res[1].cast<double>();
    c.sensor_size[0] = sensor[0].cast<double>();
    c.sensor_size[1] = sensor[1].cast<double>();
    c.z_positive = z_positive;
    return c;
}

struct SimpleLidar : Pose {
    bool z_positive = false;
    std::vector<double> batch_pos;   // n x 3
    std::vector<double> batch_rot;   // n x 9 row-major

    // src/pybind.cpp:216-247: positions (n,3), rotations (3,3,M) with M >= n
    // (the reference pads to 100000), element R_k[i][j] = rotations[i, j, k].
    void batch_set_pose(darr positions, darr rotations, int n) {
        if (n < 0) throw py::value_error("number_of_poses must be >= 0");
        if ((size_t)positions.size() < (size_t)n * 3) throw py::value_error("batch_positions has fewer than n poses");
        const size_t m = rotations.size() / 9;
        if (rotations.ndim() != 3 || rotations.shape(0) != 3 || rotations.shape(1) != 3 || m < (size_t)n)
            throw py::value_error("batch_rotations must have shape (3, 3, M) with M >= number_of_poses");
        batch_pos.assign(positions.data(), positions.data() + 3 * (size_t)n);
        batch_rot.resize(9 * (size_t)n);
        const double* r = rotations.data();
        for (int k = 0; k < n; ++k)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) batch_rot[9 * (size_t)k + 3 * i + j] = r[(size_t)i * 3 * m + (size_t)j * m + k];
    }
};

struct PointLight : Pose {
    double intensity;
};
struct AreaLight : Pose {
    double intensity, size[2];
};
struct DirectionalLight : Pose {
    double intensity, direction[3];
};

using u8arr = py::array_t<uint8_t, py::array::c_style | py::array::forcecast>;

// RGBA8 copy of an (H, W, 4) or (H, W, 3) uint8 image (alpha 255 added).
std::shared_ptr<std::vector<uint8_t>> rgba_copy(u8arr img, uint32_t& w, uint32_t& h) {
    if (img.ndim() != 3 || (img.shape(2) != 3 && img.shape(2) != 4) || img.shape(0) < 1 || img.shape(1) < 1)
        throw py::value_error("texture must be an (H, W, 3) or (H, W, 4) uint8 array");
    h = (uint32_t)img.shape(0);
    w = (uint32_t)img.shape(1);
    const size_t c = (size_t)img.shape(2);
    auto out = std::make_shared<std::vector<uint8_t>>((size_t)w * h * 4, 255);
    const uint8_t* p = img.data();
    for (size_t i = 0; i < (size_t)w * h; ++i)
        for (size_t k = 0; k < c; ++k) (*out)[4 * i + k] = p[c * i + k];
    return out;
}

struct Texture {
    uint32_t w = 0, h = 0;
    std::shared_ptr<std::vector<uint8_t>> rgba;
};

struct Entity : Pose {
    std::string geometry_path, geometry_type;
    bool smooth_shading;
    float color[3];
    double scale = 1.0;
    uint32_t id = 0;
    std::shared_ptr<crt_host::Mesh> mesh;   // loaded at construction (entity.hpp:31-67)
    // material (additive: the reference hard-codes ColoredLambertian,
    // src/crt/rendering_dynamic/entity.hpp:51-53) and per-vertex UVs
    uint32_t mat_type = CRT_MAT_LAMBERTIAN;
    float mat_alpha = 24.0f;
    float mat_coeffs[3] = {0.f, 1.f, 0.f};
    Texture tex, spec_tex;
    std::vector<float> uv;

    // kind: "lambertian" (the entity colour), "textured_lambertian",
    // "blinn_phong", "mirror"; texture: image or None (the entity colour);
    // specular: (ka, kd, ks) or an image whose R, G, B hold them.
    void set_material(const std::string& kind, py::object texture, py::object specular, double alpha) {
        Texture t, st;
        float co[3] = {0.f, 1.f, 0.f};
        uint32_t type;
        if (kind == "lambertian") type = CRT_MAT_LAMBERTIAN;
        else if (kind == "textured_lambertian") type = CRT_MAT_TEX_LAMBERTIAN;
        else if (kind == "blinn_phong") type = CRT_MAT_BLINN_PHONG;
        else if (kind == "mirror") type = CRT_MAT_MIRROR;
        else throw py::value_error("unknown material '" + kind + "' (lambertian, textured_lambertian, blinn_phong, mirror)");
        if (!texture.is_none()) t.rgba = rgba_copy(texture.cast<u8arr>(), t.w, t.h);
        if (!specular.is_none()) {
            py::array sa = py::array::ensure(specular);
            if (sa && sa.ndim() == 3) {
                st.rgba = rgba_copy(specular.cast<u8arr>(), st.w, st.h);
            } else {
                auto v = specular.cast<std::vector<double>>();
                if (v.size() != 3) throw py::value_error("specular needs (ka, kd, ks) or an image");
                for (int k = 0; k < 3; ++k) co[k] = (float)v[k];
            }
        }
        if (!std::isfinite(alpha)) throw py::value_error("alpha must be finite");
        mat_type = type;
        mat_alpha = (float)alpha;
        for (int k = 0; k < 3; ++k) mat_coeffs[k] = co[k];
        tex = t;
        spec_tex = st;
    }
    void set_uv(darr uvs) {
        if (mesh && (size_t)uvs.size() != 2 * mesh->n_vertices())
            throw py::value_error("uv needs (n_vertices, 2) values");
        uv.assign(uvs.data(), uvs.data() + uvs.size());
    }
    bool default_material() const { return mat_type == CRT_MAT_LAMBERTIAN && !tex.rgba && uv.empty(); }

    void load() {
        std::string t = geometry_type;
        std::transform(t.begin(), t.end(), t.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        if (t != "obj")
            throw py::value_error("file type of " + geometry_type + " is not valid: crt supports obj");
        try {
            mesh = std::make_shared<crt_host::Mesh>(crt_host::load_obj(geometry_path));
        } catch (const std::exception& e) {
            throw py::value_error(e.what());
        }
    }
    void set_scale(double s) { scale = s; }
};

Entity make_entity(const std::string& path, const std::string& type, bool smooth, py::list color, bool load) {
    Entity e;
    e.geometry_path = path;
    e.geometry_type = type;
    e.smooth_shading = smooth;
    for (int k = 0; k < 3; ++k) e.color[k] = color[k].cast<float>();
    if (load) e.load();
    return e;
}

struct BodyFixedEntity : Entity {};

// Merge entities into one GPU scene.  Per entity, FP64 scale -> rotate (R)
// -> translate of every vertex (src/crt/transform.hpp:42-85), vertex normals
// rotated only; ids 1..n in list order (src/pybind.cpp:107-127).
struct GpuScene {
    crt_scene* h = nullptr;
    std::vector<float> colors;
    GpuScene() = default;
    GpuScene(const GpuScene&) = delete;
    GpuScene& operator=(const GpuScene&) = delete;
    ~GpuScene() {
        if (h) crt_scene_destroy(h);
    }
};

// Entities merged in FP64 (vertices stay FP64 up to the library, which
// rounds them to FP32 only after moving them into the scene frame: an entity
// 1e7 m out keeps its own resolution, as the reference's double geometry does).
struct SceneInput {
    std::vector<double> V;
    std::vector<float> VN, col;
    std::vector<uint32_t> I, inst;
    std::vector<uint8_t> sm;
};

SceneInput gather(const std::vector<const Entity*>& ents) {
    SceneInput in;
    uint32_t base = 0;
    for (size_t e = 0; e < ents.size(); ++e) {
        const Entity& en = *ents[e];
        if (!en.mesh) throw py::value_error("entity has no geometry loaded");
        const crt_host::Mesh& m = *en.mesh;
        const double* R = en.rotation;
        for (size_t i = 0; i < m.n_vertices(); ++i) {
            const double x = en.scale * m.v[3 * i], y = en.scale * m.v[3 * i + 1], z = en.scale * m.v[3 * i + 2];
            in.V.push_back(R[0] * x + R[1] * y + R[2] * z + en.position[0]);
            in.V.push_back(R[3] * x + R[4] * y + R[5] * z + en.position[1]);
            in.V.push_back(R[6] * x + R[7] * y + R[8] * z + en.position[2]);
            const double a = m.vn[3 * i], b = m.vn[3 * i + 1], c = m.vn[3 * i + 2];
            in.VN.push_back((float)(R[0] * a + R[1] * b + R[2] * c));
            in.VN.push_back((float)(R[3] * a + R[4] * b + R[5] * c));
            in.VN.push_back((float)(R[6] * a + R[7] * b + R[8] * c));
        }
        for (size_t t = 0; t < m.f.size(); ++t) in.I.push_back(m.f[t] + base);
        in.inst.insert(in.inst.end(), m.n_tris(), (uint32_t)(e + 1));
        base += (uint32_t)m.n_vertices();
        in.col.insert(in.col.end(), en.color, en.color + 3);
        in.sm.push_back(en.smooth_shading ? 1 : 0);
    }
    if (in.I.empty()) throw py::value_error("the entities contain no triangles");
    return in;
}

void upload_materials(crt_scene* h, const std::vector<const Entity*>& ents) {
    bool any = false, any_uv = false;
    for (const Entity* e : ents) {
        any = any || !e->default_material();
        any_uv = any_uv || !e->uv.empty();
    }
    if (!any) return;
    std::vector<crt_material> mats;
    std::vector<crt_texture> texs;
    std::vector<float> uvs;
    for (const Entity* e : ents) {
        crt_material m{};
        m.type = e->mat_type;
        m.alpha = e->mat_alpha;
        for (int k = 0; k < 3; ++k) { m.color[k] = e->color[k]; m.coeffs[k] = e->mat_coeffs[k]; }
        m.texture = m.spec_texture = -1;
        if (e->tex.rgba) {
            m.texture = (int32_t)texs.size();
            texs.push_back({e->tex.w, e->tex.h, e->tex.rgba->data()});
        }
        if (e->spec_tex.rgba) {
            m.spec_texture = (int32_t)texs.size();
            texs.push_back({e->spec_tex.w, e->spec_tex.h, e->spec_tex.rgba->data()});
        }
        mats.push_back(m);
        if (any_uv) {
            if (!e->uv.empty() && e->uv.size() != 2 * e->mesh->n_vertices())
                throw py::value_error("entity uv needs (n_vertices, 2) values");
            if (e->uv.empty()) uvs.insert(uvs.end(), 2 * e->mesh->n_vertices(), 0.0f);
            else uvs.insert(uvs.end(), e->uv.begin(), e->uv.end());
        }
    }
    py::gil_scoped_release nogil;
    check(crt_scene_set_materials(h, (uint32_t)mats.size(), mats.data(), any_uv ? uvs.data() : nullptr,
                                  (uint32_t)texs.size(), texs.empty() ? nullptr : texs.data()));
}

std::shared_ptr<GpuScene> build_scene(const std::vector<const Entity*>& ents, uint32_t max_leaf) {
    ensure_device();
    const SceneInput in = gather(ents);
    auto g = std::make_shared<GpuScene>();
    g->colors = in.col;
    crt_build_opts opts{max_leaf, 0};
    crt_scene* h = nullptr;
    {
        py::gil_scoped_release nogil;
        check(crt_scene_create_f64(in.V.data(), (uint32_t)(in.V.size() / 3), in.I.data(), (uint32_t)(in.I.size() / 3),
                                   in.inst.data(), in.VN.data(), (uint32_t)ents.size(), in.col.data(), in.sm.data(),
                                   &opts, &h));
    }
    g->h = h;
    upload_materials(h, ents);
    return g;
}

// Dynamic API (src/crt/rendering_dynamic/render.hpp:19-62 rebuilds the BVH
// on every call): the last per-call scene is kept, and a call with the same
// meshes, colours, shading flags and materials -- only poses or scales
// changed -- refits it on the GPU (crt_scene_refit, which rebuilds by itself
// when the SAH has grown 1.5x).  A call checks the scene out of the cache
// while it uses it, so concurrent calls never share one.
struct DynKey {
    std::vector<const void*> meshes, textures;
    std::vector<float> params;    // colour, smooth flag, material type / alpha / coefficients per entity
    std::vector<float> uv;
    bool operator==(const DynKey& o) const {
        return meshes == o.meshes && textures == o.textures && params == o.params && uv == o.uv;
    }
};

DynKey dyn_key(const std::vector<const Entity*>& ents) {
    DynKey k;
    for (const Entity* e : ents) {
        k.meshes.push_back(e->mesh.get());
        k.textures.push_back(e->tex.rgba.get());
        k.textures.push_back(e->spec_tex.rgba.get());
        k.params.insert(k.params.end(), {e->color[0], e->color[1], e->color[2], e->smooth_shading ? 1.f : 0.f,
                                         (float)e->mat_type, e->mat_alpha, e->mat_coeffs[0], e->mat_coeffs[1],
                                         e->mat_coeffs[2]});
        k.uv.insert(k.uv.end(), e->uv.begin(), e->uv.end());
        k.uv.push_back(-1.f);   // separator: which entity the UVs belong to
    }
    return k;
}

std::mutex g_dyn_mu;
DynKey g_dyn_key;
std::shared_ptr<GpuScene> g_dyn_scene;
std::atomic<uint64_t> g_dyn_refits{0}, g_dyn_builds{0};

struct DynScene {
    std::shared_ptr<GpuScene> scene;
    DynKey key;
    ~DynScene() {   // back into the cache for the next call
        if (!scene) return;
        std::lock_guard<std::mutex> lk(g_dyn_mu);
        g_dyn_scene = scene;
        g_dyn_key = std::move(key);
    }
};

// ---------------------------------------------------------------- lights
std::vector<crt_light> convert_lights(py::list lights) {
    std::vector<crt_light> out;
    for (py::handle l : lights) {
        crt_light c{};
        if (py::isinstance<PointLight>(l)) {
            const auto& p = l.cast<const PointLight&>();
            c.type = CRT_LIGHT_POINT;
            c.intensity = p.intensity;
            std::memcpy(c.position, p.position, sizeof(c.position));
            std::memcpy(c.rotation, p.rotation, sizeof(c.rotation));
        } else if (py::isinstance<AreaLight>(l)) {
            const auto& a = l.cast<const AreaLight&>();
            c.type = CRT_LIGHT_AREA;
            c.intensity = a.intensity;
            std::memcpy(c.position, a.position, sizeof(c.position));
            std::memcpy(c.rotation, a.rotation, sizeof(c.rotation));
            c.size[0] = a.size[0];
            c.size[1] = a.size[1];
        } else if (py::isinstance<DirectionalLight>(l)) {
            const auto& d = l.cast<const DirectionalLight&>();
            c.type = CRT_LIGHT_DIRECTIONAL;
            c.intensity = d.intensity;
            std::memcpy(c.direction, d.direction, sizeof(c.direction));
        } else {
            throw py::type_error("unsupported light type (PointLight, AreaLight or DirectionalLight expected)");
        }
        out.push_back(c);
    }
    return out;
}

const SimpleCamera& as_camera(py::handle h) {
    if (!py::isinstance<SimpleCamera>(h)) throw py::type_error("unsupported camera model (SimpleCamera expected)");
    return h.cast<const SimpleCamera&>();
}
const SimpleLidar& as_lidar(py::handle h) {
    if (!py::isinstance<SimpleLidar>(h)) throw py::type_error("unsupported lidar model (SimpleLidar expected)");
    return h.cast<const SimpleLidar&>();
}

uint64_t fresh_seed() {
    // the reference draws its render seeds from std::random_device
    // (src/crt/do_render.hpp:49-50); so does the default here
    static std::random_device rd;
    return ((uint64_t)rd() << 32) ^ rd();
}

// ---------------------------------------------------------------- operations
py::array_t<uint8_t> op_render(crt_scene* s, const SimpleCamera& cam, py::list lights, int min_s, int max_s,
                               double thr, int nb, py::object seed, py::array_t<float>* radiance) {
    const crt_camera c = cam.to_c();
    const auto L = convert_lights(lights);
    crt_render_opts o{};
    o.min_samples = min_s;
    o.max_samples = max_s;
    o.noise_threshold = thr;
    o.num_bounces = nb;
    o.seed = seed.is_none() ? fresh_seed() : seed.cast<uint64_t>();
    const py::ssize_t h = cam.height(), w = cam.width();
    if (h < 1 || w < 1) throw py::value_error("camera resolution must be >= 1");
    py::array_t<uint8_t> img({h, w, (py::ssize_t)4});
    float* rad = nullptr;
    std::vector<float> tmp;
    if (radiance) {
        *radiance = py::array_t<float>({h, w, (py::ssize_t)4});
        rad = radiance->mutable_data();
    }
    uint8_t* out = img.mutable_data();
    {
        py::gil_scoped_release nogil;
        check(crt_render(s, &c, L.data(), (uint32_t)L.size(), &o, rad, out, nullptr, 0, nullptr, nullptr));
    }
    return img;
}

py::array_t<double> op_pass_xyz(crt_scene* s, const SimpleCamera& cam, bool normals) {
    const crt_camera c = cam.to_c();
    const py::ssize_t h = cam.height(), w = cam.width();
    py::array_t<double> out({h, w, (py::ssize_t)3});
    double* p = out.mutable_data();
    {
        py::gil_scoped_release nogil;
        check(crt_pass(s, &c, normals ? nullptr : p, normals ? p : nullptr, nullptr, nullptr, nullptr, 0, nullptr,
                       nullptr));
    }
    return out;
}

py::array_t<uint32_t> op_pass_instance(crt_scene* s, const SimpleCamera& cam) {
    const crt_camera c = cam.to_c();
    py::array_t<uint32_t> out({(py::ssize_t)cam.height(), (py::ssize_t)cam.width()});
    uint32_t* p = out.mutable_data();
    {
        py::gil_scoped_release nogil;
        check(crt_pass(s, &c, nullptr, nullptr, p, nullptr, nullptr, 0, nullptr, nullptr));
    }
    return out;
}

double op_lidar(crt_scene* s, const SimpleLidar& l) {
    // src/crt/do_lidar.hpp:13-77: one ray, mean of its range (0 on a miss)
    double r = 0.0;
    {
        py::gil_scoped_release nogil;
        check(crt_lidar_batch(s, 1, l.position, l.rotation, l.z_positive ? 1 : 0, &r, 0, nullptr, nullptr));
    }
    return r;
}

py::array_t<double> op_batch_lidar(crt_scene* s, const SimpleLidar& l) {
    const py::ssize_t n = (py::ssize_t)(l.batch_pos.size() / 3);
    py::array_t<double> out({(py::ssize_t)1, n});
    double* p = out.mutable_data();
    if (n > 0) {
        py::gil_scoped_release nogil;
        check(crt_lidar_batch(s, (uint64_t)n, l.batch_pos.data(), l.batch_rot.data(), l.z_positive ? 1 : 0, p, 0,
                              nullptr, nullptr));
    }
    return out;
}

struct BodyFixedGroup : Pose {
    double scale = 1.0;
    std::shared_ptr<GpuScene> scene;
    uint32_t n_entities = 0;
    void set_scale(double s) { scale = s; }
};

BodyFixedGroup make_group(py::list entities, uint32_t max_leaf) {
    std::vector<Entity> loaded;
    loaded.reserve(py::len(entities));
    for (py::handle h : entities) {
        if (!py::isinstance<BodyFixedEntity>(h) && !py::isinstance<Entity>(h))
            throw py::type_error("BodyFixedGroup expects BodyFixedEntity objects");
        Entity e = h.cast<Entity>();
        e.load();
        loaded.push_back(std::move(e));
    }
    std::vector<const Entity*> ptrs;
    for (auto& e : loaded) ptrs.push_back(&e);
    BodyFixedGroup g;
    g.scene = build_scene(ptrs, max_leaf);
    g.n_entities = (uint32_t)ptrs.size();
    return g;
}

std::unique_ptr<DynScene> dynamic_scene(py::list entities) {
    std::vector<const Entity*> ptrs;
    for (py::handle h : entities) {
        if (!py::isinstance<Entity>(h)) throw py::type_error("expected Entity objects");
        ptrs.push_back(&h.cast<const Entity&>());
    }
    ensure_device();
    auto d = std::make_unique<DynScene>();
    d->key = dyn_key(ptrs);
    {
        std::lock_guard<std::mutex> lk(g_dyn_mu);
        if (g_dyn_scene && g_dyn_key == d->key) d->scene = std::move(g_dyn_scene);
        g_dyn_scene.reset();
    }
    if (d->scene) {
        const SceneInput in = gather(ptrs);
        uint32_t rebuilt = 0;
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = crt_scene_refit(d->scene->h, in.V.data(), in.VN.data(), 0.0, &rebuilt);
        }
        if (rc != CRT_OK) d->scene.reset();   // a half-refitted scene never goes back to the cache
        check(rc);
        ++g_dyn_refits;
    } else {
        d->scene = build_scene(ptrs, 8);
        ++g_dyn_builds;
    }
    return d;
}

template <typename T>
void bind_pose(py::class_<T>& c) {
    c.def("set_position", &T::set_position)
        .def("set_rotation", &T::set_rotation)
        .def("set_pose", &T::set_pose)
        .def_property_readonly("position", [](const T& s) {
            return py::array_t<double>(3, s.position);
        })
        .def_property_readonly("rotation", [](const T& s) {
            py::array_t<double> r({3, 3});
            std::memcpy(r.mutable_data(), s.rotation, sizeof(s.rotation));
            return r;
        });
}

}  // namespace

PYBIND11_MODULE(_crt, m) {
    m.doc() = "ceres ray tracer -- MI355X HIP core";

    auto cam = py::class_<SimpleCamera>(m, "SimpleCamera").def(py::init(&make_camera));
    bind_pose(cam);
    cam.def_readonly("focal_length", &SimpleCamera::focal_length).def_readonly("z_positive", &SimpleCamera::z_positive);

    auto lid = py::class_<SimpleLidar>(m, "SimpleLidar").def(py::init([](bool z) {
        SimpleLidar l;
        l.z_positive = z;
        return l;
    }));
    bind_pose(lid);
    lid.def("batch_set_pose", &SimpleLidar::batch_set_pose);

    auto pl = py::class_<PointLight>(m, "PointLight").def(py::init([](double i) {
        PointLight l;
        l.intensity = i;
        return l;
    }));
    bind_pose(pl);
    auto al = py::class_<AreaLight>(m, "AreaLight").def(py::init([](double i, py::list size) {
        AreaLight l;
        l.intensity = i;
        l.size[0] = size[0].cast<double>();
        l.size[1] = size[1].cast<double>();
        return l;
    }));
    bind_pose(al);
    auto dl = py::class_<DirectionalLight>(m, "DirectionalLight").def(py::init([](double i, darr dir) {
        DirectionalLight l;
        l.intensity = i;
        if (dir.size() < 3) throw py::value_error("direction needs 3 values");
        for (int k = 0; k < 3; ++k) l.direction[k] = dir.data()[k];
        return l;
    }));
    bind_pose(dl);

    auto ent = py::class_<Entity>(m, "Entity").def(
        py::init([](const std::string& p, const std::string& t, bool s, py::list c) {
            return make_entity(p, t, s, c, true);
        }));
    bind_pose(ent);
    ent.def("set_scale", &Entity::set_scale).def_readonly("scale", &Entity::scale);
    ent.def("set_material", &Entity::set_material, py::arg("kind"), py::arg("texture") = py::none(),
            py::arg("specular") = py::none(), py::arg("alpha") = 24.0)
        .def("set_uv", &Entity::set_uv);

    auto bfe = py::class_<BodyFixedEntity, Entity>(m, "BodyFixedEntity").def(
        py::init([](const std::string& p, const std::string& t, bool s, py::list c) {
            BodyFixedEntity e;
            static_cast<Entity&>(e) = make_entity(p, t, s, c, false);
            return e;
        }));
    bfe.def("set_scale", &BodyFixedEntity::set_scale);

    auto grp = py::class_<BodyFixedGroup>(m, "BodyFixedGroup")
                   .def(py::init([](py::list ents, uint32_t max_leaf) { return make_group(ents, max_leaf); }),
                        py::arg("entities"), py::arg("max_leaf_size") = 8u);
    bind_pose(grp);
    grp.def("set_scale", &BodyFixedGroup::set_scale)
        .def("render",
             [](BodyFixedGroup& g, py::handle camera, py::list lights, int min_s, int max_s, double thr, int nb,
                py::object seed) {
                 return op_render(g.scene->h, as_camera(camera), lights, min_s, max_s, thr, nb, seed, nullptr);
             },
             py::arg("camera"), py::arg("lights"), py::arg("min_samples"), py::arg("max_samples"),
             py::arg("noise_threshold"), py::arg("num_bounces"), py::arg("seed") = py::none())
        .def("render_radiance",
             [](BodyFixedGroup& g, py::handle camera, py::list lights, int min_s, int max_s, double thr, int nb,
                py::object seed) {
                 py::array_t<float> rad;
                 op_render(g.scene->h, as_camera(camera), lights, min_s, max_s, thr, nb, seed, &rad);
                 return rad;
             },
             py::arg("camera"), py::arg("lights"), py::arg("min_samples") = 1, py::arg("max_samples") = 1,
             py::arg("noise_threshold") = 1.0, py::arg("num_bounces") = 1, py::arg("seed") = py::none())
        .def("simulate_lidar", [](BodyFixedGroup& g, py::handle lidar, double) { return op_lidar(g.scene->h, as_lidar(lidar)); })
        .def("batch_simulate_lidar",
             [](BodyFixedGroup& g, py::handle lidar, double) { return op_batch_lidar(g.scene->h, as_lidar(lidar)); })
        .def("intersection_pass",
             [](BodyFixedGroup& g, py::handle camera) { return op_pass_xyz(g.scene->h, as_camera(camera), false); })
        .def("instance_pass",
             [](BodyFixedGroup& g, py::handle camera) { return op_pass_instance(g.scene->h, as_camera(camera)); })
        .def("normal_pass",
             [](BodyFixedGroup& g, py::handle camera) { return op_pass_xyz(g.scene->h, as_camera(camera), true); })
        .def("depth_pass",
             [](BodyFixedGroup& g, py::handle camera) {
                 const SimpleCamera& c = as_camera(camera);
                 const crt_camera cc = c.to_c();
                 py::array_t<double> out({(py::ssize_t)c.height(), (py::ssize_t)c.width()});
                 double* p = out.mutable_data();
                 py::gil_scoped_release nogil;
                 check(crt_pass(g.scene->h, &cc, nullptr, nullptr, nullptr, p, nullptr, 0, nullptr, nullptr));
                 return out;
             })
        .def("triangle_pass",
             [](BodyFixedGroup& g, py::handle camera) {
                 const SimpleCamera& c = as_camera(camera);
                 const crt_camera cc = c.to_c();
                 py::array_t<int32_t> out({(py::ssize_t)c.height(), (py::ssize_t)c.width()});
                 int32_t* p = out.mutable_data();
                 py::gil_scoped_release nogil;
                 check(crt_pass(g.scene->h, &cc, nullptr, nullptr, nullptr, nullptr, p, 0, nullptr, nullptr));
                 return out;
             })
        .def("sun_pass",
             [](BodyFixedGroup& g, py::handle camera, darr sun) {
                 const SimpleCamera& c = as_camera(camera);
                 const crt_camera cc = c.to_c();
                 if (sun.size() < 3) throw py::value_error("sun_dir needs 3 values");
                 const double s[3] = {sun.data()[0], sun.data()[1], sun.data()[2]};
                 const py::ssize_t h = c.height(), w = c.width();
                 py::array_t<double> depth({h, w});
                 py::array_t<uint8_t> lit({h, w});
                 double* dp = depth.mutable_data();
                 uint8_t* lp = lit.mutable_data();
                 {
                     py::gil_scoped_release nogil;
                     check(crt_sun_pass(g.scene->h, &cc, s, dp, nullptr, lp, 0, nullptr, nullptr));
                 }
                 return py::make_tuple(depth, lit);
             })
        .def("flash_lidar",
             [](BodyFixedGroup& g, py::handle beams, uint32_t sub_rays, double sigma, uint64_t seed) {
                 const SimpleCamera& c = as_camera(beams);
                 crt_flash_opts o{};
                 o.beams = c.to_c();
                 o.sub_rays = sub_rays;
                 o.sigma = sigma;
                 o.seed = seed;
                 const py::ssize_t h = c.height(), w = c.width();
                 py::array_t<double> single({h, w}), first({h, w}), last({h, w});
                 py::array_t<uint8_t> nret({h, w});
                 double *a = single.mutable_data(), *b = first.mutable_data(), *d = last.mutable_data();
                 uint8_t* n = nret.mutable_data();
                 {
                     py::gil_scoped_release nogil;
                     check(crt_flash_lidar(g.scene->h, &o, a, b, d, n, 0, nullptr, nullptr));
                 }
                 py::dict r;
                 r["single"] = single;
                 r["first"] = first;
                 r["last"] = last;
                 r["n_returns"] = nret;
                 return r;
             },
             py::arg("beams"), py::arg("sub_rays") = 8u, py::arg("sigma") = 5e-4, py::arg("seed") = 0u)
        .def("radiation_pressure",
             [](BodyFixedGroup& g, darr sun, double pressure, double rho_s, double rho_d, uint32_t grid, darr ref) {
                 crt_flux_opts o{};
                 if (sun.size() < 3 || ref.size() < 3) throw py::value_error("sun_dir / ref_point need 3 values");
                 for (int k = 0; k < 3; ++k) {
                     o.sun_dir[k] = sun.data()[k];
                     o.ref_point[k] = ref.data()[k];
                 }
                 o.pressure = pressure;
                 o.rho_s = rho_s;
                 o.rho_d = rho_d;
                 o.grid = grid;
                 o.radius = 0.0;
                 double f[3], t[3];
                 uint64_t hits = 0;
                 {
                     py::gil_scoped_release nogil;
                     check(crt_radiation_pressure(g.scene->h, &o, f, t, &hits, nullptr, nullptr));
                 }
                 return py::make_tuple(py::array_t<double>(3, f), py::array_t<double>(3, t), hits);
             },
             py::arg("sun_dir"), py::arg("pressure"), py::arg("rho_s"), py::arg("rho_d"), py::arg("grid") = 2048u,
             py::arg("ref_point") = std::vector<double>{0.0, 0.0, 0.0})
        .def("bvh_info", [](BodyFixedGroup& g) {
            crt_bvh_info i{};
            check(crt_scene_info(g.scene->h, &i));
            py::dict d;
            d["n_tris"] = i.n_tris;
            d["n_wide_nodes"] = i.n_wide_nodes;
            d["n_leaves"] = i.n_leaves;
            d["max_depth"] = i.max_depth;
            d["build_ms"] = i.build_ms;
            d["sah_cost"] = i.sah_cost;
            d["device_bytes"] = i.device_bytes;
            return d;
        });

    m.def("load_obj",
          [](const std::string& path) {
              crt_host::Mesh mm;
              try {
                  mm = crt_host::load_obj(path);
              } catch (const std::exception& e) {
                  throw py::value_error(e.what());
              }
              py::array_t<double> v({(py::ssize_t)mm.n_vertices(), (py::ssize_t)3});
              py::array_t<uint32_t> f({(py::ssize_t)mm.n_tris(), (py::ssize_t)3});
              py::array_t<double> vn({(py::ssize_t)mm.n_vertices(), (py::ssize_t)3});
              std::copy(mm.v.begin(), mm.v.end(), v.mutable_data());
              std::copy(mm.f.begin(), mm.f.end(), f.mutable_data());
              std::copy(mm.vn.begin(), mm.vn.end(), vn.mutable_data());
              return py::make_tuple(v, f, vn);
          },
          "Parse an OBJ file like the reference loader: (vertices, triangles, vertex normals).");

    // ---- dynamic (per-call BVH) module functions, src/pybind.cpp:580-718
    m.def("render", [](py::handle camera, py::list lights, py::list entities, int min_s, int max_s, double thr,
                       int nb) {
        auto d = dynamic_scene(entities);
        return op_render(d->scene->h, as_camera(camera), lights, min_s, max_s, thr, nb, py::none(), nullptr);
    });
    m.def("simulate_lidar", [](py::handle lidar, py::list entities, int) {
        auto d = dynamic_scene(entities);
        return op_lidar(d->scene->h, as_lidar(lidar));
    });
    m.def("intersection_pass", [](py::handle camera, py::list entities) {
        auto d = dynamic_scene(entities);
        return op_pass_xyz(d->scene->h, as_camera(camera), false);
    });
    m.def("instance_pass", [](py::handle camera, py::list entities) {
        auto d = dynamic_scene(entities);
        return op_pass_instance(d->scene->h, as_camera(camera));
    });
    m.def("normal_pass", [](py::handle camera, py::list entities) {
        auto d = dynamic_scene(entities);
        return op_pass_xyz(d->scene->h, as_camera(camera), true);
    });
    m.def("dynamic_scene_stats", []() {
        std::lock_guard<std::mutex> lk(g_dyn_mu);
        py::dict r;
        r["refits"] = g_dyn_refits.load();
        r["builds"] = g_dyn_builds.load();
        return r;
    }, "Counts of per-call scenes built and refitted by the dynamic API (additive).");
}
