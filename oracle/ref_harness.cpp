// ref_harness.cpp -- C entry points around the reference CPU tracer.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile.ref against the
// ceres-raytracer headers where they lie (nothing is copied) into
// oracle/_ref/libcrtref.so.  It drives the reference's own classes --
// Entity + BodyFixedGroup (SweepSAH + PRO + NLO BVH), SimpleCamera,
// SingleRayTraverser with Statistics, unidirectional(), do_render,
// do_lidar -- so that golden vectors, traversal statistics and the CPU
// baseline come from the reference itself.
//
// <iostream> must come first: src/crt/materials/material.hpp:47,84 use
// std::cout without including it.
#include <iostream>
#include <sstream>
#include <cstdint>
#include <memory>
#include <random>
#include <vector>
#include <omp.h>

// same include order as src/pybind.cpp:8-25 (several headers rely on it)
#include "crt/cameras/camera.hpp"
#include "crt/cameras/simple_camera.hpp"
#include "crt/lidars/lidar.hpp"
#include "crt/lidars/simple_lidar.hpp"
#include "crt/lights/light.hpp"
#include "crt/lights/point_light.hpp"
#include "crt/lights/area_light.hpp"
#include "crt/rendering_body_fixed/body_fixed_entity.hpp"
#include "crt/rendering_body_fixed/body_fixed_group.hpp"
#include "crt/rendering_dynamic/entity.hpp"
#include "crt/passes.hpp"

using Scalar = double;
using Group = BodyFixedGroup<Scalar>;

namespace {

std::stringstream g_sink;
std::streambuf* g_saved = nullptr;

struct CamParams {   // flat layout shared with oracle/ref.py
    double pos[3], rot[9], f, res[2], sensor[2], zpos;
};

std::unique_ptr<Camera<Scalar>> make_camera(const double* p) {
    const CamParams* c = reinterpret_cast<const CamParams*>(p);
    Scalar res[2] = {c->res[0], c->res[1]};
    Scalar sen[2] = {c->sensor[0], c->sensor[1]};
    auto cam = std::make_unique<SimpleCamera<Scalar>>(c->f, res, sen, c->zpos != 0.0);
    Scalar r[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[i][j] = c->rot[3 * i + j];
    cam->set_pose(bvh::Vector3<Scalar>(c->pos[0], c->pos[1], c->pos[2]), r);
    return cam;
}

struct LightParams {  // type(0 point,1 area), intensity, pos[3], rot[9], size[2]
    double type, intensity, pos[3], rot[9], size[2];
};

std::vector<std::unique_ptr<Light<Scalar>>> make_lights(int nl, const double* p) {
    std::vector<std::unique_ptr<Light<Scalar>>> out;
    for (int k = 0; k < nl; ++k) {
        const LightParams* L = reinterpret_cast<const LightParams*>(p) + k;
        Scalar r[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r[i][j] = L->rot[3 * i + j];
        const bvh::Vector3<Scalar> pos(L->pos[0], L->pos[1], L->pos[2]);
        if (L->type == 0.0) {
            auto l = std::make_unique<PointLight<Scalar>>(L->intensity);
            l->set_pose(pos, r);
            out.push_back(std::move(l));
        } else {
            Scalar sz[2] = {L->size[0], L->size[1]};
            auto l = std::make_unique<AreaLight<Scalar>>(L->intensity, sz);
            l->set_pose(pos, r);
            out.push_back(std::move(l));
        }
    }
    return out;
}

}  // namespace

extern "C" {

void ref_quiet(int on) {
    if (on && !g_saved) g_saved = std::cout.rdbuf(g_sink.rdbuf());
    if (!on && g_saved) { std::cout.rdbuf(g_saved); g_saved = nullptr; }
    g_sink.str("");
}

int ref_num_threads(void) { return omp_get_max_threads(); }

// Build a BodyFixedGroup from OBJ files exactly like create_body_fixed_group
// (src/pybind.cpp:107-127): Entity per file, scale/pose, ids 1..n in order.
void* ref_group_create(int n, const char** paths, const double* pos, const double* rot, const double* scale,
                       const float* color, const int* smooth, int build_threads) {
    std::vector<Entity<Scalar>*> ents;
    for (int k = 0; k < n; ++k) {
        Color c(color[3 * k], color[3 * k + 1], color[3 * k + 2]);
        auto* e = new Entity<Scalar>(paths[k], "obj", smooth[k] != 0, c);
        e->set_scale(scale[k]);
        e->set_position(bvh::Vector3<Scalar>(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]));
        Scalar r[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r[i][j] = rot[9 * k + 3 * i + j];
        e->set_rotation(r);
        e->set_id((uint32_t)(k + 1));
        ents.push_back(e);
    }
    const int saved = omp_get_max_threads();
    if (build_threads > 0) omp_set_num_threads(build_threads);
    auto* g = new Group(ents);
    omp_set_num_threads(saved);
    return g;   // entities are kept alive: triangles hold raw parent pointers
}

void ref_group_destroy(void* g) { delete static_cast<Group*>(g); }

int64_t ref_group_ntris(void* g) { return (int64_t) static_cast<Group*>(g)->triangles.size(); }

// Triangle data in group order: p0,p1,p2 (9 doubles), instance id, vertex normals (9 doubles).
void ref_group_tris(void* gp, double* p, uint32_t* inst, double* vn) {
    Group* g = static_cast<Group*>(gp);
    for (size_t t = 0; t < g->triangles.size(); ++t) {
        const auto& tr = g->triangles[t];
        const auto a = tr.p0, b = tr.p1(), c = tr.p2();
        for (int k = 0; k < 3; ++k) { p[9 * t + k] = a[k]; p[9 * t + 3 + k] = b[k]; p[9 * t + 6 + k] = c[k]; }
        if (inst) inst[t] = tr.parent ? tr.parent->id : 0u;
        if (vn)
            for (int k = 0; k < 3; ++k) {
                vn[9 * t + k] = tr.vn0[k]; vn[9 * t + 3 + k] = tr.vn1[k]; vn[9 * t + 6 + k] = tr.vn2[k];
            }
    }
}

// Node count and SAH cost (formula of lib/bvh/sah_based_algorithm.hpp:21-32,
// traversal_cost = 1) of the reference BVH.
void ref_group_bvh_info(void* gp, int64_t* node_count, double* sah_cost) {
    Group* g = static_cast<Group*>(gp);
    auto& bvh = g->bvh_cache;
    double cost = 0.0;
    for (size_t i = 0; i < bvh.node_count; ++i) {
        const auto bb = bvh.nodes[i].bounding_box_proxy().to_bounding_box();
        const double ha = bb.half_area();
        cost += bvh.nodes[i].is_leaf() ? ha * (double)bvh.nodes[i].primitive_count : ha;
    }
    *node_count = (int64_t)bvh.node_count;
    *sah_cost = cost / bvh.nodes[0].bounding_box_proxy().to_bounding_box().half_area();
}

// Rays through the reference traverser with statistics
// (lib/bvh/single_ray_traverser.hpp:132-163).  prim = triangle index in
// group order (-1 on miss).
void ref_trace(void* gp, int64_t n, const double* o, const double* d, const double* tmin, const double* tmax,
               int any_hit, int64_t* prim, double* t, double* u, double* v, uint32_t* steps, uint32_t* isects) {
    Group* g = static_cast<Group*>(gp);
    auto& bvh = g->bvh_cache;
    auto* tri = g->triangles.data();
    bvh::SingleRayTraverser<bvh::Bvh<Scalar>> trav(bvh);
    bvh::ClosestPrimitiveIntersector<bvh::Bvh<Scalar>, bvh::Triangle<Scalar>, false> ci(bvh, tri);
    bvh::AnyPrimitiveIntersector<bvh::Bvh<Scalar>, bvh::Triangle<Scalar>, false> ai(bvh, tri);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        bvh::Ray<Scalar> ray(bvh::Vector3<Scalar>(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                             bvh::Vector3<Scalar>(d[3 * i], d[3 * i + 1], d[3 * i + 2]),
                             tmin ? tmin[i] : 0.0,
                             tmax ? tmax[i] : std::numeric_limits<Scalar>::max());
        typename bvh::SingleRayTraverser<bvh::Bvh<Scalar>>::Statistics st;
        if (any_hit) {
            auto h = trav.traverse(ray, ai, st);
            if (prim) prim[i] = h ? 0 : -1;
            if (t) t[i] = h ? h->t : -1.0;
        } else {
            auto h = trav.traverse(ray, ci, st);
            if (prim) prim[i] = h ? (int64_t)h->primitive_index : -1;
            if (t) t[i] = h ? h->intersection.t : -1.0;
            if (u) u[i] = h ? h->intersection.u : 0.0;
            if (v) v[i] = h ? h->intersection.v : 0.0;
        }
        if (steps) steps[i] = (uint32_t)st.traversal_steps;
        if (isects) isects[i] = (uint32_t)st.intersections;
    }
}

// BodyFixedGroup passes (body_fixed_group.hpp:111-124); any output may be null.
void ref_pass(void* gp, const double* cam, double* xyz, uint32_t* inst, double* normal) {
    Group* g = static_cast<Group*>(gp);
    auto c = make_camera(cam);
    const size_t w = (size_t)floor(c->get_resolutionX()), h = (size_t)floor(c->get_resolutionY());
    if (xyz) {
        auto r = g->intersection_pass(c);
        std::copy(r.data(), r.data() + 3 * w * h, xyz);
    }
    if (inst) {
        auto r = g->instance_pass(c);
        std::copy(r.data(), r.data() + w * h, inst);
    }
    if (normal) {
        auto r = g->normal_pass(c);
        std::copy(r.data(), r.data() + 3 * w * h, normal);
    }
}

// BodyFixedGroup::render (uint8 RGBA, do_render.hpp).
void ref_render_u8(void* gp, const double* cam, int nl, const double* lights, int min_s, int max_s, double thr,
                   int nb, uint8_t* out) {
    Group* g = static_cast<Group*>(gp);
    auto c = make_camera(cam);
    auto L = make_lights(nl, lights);
    auto img = g->render(c, L, min_s, max_s, thr, nb);
    std::copy(img.begin(), img.end(), out);
}

// Float radiance with the per-pixel loop of do_render.hpp:65-116 (fixed spp,
// jitter U(-0.5,0.5)), calling the reference integrator unidirectional()
// (path_tracing/unidirectional.hpp:24-96).  Jitter RNG: one minstd_rand per
// thread seeded from random_device (the reference shares one, racily).
void ref_render_float(void* gp, const double* cam, int nl, const double* lights, int spp, int nb, int threads,
                      float* out) {
    Group* g = static_cast<Group*>(gp);
    auto c = make_camera(cam);
    auto L = make_lights(nl, lights);
    auto& bvh = g->bvh_cache;
    auto* tri = g->triangles.data();
    bvh::ClosestPrimitiveIntersector<bvh::Bvh<Scalar>, bvh::Triangle<Scalar>, false> ci(bvh, tri);
    bvh::AnyPrimitiveIntersector<bvh::Bvh<Scalar>, bvh::Triangle<Scalar>, false> ai(bvh, tri);
    bvh::SingleRayTraverser<bvh::Bvh<Scalar>> trav(bvh);
    const size_t w = (size_t)floor(c->get_resolutionX()), h = (size_t)floor(c->get_resolutionY());
    std::random_device rd;
    const int nt = threads > 0 ? threads : omp_get_max_threads();
    std::vector<unsigned> seeds(nt);
    for (auto& s : seeds) s = rd();
#pragma omp parallel num_threads(nt)
    {
        std::minstd_rand eng(seeds[omp_get_thread_num()]);
        std::uniform_real_distribution<Scalar> distr(-0.5, 0.5);
#pragma omp for schedule(dynamic, 1)
        for (size_t j = 0; j < h; ++j)
            for (size_t i = 0; i < w; ++i) {
                Color px(0);
                for (int s = 1; s < spp + 1; ++s) {
                    const auto di = distr(eng), dj = distr(eng);
                    auto ray = spp == 1 ? c->pixel_to_ray(i, j) : c->pixel_to_ray(i + di, j + dj);
                    Color pr = unidirectional(L, tri, ci, ai, trav, ray, nb);
                    px += (pr - px) * (1.0f / s);
                }
                out[3 * (w * j + i)] = px[0];
                out[3 * (w * j + i) + 1] = px[1];
                out[3 * (w * j + i) + 2] = px[2];
            }
    }
}

// BodyFixedGroup::simulate_lidar / batch_simulate_lidar (do_lidar.hpp).
double ref_lidar(void* gp, const double* pos, const double* rot, int z_positive) {
    Group* g = static_cast<Group*>(gp);
    auto l = std::make_unique<SimpleLidar<Scalar>>(z_positive != 0);
    Scalar r[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[i][j] = rot[3 * i + j];
    l->set_pose(bvh::Vector3<Scalar>(pos[0], pos[1], pos[2]), r);
    std::unique_ptr<Lidar<Scalar>> lp(l.release());
    return g->simulate_lidar(lp, 1);
}

void ref_batch_lidar(void* gp, int n, const double* pos, const double* rot /* n x 9 row-major */, int z_positive,
                     double* out) {
    Group* g = static_cast<Group*>(gp);
    auto l = std::make_unique<SimpleLidar<Scalar>>(z_positive != 0);
    std::vector<bvh::Vector3<Scalar>> P;
    for (int k = 0; k < n; ++k) P.emplace_back(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]);
    // batch_rotations[i][j][k] (lidar.hpp:11); 7.2 MB, so heap-allocated here
    auto R = std::make_unique<Scalar[][3][100000]>(3);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 100000; ++k) R[i][j][k] = k < n ? rot[9 * k + 3 * i + j] : 0.0;
    l->batch_set_pose(P, R.get());
    std::unique_ptr<Lidar<Scalar>> lp(l.release());
    auto d = g->batch_simulate_lidar(lp, 1);
    std::copy(d.begin(), d.end(), out);
}

}  // extern "C"
