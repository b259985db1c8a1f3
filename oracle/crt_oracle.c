/*
 * crt_oracle.c -- FP64 CPU restatement of the ceres-raytracer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP core: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product path never calls it (the HIP library has no CPU fall-back).
 *
 * Each function restates the reference's algorithm in plain C, in double
 * precision, and cites the file:line it follows (paths relative to the
 * ceres-raytracer v0.6a4 tree).  Where the reference depends on state it
 * cannot reproduce (random_device seeds, OpenMP-order tie winners), the
 * oracle documents the deterministic choice it makes instead.
 *
 * Acceleration: a binary median-split BVH (FP64 boxes) is used only to make
 * the oracle fast enough; it returns the same closest hit as the brute-force
 * loop over all triangles except where two hits have exactly equal t
 * (crt_oracle_closest_brute pins that on small inputs).
 */
#include "crt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- vectors */
typedef struct { double x, y, z; } v3;
static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* lib/bvh/vector.hpp:166-173 */
static inline v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline double len(v3 a) { return sqrt(dot(a, a)); }
/* lib/bvh/vector.hpp:159-163: normalize = v * (1 / length) */
static inline v3 nrm(v3 a) { double inv = 1.0 / len(a); return scl(a, inv); }

/* ---------------------------------------------------------------- scene */
typedef struct { double lo[3], hi[3]; int32_t left, right, first, count; } bnode;

struct orc_scene {
    int64_t nt, nv, ni;
    v3 *p0, *p1, *p2;     /* triangle vertices (FP32 input promoted to FP64) */
    v3 *e1, *e2, *n;      /* lib/bvh/triangle.hpp:36-40: e1 = p0-p1, e2 = p2-p0, n = e1 x e2 */
    uint32_t* inst;       /* per triangle instance id */
    uint32_t* vidx;       /* per triangle vertex indices */
    v3* vn;               /* per vertex normals or NULL */
    float* color;         /* per instance RGB */
    uint8_t* smooth;      /* per instance */
    bnode* nodes;
    int32_t* prim;
    int64_t nnodes;
    crt_oracle_material* mat; /* per instance, NULL: ColoredLambertian with `color` */
    float* uv;                /* per vertex (u, v), NULL: zero */
    int ntex;
    crt_oracle_texture* tex;  /* pixel data owned by the scene */
};

static void bbox_of(const orc_scene* s, const int32_t* idx, int64_t cnt, double lo[3], double hi[3]) {
    for (int k = 0; k < 3; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
    for (int64_t i = 0; i < cnt; ++i) {
        const int32_t t = idx[i];
        const v3 q[3] = {s->p0[t], s->p1[t], s->p2[t]};
        for (int j = 0; j < 3; ++j) {
            const double c[3] = {q[j].x, q[j].y, q[j].z};
            for (int k = 0; k < 3; ++k) {
                if (c[k] < lo[k]) lo[k] = c[k];
                if (c[k] > hi[k]) hi[k] = c[k];
            }
        }
    }
}

static const orc_scene* g_sort_scene;
static int g_sort_axis;
static double centroid_axis(const orc_scene* s, int32_t t, int a) {
    const double c0[3] = {s->p0[t].x, s->p0[t].y, s->p0[t].z};
    const double c1[3] = {s->p1[t].x, s->p1[t].y, s->p1[t].z};
    const double c2[3] = {s->p2[t].x, s->p2[t].y, s->p2[t].z};
    return c0[a] + c1[a] + c2[a];
}
static int cmp_centroid(const void* a, const void* b) {
    const double ca = centroid_axis(g_sort_scene, *(const int32_t*)a, g_sort_axis);
    const double cb = centroid_axis(g_sort_scene, *(const int32_t*)b, g_sort_axis);
    if (ca < cb) return -1;
    if (ca > cb) return 1;
    return (*(const int32_t*)a) - (*(const int32_t*)b);
}

static int32_t build_rec(orc_scene* s, int32_t first, int32_t count) {
    const int32_t id = (int32_t)s->nnodes++;
    bnode* nd = &s->nodes[id];
    bbox_of(s, s->prim + first, count, nd->lo, nd->hi);
    nd->first = first;
    nd->count = count;
    nd->left = nd->right = -1;
    if (count <= 4) return id;
    int axis = 0;
    double ext = -1;
    for (int k = 0; k < 3; ++k)
        if (nd->hi[k] - nd->lo[k] > ext) { ext = nd->hi[k] - nd->lo[k]; axis = k; }
    g_sort_scene = s;
    g_sort_axis = axis;
    qsort(s->prim + first, (size_t)count, sizeof(int32_t), cmp_centroid);
    const int32_t half = count / 2;
    const int32_t l = build_rec(s, first, half);
    const int32_t r = build_rec(s, first + half, count - half);
    s->nodes[id].left = l;
    s->nodes[id].right = r;
    s->nodes[id].count = 0;
    return id;
}

/* Vertices come as FP32 (vf) or FP64 (vd); the reference holds its
 * transformed geometry in double (src/pybind.cpp:30, src/crt/transform.hpp:42-85). */
static orc_scene* scene_create(const float* vf, const double* vd, int64_t nv, const uint32_t* indices, int64_t nt,
                               const uint32_t* tri_inst, const float* vnormals, int64_t ni, const float* colors,
                               const uint8_t* smooth);

orc_scene* crt_oracle_scene_create(const float* vertices, int64_t nv, const uint32_t* indices, int64_t nt,
                                   const uint32_t* tri_inst, const float* vnormals, int64_t ni,
                                   const float* colors, const uint8_t* smooth) {
    return scene_create(vertices, NULL, nv, indices, nt, tri_inst, vnormals, ni, colors, smooth);
}

orc_scene* crt_oracle_scene_create_f64(const double* vertices, int64_t nv, const uint32_t* indices, int64_t nt,
                                       const uint32_t* tri_inst, const float* vnormals, int64_t ni,
                                       const float* colors, const uint8_t* smooth) {
    return scene_create(NULL, vertices, nv, indices, nt, tri_inst, vnormals, ni, colors, smooth);
}

#define VTX(i) (vd ? vd[(i)] : (double)vf[(i)])

static orc_scene* scene_create(const float* vf, const double* vd, int64_t nv, const uint32_t* indices, int64_t nt,
                               const uint32_t* tri_inst, const float* vnormals, int64_t ni, const float* colors,
                               const uint8_t* smooth) {
    if (nt <= 0 || nv <= 0) return NULL;
    orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
    s->nt = nt; s->nv = nv; s->ni = ni > 0 ? ni : 1;
    s->p0 = (v3*)malloc(sizeof(v3) * nt); s->p1 = (v3*)malloc(sizeof(v3) * nt); s->p2 = (v3*)malloc(sizeof(v3) * nt);
    s->e1 = (v3*)malloc(sizeof(v3) * nt); s->e2 = (v3*)malloc(sizeof(v3) * nt); s->n = (v3*)malloc(sizeof(v3) * nt);
    s->inst = (uint32_t*)malloc(4 * nt);
    s->vidx = (uint32_t*)malloc(12 * nt);
    for (int64_t t = 0; t < nt; ++t) {
        const uint32_t a = indices[3 * t], b = indices[3 * t + 1], c = indices[3 * t + 2];
        s->vidx[3 * t] = a; s->vidx[3 * t + 1] = b; s->vidx[3 * t + 2] = c;
        s->p0[t] = mk(VTX(3 * a), VTX(3 * a + 1), VTX(3 * a + 2));
        s->p1[t] = mk(VTX(3 * b), VTX(3 * b + 1), VTX(3 * b + 2));
        s->p2[t] = mk(VTX(3 * c), VTX(3 * c + 1), VTX(3 * c + 2));
        s->e1[t] = sub(s->p0[t], s->p1[t]);
        s->e2[t] = sub(s->p2[t], s->p0[t]);
        s->n[t] = cross(s->e1[t], s->e2[t]);
        s->inst[t] = tri_inst ? tri_inst[t] : 1u;
    }
    if (vnormals) {
        s->vn = (v3*)malloc(sizeof(v3) * nv);
        for (int64_t i = 0; i < nv; ++i) s->vn[i] = mk(vnormals[3 * i], vnormals[3 * i + 1], vnormals[3 * i + 2]);
    }
    s->color = (float*)malloc(12 * s->ni);
    s->smooth = (uint8_t*)calloc((size_t)s->ni, 1);
    for (int64_t i = 0; i < s->ni; ++i) {
        for (int k = 0; k < 3; ++k) s->color[3 * i + k] = colors ? colors[3 * i + k] : 1.0f;
        if (smooth) s->smooth[i] = smooth[i];
    }
    s->prim = (int32_t*)malloc(4 * nt);
    for (int64_t t = 0; t < nt; ++t) s->prim[t] = (int32_t)t;
    s->nodes = (bnode*)malloc(sizeof(bnode) * (2 * nt + 1));
    s->nnodes = 0;
    build_rec(s, 0, (int32_t)nt);
    return s;
}

static void free_materials(orc_scene* s) {
    for (int k = 0; k < s->ntex; ++k) free((void*)s->tex[k].rgba);
    free(s->tex); free(s->mat); free(s->uv);
    s->tex = NULL; s->mat = NULL; s->uv = NULL; s->ntex = 0;
}

void crt_oracle_scene_destroy(orc_scene* s) {
    if (!s) return;
    free(s->p0); free(s->p1); free(s->p2); free(s->e1); free(s->e2); free(s->n);
    free(s->inst); free(s->vidx); free(s->vn); free(s->color); free(s->smooth);
    free(s->prim); free(s->nodes);
    free_materials(s);
    free(s);
}

void crt_oracle_scene_set_materials(orc_scene* s, int64_t n, const crt_oracle_material* mats, const float* uv,
                                    int64_t nv, int ntex, const crt_oracle_texture* tex) {
    free_materials(s);
    s->mat = (crt_oracle_material*)calloc((size_t)s->ni, sizeof(crt_oracle_material));
    for (int64_t i = 0; i < s->ni; ++i) {
        if (i < n) s->mat[i] = mats[i];
        else { s->mat[i].type = 0; s->mat[i].texture = s->mat[i].spec_texture = -1;
               for (int k = 0; k < 3; ++k) s->mat[i].color[k] = s->color[3 * i + k]; }
    }
    if (uv && nv == s->nv) {
        s->uv = (float*)malloc(sizeof(float) * 2 * nv);
        memcpy(s->uv, uv, sizeof(float) * 2 * nv);
    }
    s->ntex = ntex;
    s->tex = (crt_oracle_texture*)calloc((size_t)(ntex > 0 ? ntex : 1), sizeof(crt_oracle_texture));
    for (int k = 0; k < ntex; ++k) {
        const size_t b = (size_t)tex[k].width * tex[k].height * 4;
        uint8_t* px = (uint8_t*)malloc(b > 0 ? b : 1);
        memcpy(px, tex[k].rgba, b);
        s->tex[k].width = tex[k].width;
        s->tex[k].height = tex[k].height;
        s->tex[k].rgba = px;
    }
}

/* ------------------------------------------------------ triangle test */
/* lib/bvh/triangle.hpp:140-160 (LeftHandedNormal = true):
 *   c = p0 - o; r = d x c; inv = 1/(n.d); u = (r.e2) inv; v = (r.e1) inv; w = 1-u-v
 *   accept if u,v,w >= 0 and tmin <= t = (n.c) inv <= tmax (all inclusive)      */
static inline int tri_isect(const orc_scene* s, int64_t t_i, v3 o, v3 d, double tmin, double tmax,
                            double* t, double* u, double* v) {
    const v3 c = sub(s->p0[t_i], o);
    const v3 r = cross(d, c);
    const double inv_det = 1.0 / dot(s->n[t_i], d);
    const double uu = dot(r, s->e2[t_i]) * inv_det;
    const double vv = dot(r, s->e1[t_i]) * inv_det;
    const double ww = 1.0 - uu - vv;
    if (uu >= 0 && vv >= 0 && ww >= 0) {
        const double tt = dot(s->n[t_i], c) * inv_det;
        if (tt >= tmin && tt <= tmax) { *t = tt; *u = uu; *v = vv; return 1; }
    }
    return 0;
}

int crt_oracle_tri_intersect(const double p0[3], const double p1[3], const double p2[3], const double o[3],
                             const double d[3], double tmin, double tmax, double out[3]) {
    orc_scene tmp;
    v3 P0 = mk(p0[0], p0[1], p0[2]), P1 = mk(p1[0], p1[1], p1[2]), P2 = mk(p2[0], p2[1], p2[2]);
    v3 E1 = sub(P0, P1), E2 = sub(P2, P0), N = cross(E1, E2);
    memset(&tmp, 0, sizeof(tmp));
    tmp.p0 = &P0; tmp.e1 = &E1; tmp.e2 = &E2; tmp.n = &N;
    return tri_isect(&tmp, 0, mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]), tmin, tmax, &out[0], &out[1], &out[2]);
}

/* ------------------------------------------------------ traversal */
/* Slab test in FP64 with the safe inverse of lib/bvh/vector.hpp:75-80 and the
 * entry/exit rule of lib/bvh/node_intersectors.hpp:35-47 (hit iff entry <= exit). */
static inline int box_hit(const bnode* nd, v3 o, v3 inv, double tmin, double tmax, double* tent) {
    const double oo[3] = {o.x, o.y, o.z}, ii[3] = {inv.x, inv.y, inv.z};
    double t0 = tmin, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
        double a = (nd->lo[k] - oo[k]) * ii[k], b = (nd->hi[k] - oo[k]) * ii[k];
        if (a > b) { double x = a; a = b; b = x; }
        if (a > t0) t0 = a;
        if (b < t1) t1 = b;
    }
    *tent = t0;
    /* pad the exit by a few ulp: the oracle must not lose grazing hits */
    return t0 <= t1 * (1.0 + 1e-12) + 1e-300;
}

static inline double safe_inv(double x) {
    const double eps = 2.220446049250313e-16; /* DBL_EPSILON, vector.hpp:76 */
    return 1.0 / (fabs(x) < eps ? copysign(eps, x) : x);
}

/* Closest hit (any_hit=0) or first hit (any_hit=1).  Tie rule: a later hit
 * with t <= current best replaces it (single_ray_traverser.hpp:55-60). */
static int64_t trace(const orc_scene* s, v3 o, v3 d, double tmin, double tmax, int any_hit, double* tb, double* ub,
                     double* vb) {
    const v3 inv = mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    int32_t stack[128];
    int sp = 0;
    int64_t best = -1;
    double bt = tmax;
    stack[sp++] = 0;
    while (sp > 0) {
        const bnode* nd = &s->nodes[stack[--sp]];
        double te;
        if (!box_hit(nd, o, inv, tmin, bt, &te)) continue;
        if (nd->count > 0) {
            for (int32_t i = 0; i < nd->count; ++i) {
                const int32_t t = s->prim[nd->first + i];
                double tt, uu, vv;
                if (tri_isect(s, t, o, d, tmin, bt, &tt, &uu, &vv)) {
                    best = t; bt = tt; *tb = tt; *ub = uu; *vb = vv;
                    if (any_hit) return best;
                }
            }
        } else {
            stack[sp++] = nd->right;
            stack[sp++] = nd->left;
        }
    }
    return best;
}

int64_t crt_oracle_closest_brute(const orc_scene* s, const double o[3], const double d[3], double tmin, double tmax,
                                 double out[3]) {
    int64_t best = -1;
    double bt = tmax;
    const v3 O = mk(o[0], o[1], o[2]), D = mk(d[0], d[1], d[2]);
    for (int64_t t = 0; t < s->nt; ++t) {
        double tt, uu, vv;
        if (tri_isect(s, t, O, D, tmin, bt, &tt, &uu, &vv)) { best = t; bt = tt; out[0] = tt; out[1] = uu; out[2] = vv; }
    }
    return best;
}

void crt_oracle_trace(const orc_scene* s, int64_t n, const double* o, const double* d, const double* tmin,
                      const double* tmax, int any_hit, double* t, int32_t* tri, double* u, double* v) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        double tt = -1, uu = 0, vv = 0;
        const int64_t h = trace(s, mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]), mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]),
                                tmin ? tmin[i] : 0.0, tmax ? tmax[i] : INFINITY, any_hit, &tt, &uu, &vv);
        if (t) t[i] = h >= 0 ? tt : -1.0;
        if (tri) tri[i] = (int32_t)h;
        if (u) u[i] = h >= 0 ? uu : 0.0;
        if (v) v[i] = h >= 0 ? vv : 0.0;
    }
}

/* ------------------------------------------------------ fragility */
/* Triangle test with every edge moved outwards by `slack` / |cos| (negative:
 * inwards), cos = angle between the unit ray and the triangle's normal: an
 * uncertainty of `slack` across the ray moves a hit by slack / |cos| within
 * the triangle's plane.  The barycentric coordinate of a vertex times the
 * triangle's doubled area |n| over the opposite edge's length is the
 * distance of the point from that edge. */
static inline int tri_isect_slack(const orc_scene* s, int64_t t_i, v3 o, v3 d, double tmin, double tmax,
                                  double slack, double* t) {
    const v3 c = sub(s->p0[t_i], o);
    const v3 r = cross(d, c);
    const double nd = dot(s->n[t_i], d);
    if (nd == 0.0) return 0;
    const double inv_det = 1.0 / nd;
    const double uu = dot(r, s->e2[t_i]) * inv_det;
    const double vv = dot(r, s->e1[t_i]) * inv_det;
    const double ww = 1.0 - uu - vv;
    const double an = len(s->n[t_i]);
    if (an == 0.0) return 0;
    double cs = fabs(nd) / (an * len(d));
    if (cs < 1e-3) cs = 1e-3;
    const double sl = slack / cs;
    const double l0 = len(s->e2[t_i]), l1 = len(s->e1[t_i]), l2 = len(sub(s->p2[t_i], s->p1[t_i]));
    if (uu * an < -sl * l0 || vv * an < -sl * l1 || ww * an < -sl * l2) return 0;
    const double tt = dot(s->n[t_i], c) * inv_det;
    if (!(tt >= tmin && tt <= tmax)) return 0;
    *t = tt;
    return 1;
}

/* Smallest distance from a hit point (barycentrics u, v) to an edge of t_i,
 * times |cos| of the ray (unit d) with the triangle's normal (see above). */
static double edge_margin(const orc_scene* s, int64_t t_i, v3 d, double u, double v) {
    const double an = len(s->n[t_i]);
    double cs = fabs(dot(s->n[t_i], d)) / (an * len(d));
    if (cs < 1e-3) cs = 1e-3;
    const double w = 1.0 - u - v;
    const double d0 = u * an / len(s->e2[t_i]), d1 = v * an / len(s->e1[t_i]);
    const double d2 = w * an / len(sub(s->p2[t_i], s->p1[t_i]));
    const double m = d0 < d1 ? d0 : d1;
    return (d2 < m ? d2 : m) * cs;
}

/* Closest / any hit over triangles grown (slack > 0) or shrunk (slack < 0);
 * node boxes are padded by |slack|. */
static int64_t trace_slack(const orc_scene* s, v3 o, v3 d, double tmin, double tmax, int any_hit, double slack,
                           double* tb) {
    const v3 inv = mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    const double pad = fabs(slack);
    int32_t stack[128];
    int sp = 0;
    int64_t best = -1;
    double bt = tmax;
    stack[sp++] = 0;
    while (sp > 0) {
        bnode nd = s->nodes[stack[--sp]];
        for (int k = 0; k < 3; ++k) { nd.lo[k] -= pad; nd.hi[k] += pad; }
        double te;
        if (!box_hit(&nd, o, inv, tmin, bt, &te)) continue;
        if (nd.count > 0) {
            for (int32_t i = 0; i < nd.count; ++i) {
                const int32_t t = s->prim[nd.first + i];
                double tt;
                if (tri_isect_slack(s, t, o, d, tmin, bt, slack, &tt)) {
                    best = t; bt = tt; *tb = tt;
                    if (any_hit) return best;
                }
            }
        } else {
            stack[sp++] = nd.right;
            stack[sp++] = nd.left;
        }
    }
    return best;
}

/* A closest-hit decision is fragile when the hit lies within eps of an edge
 * of its triangle, or when triangles grown by eps give another closest hit
 * (a near miss in front, or a near miss where the strict ray misses). */
static int closest_fragile(const orc_scene* s, v3 o, v3 d, double tmin, int64_t h, double u, double v, double eps) {
    if (h >= 0 && edge_margin(s, h, d, u, v) < eps) return 1;
    double tl;
    const int64_t hl = trace_slack(s, o, d, tmin, INFINITY, 0, eps, &tl);
    return hl != h;
}

/* A shadow ray (unit direction) is fragile when its occlusion differs between
 * triangles grown by eps over [tmin, tmax + eps] and shrunk by eps over
 * [tmin + eps, tmax - eps]. */
static int shadow_fragile(const orc_scene* s, v3 o, v3 d, double tmin, double tmax, double eps) {
    double t;
    const int loose = trace_slack(s, o, d, tmin, tmax + eps, 1, eps, &t) >= 0;
    const int tight = (tmin + eps <= tmax - eps) && trace_slack(s, o, d, tmin + eps, tmax - eps, 1, -eps, &t) >= 0;
    return loose != tight;
}

void crt_oracle_closest_fragile(const orc_scene* s, int64_t n, const double* o, const double* d, double eps,
                                uint8_t* flags) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const v3 O = mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]), D = nrm(mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
        double tt = 0, uu = 0, vv = 0;
        const int64_t h = trace(s, O, D, 0.0, INFINITY, 0, &tt, &uu, &vv);
        flags[i] = (uint8_t)closest_fragile(s, O, D, 0.0, h, uu, vv, eps);
    }
}

void crt_oracle_shadow_fragile(const orc_scene* s, int64_t n, const double* o, const double* d, const double* tmax,
                               double eps, uint8_t* flags) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const v3 O = mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]), D = nrm(mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
        flags[i] = (uint8_t)shadow_fragile(s, O, D, 0.0, tmax ? tmax[i] : INFINITY, eps);
    }
}

/* Tie / edge classification on the brute-force hit list: a ray is flagged
 * when a second triangle is hit within rel_t of the closest t, or when the
 * closest hit's smallest barycentric is below bary_eps. */
void crt_oracle_tie_flags(const orc_scene* s, int64_t n, const double* o, const double* d, double rel_t,
                          double bary_eps, uint8_t* flags) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
        const v3 O = mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]), D = mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        double bt = -1, bu = 0, bv = 0;
        const int64_t h = trace(s, O, D, 0.0, INFINITY, 0, &bt, &bu, &bv);
        uint8_t f = 0;
        if (h >= 0) {
            const double w = 1.0 - bu - bv;
            double mb = bu < bv ? bu : bv;
            if (w < mb) mb = w;
            if (mb <= bary_eps) f = 1;
            /* any other triangle within the window (t in [bt - tol, bt + tol]) */
            const double tol = rel_t * fabs(bt) + 1e-12;
            double t2, u2, v2;
            const v3 inv = mk(safe_inv(D.x), safe_inv(D.y), safe_inv(D.z));
            int32_t stack[128];
            int sp = 0;
            stack[sp++] = 0;
            while (sp > 0 && !f) {
                const bnode* nd = &s->nodes[stack[--sp]];
                double te;
                if (!box_hit(nd, O, inv, bt - tol, bt + tol, &te)) continue;
                if (nd->count > 0) {
                    for (int32_t k = 0; k < nd->count; ++k) {
                        const int32_t tt = s->prim[nd->first + k];
                        if (tt == h) continue;
                        if (tri_isect(s, tt, O, D, bt - tol, bt + tol, &t2, &u2, &v2)) { f = 1; break; }
                    }
                } else {
                    stack[sp++] = nd->right;
                    stack[sp++] = nd->left;
                }
            }
        }
        flags[i] = f;
    }
}

/* Edge classification for a disagreement: the ray (o,d) and a triangle
 * `other` chosen by another tracer.  Flags the ray when `other`'s barycentric
 * coordinates for this ray (computed without the acceptance test) are all
 * >= -bary_eps and one of them is <= bary_eps, i.e. the ray grazes an edge or
 * vertex of `other`, or when `other` is hit at a distance within rel_t of the
 * oracle's closest hit (an exact tie). */
void crt_oracle_edge_flags(const orc_scene* s, int64_t n, const double* o, const double* d, const int32_t* other,
                           double rel_t, double bary_eps, uint8_t* flags) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
        const v3 O = mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]), D = mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        uint8_t f = 0;
        const int32_t t_i = other[i];
        double bt = -1, bu = 0, bv = 0;
        const int64_t h = trace(s, O, D, 0.0, INFINITY, 0, &bt, &bu, &bv);
        if (h >= 0) {
            const double w = 1.0 - bu - bv;
            double mb = bu < bv ? bu : bv;
            if (w < mb) mb = w;
            if (mb <= bary_eps) f = 1;
        }
        if (!f && t_i >= 0 && t_i < s->nt) {
            const v3 c = sub(s->p0[t_i], O);
            const v3 r = cross(D, c);
            const double inv_det = 1.0 / dot(s->n[t_i], D);
            const double uu = dot(r, s->e2[t_i]) * inv_det;
            const double vv = dot(r, s->e1[t_i]) * inv_det;
            const double ww = 1.0 - uu - vv;
            const double tt = dot(s->n[t_i], c) * inv_det;
            double mn = uu < vv ? uu : vv;
            if (ww < mn) mn = ww;
            if (mn >= -bary_eps && mn <= bary_eps) f = 1;
            if (h >= 0 && fabs(tt - bt) <= rel_t * fabs(bt) && mn >= -bary_eps) f = 1;
        }
        flags[i] = f;
    }
}

/* ------------------------------------------------------ camera */
/* src/crt/cameras/simple_camera.hpp:14-27 (centre, scale) and :61-81 (ray). */
static void cam_ray(const crt_oracle_camera* c, double u, double v, v3* o, v3* d) {
    const double cx = c->resolution[0] / 2.0, cy = c->resolution[1] / 2.0;
    const double sx = c->resolution[0] / c->sensor_size[0], sy = c->resolution[1] / c->sensor_size[1];
    const double f = c->z_positive ? c->focal_length : -c->focal_length;
    const v3 dc = nrm(mk((-cx + u) / sx, (cy - v) / sy, f));
    const double* R = c->rotation;
    /* world = R^T * dir (simple_camera.hpp:71-76) */
    *d = mk(R[0] * dc.x + R[3] * dc.y + R[6] * dc.z, R[1] * dc.x + R[4] * dc.y + R[7] * dc.z,
            R[2] * dc.x + R[5] * dc.y + R[8] * dc.z);
    *o = mk(c->position[0], c->position[1], c->position[2]);
}

void crt_oracle_camera_rays(const crt_oracle_camera* c, double* o, double* d) {
    const int64_t w = (int64_t)floor(c->resolution[0]), h = (int64_t)floor(c->resolution[1]);
    for (int64_t j = 0; j < h; ++j)
        for (int64_t i = 0; i < w; ++i) {
            v3 O, D;
            cam_ray(c, (double)i, (double)j, &O, &D);
            const int64_t k = j * w + i;
            o[3 * k] = O.x; o[3 * k + 1] = O.y; o[3 * k + 2] = O.z;
            d[3 * k] = D.x; d[3 * k + 1] = D.y; d[3 * k + 2] = D.z;
        }
}

/* Barycentric point (src/crt/passes.hpp:53-58) and normal (passes.hpp:176-181). */
static v3 hit_point(const orc_scene* s, int64_t t, double u, double v) {
    return add(add(scl(s->p1[t], u), scl(s->p2[t], v)), scl(s->p0[t], 1.0 - u - v));
}
static v3 hit_normal(const orc_scene* s, int64_t t, double u, double v) {
    const uint32_t in = s->inst[t];
    if (s->vn && in >= 1 && in <= (uint32_t)s->ni && s->smooth[in - 1]) {
        const v3 a = s->vn[s->vidx[3 * t]], b = s->vn[s->vidx[3 * t + 1]], c = s->vn[s->vidx[3 * t + 2]];
        return nrm(add(add(scl(b, u), scl(c, v)), scl(a, 1.0 - u - v)));
    }
    return nrm(s->n[t]);
}

void crt_oracle_pass(const orc_scene* s, const crt_oracle_camera* c, double* xyz, double* normal, uint32_t* instance,
                     double* depth, int32_t* tri) {
    const int64_t w = (int64_t)floor(c->resolution[0]), h = (int64_t)floor(c->resolution[1]);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t j = 0; j < h; ++j)
        for (int64_t i = 0; i < w; ++i) {
            v3 O, D;
            cam_ray(c, (double)i, (double)j, &O, &D);
            double tt = 0, uu = 0, vv = 0;
            const int64_t hh = trace(s, O, D, 0.0, INFINITY, 0, &tt, &uu, &vv);
            const int64_t k = j * w + i;
            v3 p = mk(0, 0, 0), nn = mk(0, 0, 0);
            if (hh >= 0) { p = hit_point(s, hh, uu, vv); nn = hit_normal(s, hh, uu, vv); }
            if (xyz) { xyz[3 * k] = p.x; xyz[3 * k + 1] = p.y; xyz[3 * k + 2] = p.z; }
            if (normal) { normal[3 * k] = nn.x; normal[3 * k + 1] = nn.y; normal[3 * k + 2] = nn.z; }
            if (instance) instance[k] = hh >= 0 ? s->inst[hh] : 0u;
            if (depth) depth[k] = hh >= 0 ? len(sub(p, O)) : 0.0;
            if (tri) tri[k] = (int32_t)hh;
        }
}

/* ------------------------------------------------------ LiDAR */
/* src/crt/lidars/simple_lidar.hpp:42-96 + src/crt/do_lidar.hpp:43-70,111-143:
 * one ray along +/-z rotated by R^T from the pose; range = |hit - position|
 * (hit from barycentrics), 0 on a miss. */
void crt_oracle_lidar_batch(const orc_scene* s, int64_t n, const double* pos, const double* rot, int z_positive,
                            double* range) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const double* R = rot + 9 * i;
        const double z = z_positive ? 1.0 : -1.0;
        const v3 d = mk(R[6] * z, R[7] * z, R[8] * z);
        const v3 o = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        double tt, uu, vv;
        const int64_t h = trace(s, o, d, 0.0, INFINITY, 0, &tt, &uu, &vv);
        range[i] = h >= 0 ? len(sub(hit_point(s, h, uu, vv), o)) : 0.0;
    }
}

/* ------------------------------------------------------ counter RNG */
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double crt_oracle_u01(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
    uint64_t h = mix64(seed + 0x9E3779B97F4A7C15ull);
    h = mix64(h ^ a);
    h = mix64(h ^ (b + 0x632BE59BD9B4E019ull));
    h = mix64(h ^ (c * 0x8CB92BA72F3D8DD7ull + 1ull));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------ path tracer */
/* src/crt/materials/brdfs.hpp:17-42 (Malley cosine sampling, left-handed normal) */
static v3 cosine_importance(v3 n, double r1, double r2) {
    const double r = sqrt(r1);
    const double th = r2 * 2 * 3.1415926;
    const double x = -r * cos(th), y = -r * sin(th);
    double zz = 1.0 - x * x - y * y;
    const double z = -sqrt(zz > 0.0 ? zz : 0.0);
    v3 nx;
    if (fabs(n.x) > fabs(n.y)) nx = scl(mk(n.z, 0.0, -n.x), 1.0 / sqrt(n.x * n.x + n.z * n.z));
    else nx = scl(mk(0.0, n.z, -n.y), 1.0 / sqrt(n.y * n.y + n.z * n.z));
    const v3 ny = cross(n, nx);
    return add(add(scl(nx, x), scl(ny, y)), scl(n, z));
}

#define RNG_JITTER 0u
#define RNG_LIGHT 2u
#define RNG_BRDF 62u
static inline uint64_t rng_c(int bounce, unsigned dim) { return (uint64_t)bounce * 64u + dim; }

/* Light sample (point: src/crt/lights/point_light.hpp:32-39; area:
 * area_light.hpp:54-72, sample placed with R; directional: new). */
static void light_sample(const crt_oracle_light* L, v3 p, uint64_t seed, uint64_t key_a, uint64_t sample, int bounce,
                         int li, v3* dir, double* tmax, double* irr) {
    if (L->type == 0) {
        const v3 lp = mk(L->position[0], L->position[1], L->position[2]);
        const v3 dd = sub(lp, p);
        *dir = nrm(dd);
        *tmax = len(dd);
        const double q = L->intensity / dot(sub(p, lp), sub(p, lp));
        *irr = q < 10000.0 ? q : 10000.0;
    } else if (L->type == 1) {
        const double r1 = crt_oracle_u01(seed, key_a, sample, rng_c(bounce, RNG_LIGHT + 2 * li));
        const double r2 = crt_oracle_u01(seed, key_a, sample, rng_c(bounce, RNG_LIGHT + 2 * li + 1));
        const double x = L->size[0] * r1 - L->size[0] / 2, y = L->size[1] * r2 - L->size[1] / 2;
        const double* R = L->rotation;
        const v3 sp = mk(R[0] * x + R[1] * y + L->position[0], R[3] * x + R[4] * y + L->position[1],
                         R[6] * x + R[7] * y + L->position[2]);
        const v3 dd = sub(sp, p);
        *dir = nrm(dd);
        *tmax = len(dd);
        *irr = L->intensity / dot(sub(p, sp), sub(p, sp));
    } else {
        *dir = nrm(mk(L->direction[0], L->direction[1], L->direction[2]));
        *tmax = INFINITY;
        *irr = L->intensity;
    }
}

/* ImageUVMap::operator() (src/crt/materials/material.hpp:50-60): nearest
 * texel x = (size_t)(u W + 0.5), y = (size_t)(v H + 0.5), channels / 255.
 * The indices are clamped to the image, where the reference reads past the
 * end of its pixel array for u or v near 1. */
static void tex_lookup(const orc_scene* s, int idx, float u, float v, float out[3]) {
    const crt_oracle_texture* T = &s->tex[idx];
    const double fx = (double)(u * (float)T->width) + 0.5, fy = (double)(v * (float)T->height) + 0.5;
    int64_t x = fx > 0.0 ? (int64_t)fx : 0, y = fy > 0.0 ? (int64_t)fy : 0;
    if (x > (int64_t)T->width - 1) x = (int64_t)T->width - 1;
    if (y > (int64_t)T->height - 1) y = (int64_t)T->height - 1;
    const uint8_t* p = T->rgba + 4 * ((size_t)T->width * (size_t)y + (size_t)x);
    for (int k = 0; k < 3; ++k) out[k] = p[k] / 255.f;
}

/* Texel coordinate within 2e-3 of a texel boundary: FP32 barycentrics may
 * pick the neighbouring texel. */
static int texel_fragile(const orc_scene* s, int idx, float u, float v) {
    const crt_oracle_texture* T = &s->tex[idx];
    const double fx = (double)(u * (float)T->width) + 0.5, fy = (double)(v * (float)T->height) + 0.5;
    const double ex = fabs(fx - floor(fx + 0.5)), ey = fabs(fy - floor(fy + 0.5));
    return (ex < 2e-3 && fx > 0.5 && fx < T->width - 0.5) || (ey < 2e-3 && fy > 0.5 && fy < T->height - 0.5);
}

static crt_oracle_material material_of(const orc_scene* s, int64_t h) {
    const uint32_t in = s->inst[h];
    crt_oracle_material m;
    if (s->mat && in >= 1 && in <= (uint32_t)s->ni) return s->mat[in - 1];
    memset(&m, 0, sizeof(m));
    m.texture = m.spec_texture = -1;
    for (int k = 0; k < 3; ++k) m.color[k] = (in >= 1 && in <= (uint32_t)s->ni) ? s->color[3 * (in - 1) + k] : 1.f;
    return m;
}

/* Per-vertex UVs interpolated like unidirectional.hpp:55 (float arithmetic,
 * u * uv1 + v * uv2 first, then + w * uv0). */
static void interp_uv(const orc_scene* s, int64_t h, double u, double v, float* iu, float* iv) {
    *iu = 0.f; *iv = 0.f;
    if (!s->uv) return;
    const float* a = s->uv + 2 * s->vidx[3 * h];
    const float* b = s->uv + 2 * s->vidx[3 * h + 1];
    const float* c = s->uv + 2 * s->vidx[3 * h + 2];
    const float fu = (float)u, fv = (float)v, fw = (float)(1.0 - u - v);
    *iu = fu * b[0] + fv * c[0] + fw * a[0];
    *iv = fu * b[1] + fv * c[1] + fw * a[1];
}

/* Material::compute for the light direction ld (unit, towards the light),
 * view direction vd (the incoming path ray) and shading normal sn:
 * ColoredLambertian / TexturedLambertian (material.hpp:100-103, 123-126):
 * map(u,v) * (float)(-L.n); TexturedBlinnPhong (material.hpp:152-160):
 * clamp(map (ka + kd diffuse) + ks spec), spec = pow(n^.normalize(L - vd),
 * alpha); Mirror (material.hpp:176-178): 0. */
static void mat_compute(const orc_scene* s, const crt_oracle_material* m, float iu, float iv, v3 ld, v3 vd, v3 sn,
                        float c[3]) {
    float col[3];
    if (m->type == 3) { c[0] = c[1] = c[2] = 0.f; return; }
    if (m->type == 0 || m->texture < 0) for (int k = 0; k < 3; ++k) col[k] = m->color[k];
    else tex_lookup(s, m->texture, iu, iv, col);
    if (m->type != 2) {
        const float ldn = (float)(-dot(ld, sn));
        for (int k = 0; k < 3; ++k) c[k] = col[k] * ldn;
        return;
    }
    float co[3];
    if (m->spec_texture < 0) for (int k = 0; k < 3; ++k) co[k] = m->coeffs[k];
    else tex_lookup(s, m->spec_texture, iu, iv, co);
    const float diffuse = -(float)dot(ld, sn);
    const float spec = (float)pow(dot(nrm(sn), nrm(add(scl(vd, -1.0), ld))), (double)m->alpha);
    for (int k = 0; k < 3; ++k) {
        const float x = col[k] * (co[0] + co[1] * diffuse) + co[2] * spec * 1.0f;
        c[k] = x < 0.f ? 0.f : (1.f < x ? 1.f : x);   /* std::clamp keeps NaN */
    }
}

/* src/crt/path_tracing/unidirectional.hpp:24-96 with float Color arithmetic;
 * the material of the hit triangle's instance computes the direct light and
 * samples the next direction (Material::compute / sample).  `frag` (may be
 * NULL) counts fragile decisions: [0] closest hits, [1] shadow rays, [2]
 * texel choices. */
static void unidirectional(const orc_scene* s, const crt_oracle_light* lights, int nl, v3 o, v3 d, int nb,
                           uint64_t seed, uint64_t key_a, uint64_t sample, double offset, double eps, float out[3],
                           int32_t* frag) {
    double tt, uu, vv;
    int64_t h = trace(s, o, d, 0.0, INFINITY, 0, &tt, &uu, &vv);
    if (frag && closest_fragile(s, o, d, 0.0, h, uu, vv, eps)) frag[0]++;
    float path[3] = {0.f, 0.f, 0.f};
    float weight[3];
    weight[0] = weight[1] = weight[2] = (float)(2 * M_PI);
    for (int b = 0; b < nb; ++b) {
        if (h < 0) break;
        const v3 gn = nrm(s->n[h]);
        const v3 sn = hit_normal(s, h, uu, vv);
        const crt_oracle_material m = material_of(s, h);
        float iu, iv;
        interp_uv(s, h, uu, vv, &iu, &iv);
        if (frag && m.type != 0 && m.type != 3 &&
            ((m.texture >= 0 && texel_fragile(s, m.texture, iu, iv)) ||
             (m.type == 2 && m.spec_texture >= 0 && texel_fragile(s, m.spec_texture, iu, iv))))
            frag[2]++;
        v3 p = hit_point(s, h, uu, vv);
        p = sub(p, scl(gn, offset));
        float lr[3] = {0.f, 0.f, 0.f};
        if (m.type != 3) {   /* a mirror's compute() is 0 whatever the shadow ray finds */
            for (int l = 0; l < nl; ++l) {
                v3 ld;
                double ltmax, irr;
                light_sample(&lights[l], p, seed, key_a, sample, b, l, &ld, &ltmax, &irr);
                double t2, u2, v2;
                const int64_t occ = trace(s, p, ld, 0.0, ltmax, 1, &t2, &u2, &v2);
                if (frag && shadow_fragile(s, p, ld, 0.0, ltmax, eps)) frag[1]++;
                float c[3] = {0.f, 0.f, 0.f};
                if (occ < 0) mat_compute(s, &m, iu, iv, ld, d, sn, c);
                for (int k = 0; k < 3; ++k) lr[k] += c[k] * (float)irr;
            }
        }
        if (b >= 1)
            for (int k = 0; k < 3; ++k) lr[k] = lr[k] < 0.f ? 0.f : (lr[k] > 1.f ? 1.f : lr[k]);
        for (int k = 0; k < 3; ++k) path[k] += lr[k] * weight[k];
        if (b == nb - 1) break;
        v3 nd;
        if (m.type == 3) {
            /* MirrorMaterial::sample (material.hpp:179-182): d - 2 n (d.n), weight x 1 */
            nd = sub(d, scl(scl(sn, 2.0), dot(d, sn)));
        } else {
            const double r1 = crt_oracle_u01(seed, key_a, sample, rng_c(b, RNG_BRDF));
            const double r2 = crt_oracle_u01(seed, key_a, sample, rng_c(b, RNG_BRDF + 1));
            nd = cosine_importance(sn, r1, r2);
            const float bc = (float)(1 - r1);
            for (int k = 0; k < 3; ++k) weight[k] *= bc;
        }
        o = p;
        d = nd;
        h = trace(s, o, d, 0.0, INFINITY, 0, &tt, &uu, &vv);
        if (frag && closest_fragile(s, o, d, 0.0, h, uu, vv, eps)) frag[0]++;
    }
    out[0] = path[0]; out[1] = path[1]; out[2] = path[2];
}

/* src/crt/do_render.hpp:65-116: per pixel running mean with adaptive stop. */
void crt_oracle_render_ex(const orc_scene* s, const crt_oracle_camera* c, const crt_oracle_light* lights, int nl,
                          int min_samples, int max_samples, double noise_threshold, int num_bounces, uint64_t seed,
                          uint64_t key_offset, double offset, double eps, float* rgb, int32_t* spp_used,
                          int32_t* fragile) {
    const int64_t w = (int64_t)floor(c->resolution[0]), h = (int64_t)floor(c->resolution[1]);
#pragma omp parallel for schedule(dynamic, 2)
    for (int64_t j = 0; j < h; ++j)
        for (int64_t i = 0; i < w; ++i) {
            const int64_t k = j * w + i;
            const uint64_t key_a = key_offset + (uint64_t)k;
            float px[3] = {0.f, 0.f, 0.f};
            int32_t fr[3] = {0, 0, 0};
            int used = 0;
            for (int smp = 1; smp < max_samples + 1; ++smp) {
                double ju = (double)i, jv = (double)j;
                if (max_samples != 1) {
                    ju += crt_oracle_u01(seed, key_a, (uint64_t)smp, rng_c(0, RNG_JITTER)) - 0.5;
                    jv += crt_oracle_u01(seed, key_a, (uint64_t)smp, rng_c(0, RNG_JITTER + 1)) - 0.5;
                }
                v3 O, D;
                cam_ray(c, ju, jv, &O, &D);
                float pr[3];
                unidirectional(s, lights, nl, O, D, num_bounces, seed, key_a, (uint64_t)smp, offset, eps, pr,
                               fragile ? fr : NULL);
                float rc[3];
                const float inv = 1.0f / (float)smp;
                for (int q = 0; q < 3; ++q) { rc[q] = (pr[q] - px[q]) * inv; px[q] += rc[q]; }
                used = smp;
                if (smp >= min_samples) {
                    const double noise = sqrt((double)(rc[0] * rc[0] + rc[1] * rc[1] + rc[2] * rc[2]));
                    if (noise < noise_threshold) break;
                }
            }
            rgb[3 * k] = px[0]; rgb[3 * k + 1] = px[1]; rgb[3 * k + 2] = px[2];
            if (spp_used) spp_used[k] = used;
            if (fragile) { fragile[3 * k] = fr[0]; fragile[3 * k + 1] = fr[1]; fragile[3 * k + 2] = fr[2]; }
        }
}

void crt_oracle_render(const orc_scene* s, const crt_oracle_camera* c, const crt_oracle_light* lights, int nl,
                       int min_samples, int max_samples, double noise_threshold, int num_bounces, uint64_t seed,
                       uint64_t key_offset, float* rgb, int32_t* spp_used) {
    /* the reference's self-intersection offset, src/crt/path_tracing/unidirectional.hpp:61 */
    crt_oracle_render_ex(s, c, lights, nl, min_samples, max_samples, noise_threshold, num_bounces, seed, key_offset,
                         1e-4, 0.0, rgb, spp_used, NULL);
}

/* ------------------------------------------------------ radiation pressure */
/* New (no reference counterpart).  Parallel rays on an N x N grid over the
 * square circumscribing the bounding disk perpendicular to the sun; photon
 * direction dph = -sun.  Per hit (flat outward normal n = -normalize(tri.n),
 * cos = n.sun):  dF = P A_ray [ (1-rs) dph - 2 (rs cos + rd/3) n ],
 * dT = (x - ref) x dF.  Rays hitting a back face (cos <= 0) contribute 0. */
void crt_oracle_radiation_pressure(const orc_scene* s, const double sun[3], double pressure, double rho_s,
                                   double rho_d, int64_t grid, const double center[3], double radius,
                                   const double ref[3], double force[3], double torque[3], int64_t* n_hit) {
    const v3 S = nrm(mk(sun[0], sun[1], sun[2]));
    const v3 dph = scl(S, -1.0);
    const v3 a = fabs(S.x) < 0.9 ? mk(1, 0, 0) : mk(0, 1, 0);
    const v3 e1 = nrm(cross(S, a));
    const v3 e2 = cross(S, e1);
    const double hcell = 2.0 * radius / (double)grid;
    const double area = hcell * hcell;
    const v3 C = mk(center[0], center[1], center[2]);
    const v3 Rf = mk(ref[0], ref[1], ref[2]);
    double F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
    int64_t hits = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : F[:3], T[:3], hits)
    for (int64_t j = 0; j < grid; ++j)
        for (int64_t i = 0; i < grid; ++i) {
            const double x = -radius + (i + 0.5) * hcell, y = -radius + (j + 0.5) * hcell;
            const v3 O = add(add(add(C, scl(S, 2.0 * radius)), scl(e1, x)), scl(e2, y));
            double tt, uu, vv;
            const int64_t h = trace(s, O, dph, 0.0, INFINITY, 0, &tt, &uu, &vv);
            if (h < 0) continue;
            const v3 n = scl(nrm(s->n[h]), -1.0);
            const double cs = dot(n, S);
            if (cs <= 0.0) continue;
            const v3 p = hit_point(s, h, uu, vv);
            const v3 dF = scl(sub(scl(dph, 1.0 - rho_s), scl(n, 2.0 * (rho_s * cs + rho_d / 3.0))), pressure * area);
            const v3 dT = cross(sub(p, Rf), dF);
            F[0] += dF.x; F[1] += dF.y; F[2] += dF.z;
            T[0] += dT.x; T[1] += dT.y; T[2] += dT.z;
            hits += 1;
        }
    for (int k = 0; k < 3; ++k) { force[k] = F[k]; torque[k] = T[k]; }
    if (n_hit) *n_hit = hits;
}

/* ------------------------------------------------------ OBJ vertex normals */
/* lib/model_loaders/obj.hpp:90-114: each vertex accumulates the (unnormalised,
 * inward) normal n = (p0-p1)x(p2-p0) of every triangle using it, then is
 * normalised. */
void crt_oracle_vertex_normals(const float* vertices, int64_t nv, const uint32_t* indices, int64_t nt, float* out) {
    double* acc = (double*)calloc((size_t)nv * 3, sizeof(double));
    for (int64_t t = 0; t < nt; ++t) {
        const uint32_t a = indices[3 * t], b = indices[3 * t + 1], c = indices[3 * t + 2];
        const v3 p0 = mk(vertices[3 * a], vertices[3 * a + 1], vertices[3 * a + 2]);
        const v3 p1 = mk(vertices[3 * b], vertices[3 * b + 1], vertices[3 * b + 2]);
        const v3 p2 = mk(vertices[3 * c], vertices[3 * c + 1], vertices[3 * c + 2]);
        const v3 n = cross(sub(p0, p1), sub(p2, p0));
        const uint32_t id[3] = {a, b, c};
        for (int k = 0; k < 3; ++k) { acc[3 * id[k]] += n.x; acc[3 * id[k] + 1] += n.y; acc[3 * id[k] + 2] += n.z; }
    }
    for (int64_t i = 0; i < nv; ++i) {
        const v3 n = nrm(mk(acc[3 * i], acc[3 * i + 1], acc[3 * i + 2]));
        out[3 * i] = (float)n.x; out[3 * i + 1] = (float)n.y; out[3 * i + 2] = (float)n.z;
    }
    free(acc);
}

int crt_oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
