/* crt_oracle.h -- FP64 CPU restatement of the ceres-raytracer hot path.
 * TEST INFRASTRUCTURE ONLY (the checker): used by tests/, the smoke check and
 * the CPU-baseline leg of bench.py; never by the product path. */
#ifndef CRT_ORACLE_H
#define CRT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

typedef struct {
    double position[3];
    double rotation[9];      /* row-major R, rays rotated by R^T */
    double focal_length;
    double resolution[2];
    double sensor_size[2];
    int32_t z_positive;
    int32_t pad_;
} crt_oracle_camera;

typedef struct {
    int32_t type;            /* 0 point, 1 area, 2 directional */
    int32_t pad_;
    double intensity;
    double position[3];
    double rotation[9];
    double size[2];
    double direction[3];
} crt_oracle_light;

/* Per-instance material, src/crt/materials/material.hpp:90-184:
 * type 0 ColoredLambertianMaterial(color), 1 TexturedLambertianMaterial(map),
 * 2 TexturedBlinnPhongMaterial(map, specular map, alpha), 3 MirrorMaterial.
 * A map is a texture index, or -1 for a constant map (color / coeffs =
 * ka, kd, ks).  Same layout as crt_material of the HIP library. */
typedef struct {
    uint32_t type;
    int32_t texture;
    int32_t spec_texture;
    float alpha;
    float color[3];
    float coeffs[3];
} crt_oracle_material;

typedef struct {
    uint32_t width, height;
    const uint8_t* rgba;      /* width x height RGBA8, row-major, row 0 first */
} crt_oracle_texture;

orc_scene* crt_oracle_scene_create(const float* vertices, int64_t nv, const uint32_t* indices, int64_t nt,
                                   const uint32_t* tri_inst, const float* vnormals, int64_t ni,
                                   const float* colors, const uint8_t* smooth);
orc_scene* crt_oracle_scene_create_f64(const double* vertices, int64_t nv, const uint32_t* indices, int64_t nt,
                                       const uint32_t* tri_inst, const float* vnormals, int64_t ni,
                                       const float* colors, const uint8_t* smooth);
void crt_oracle_scene_destroy(orc_scene* s);
/* Replace the per-instance materials (n = number of instances), the per-vertex
 * UVs (nv x 2 or NULL: all zero, as the reference OBJ loader leaves them) and
 * the textures (copied). */
void crt_oracle_scene_set_materials(orc_scene* s, int64_t n, const crt_oracle_material* mats, const float* uv,
                                    int64_t nv, int ntex, const crt_oracle_texture* tex);

int crt_oracle_tri_intersect(const double p0[3], const double p1[3], const double p2[3], const double o[3],
                             const double d[3], double tmin, double tmax, double out[3]);
int64_t crt_oracle_closest_brute(const orc_scene* s, const double o[3], const double d[3], double tmin, double tmax,
                                 double out[3]);
void crt_oracle_trace(const orc_scene* s, int64_t n, const double* o, const double* d, const double* tmin,
                      const double* tmax, int any_hit, double* t, int32_t* tri, double* u, double* v);
void crt_oracle_tie_flags(const orc_scene* s, int64_t n, const double* o, const double* d, double rel_t,
                          double bary_eps, uint8_t* flags);
void crt_oracle_edge_flags(const orc_scene* s, int64_t n, const double* o, const double* d, const int32_t* other,
                           double rel_t, double bary_eps, uint8_t* flags);
/* Per ray (directions normalised here): 1 where the closest-hit decision is
 * fragile at distance eps (hit within eps / |cos| of an edge of its triangle,
 * or triangles grown by eps give another closest hit), resp. where a shadow
 * ray over [0, tmax] (NULL: infinite) is fragile (occlusion differs between
 * triangles grown and shrunk by eps). */
void crt_oracle_closest_fragile(const orc_scene* s, int64_t n, const double* o, const double* d, double eps,
                                uint8_t* flags);
void crt_oracle_shadow_fragile(const orc_scene* s, int64_t n, const double* o, const double* d, const double* tmax,
                               double eps, uint8_t* flags);
void crt_oracle_camera_rays(const crt_oracle_camera* c, double* o, double* d);
void crt_oracle_pass(const orc_scene* s, const crt_oracle_camera* c, double* xyz, double* normal, uint32_t* instance,
                     double* depth, int32_t* tri);
void crt_oracle_lidar_batch(const orc_scene* s, int64_t n, const double* pos, const double* rot, int z_positive,
                            double* range);
double crt_oracle_u01(uint64_t seed, uint64_t a, uint64_t b, uint64_t c);
void crt_oracle_render(const orc_scene* s, const crt_oracle_camera* c, const crt_oracle_light* lights, int nl,
                       int min_samples, int max_samples, double noise_threshold, int num_bounces, uint64_t seed,
                       uint64_t key_offset, float* rgb, int32_t* spp_used);
/* crt_oracle_render with the self-intersection offset as a parameter (the
 * reference's is 1e-4, src/crt/path_tracing/unidirectional.hpp:61) and, when
 * `fragile` is not NULL, per pixel three counts over all its paths: [0]
 * closest hits decided within `eps` (distance) of a triangle edge or of a
 * closer triangle, [1] shadow rays whose occlusion changes when every
 * triangle is grown or shrunk by `eps` (or the blocker moved by eps along
 * the ray), [2] texture lookups within 2e-3 texel of a texel boundary. */
void crt_oracle_render_ex(const orc_scene* s, const crt_oracle_camera* c, const crt_oracle_light* lights, int nl,
                          int min_samples, int max_samples, double noise_threshold, int num_bounces, uint64_t seed,
                          uint64_t key_offset, double offset, double eps, float* rgb, int32_t* spp_used,
                          int32_t* fragile);
void crt_oracle_radiation_pressure(const orc_scene* s, const double sun[3], double pressure, double rho_s,
                                   double rho_d, int64_t grid, const double center[3], double radius,
                                   const double ref[3], double force[3], double torque[3], int64_t* n_hit);
void crt_oracle_vertex_normals(const float* vertices, int64_t nv, const uint32_t* indices, int64_t nt, float* out);
int crt_oracle_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
