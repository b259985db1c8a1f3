/* Drives every entry point of the FP64 oracle (test infrastructure) on a
 * small scene, for a build with AddressSanitizer and UndefinedBehavior-
 * Sanitizer: a cube on a floor, Lambertian / textured Blinn-Phong / mirror
 * materials, point, area and directional lights, camera passes, ray queries
 * with classification, LiDAR, renders and the flux integral.  Exit status 0
 * and a clean sanitizer report are the test. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "crt_oracle.h"

static void camera(crt_oracle_camera* c, int w, int h) {
    memset(c, 0, sizeof(*c));
    c->position[0] = 0.3; c->position[1] = -0.2; c->position[2] = 6.0;
    c->rotation[0] = 1.0; c->rotation[4] = 1.0; c->rotation[8] = 1.0;
    c->focal_length = 1.0;
    c->resolution[0] = w; c->resolution[1] = h;
    c->sensor_size[0] = 1.0; c->sensor_size[1] = 1.0;
}

int main(void) {
    /* cube [-1,1]^3 (8 vertices, 12 triangles) + floor quad at z = -1.5 */
    const float v[12 * 3] = {-1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, -1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1,
                             -4, -4, -1.5f, 4, -4, -1.5f, 4, 4, -1.5f, -4, 4, -1.5f};
    const uint32_t f[14 * 3] = {0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4,
                                1, 2, 6, 1, 6, 5, 2, 3, 7, 2, 7, 6, 3, 0, 4, 3, 4, 7, 8, 9, 10, 8, 10, 11};
    const int64_t nv = 12, nt = 14;
    uint32_t inst[14];
    for (int t = 0; t < nt; ++t) inst[t] = t < 12 ? 1u : 2u;
    float vn[12 * 3];
    crt_oracle_vertex_normals(v, nv, f, nt, vn);
    const float colors[6] = {0.8f, 0.7f, 0.6f, 0.5f, 0.5f, 0.5f};
    const uint8_t smooth[2] = {1, 0};
    orc_scene* s = crt_oracle_scene_create(v, nv, f, nt, inst, vn, 2, colors, smooth);
    if (!s) return 2;

    /* rays: camera rays of a 24x16 image */
    crt_oracle_camera cam;
    camera(&cam, 24, 16);
    const int64_t n = 24 * 16;
    double* o = malloc(sizeof(double) * 3 * n);
    double* d = malloc(sizeof(double) * 3 * n);
    crt_oracle_camera_rays(&cam, o, d);
    double* t = malloc(sizeof(double) * n);
    double* u = malloc(sizeof(double) * n);
    double* w = malloc(sizeof(double) * n);
    int32_t* tri = malloc(sizeof(int32_t) * n);
    uint8_t* flags = malloc(n);
    crt_oracle_trace(s, n, o, d, NULL, NULL, 0, t, tri, u, w);
    crt_oracle_trace(s, n, o, d, NULL, NULL, 1, t, tri, u, w);
    crt_oracle_tie_flags(s, n, o, d, 1e-6, 1e-5, flags);
    crt_oracle_edge_flags(s, n, o, d, tri, 1e-6, 1e-5, flags);
    crt_oracle_closest_fragile(s, n, o, d, 1e-6, flags);
    crt_oracle_shadow_fragile(s, n, o, d, NULL, 1e-6, flags);
    double out3[3];
    const double p0[3] = {0, 0, 0}, p1[3] = {1, 0, 0}, p2[3] = {0, 1, 0}, ro[3] = {0.2, 0.2, 1}, rd[3] = {0, 0, -1};
    if (!crt_oracle_tri_intersect(p0, p1, p2, ro, rd, 0.0, 1e30, out3)) return 3;
    (void)crt_oracle_closest_brute(s, ro, rd, 0.0, 1e30, out3);

    /* passes */
    double* xyz = malloc(sizeof(double) * 3 * n);
    double* nrm = malloc(sizeof(double) * 3 * n);
    double* depth = malloc(sizeof(double) * n);
    uint32_t* ids = malloc(sizeof(uint32_t) * n);
    crt_oracle_pass(s, &cam, xyz, nrm, ids, depth, tri);

    /* LiDAR: 32 poses looking down */
    double pos[32 * 3], rot[32 * 9];
    for (int k = 0; k < 32; ++k) {
        pos[3 * k] = -2.0 + 0.125 * k; pos[3 * k + 1] = 0.1; pos[3 * k + 2] = 5.0;
        memset(rot + 9 * k, 0, 9 * sizeof(double));
        rot[9 * k] = 1.0; rot[9 * k + 4] = 1.0; rot[9 * k + 8] = 1.0;
    }
    double range[32];
    crt_oracle_lidar_batch(s, 32, pos, rot, 0, range);

    /* renders: three light types, then materials with a 2x2 texture */
    crt_oracle_light L[3];
    memset(L, 0, sizeof(L));
    L[0].type = 0; L[0].intensity = 50.0; L[0].position[2] = 5.0;
    L[1].type = 1; L[1].intensity = 20.0; L[1].position[2] = 4.0; L[1].size[0] = 1.0; L[1].size[1] = 1.0;
    L[1].rotation[0] = 1.0; L[1].rotation[4] = 1.0; L[1].rotation[8] = 1.0;
    L[2].type = 2; L[2].intensity = 1.0; L[2].direction[0] = 0.3; L[2].direction[2] = 1.0;
    float* rgb = malloc(sizeof(float) * 3 * n);
    int32_t* spp = malloc(sizeof(int32_t) * n);
    int32_t* frag = malloc(sizeof(int32_t) * 3 * n);
    crt_oracle_render(s, &cam, L, 3, 2, 4, 0.05, 3, 7, 0, rgb, spp);
    const uint8_t texels[16] = {255, 0, 0, 255, 0, 255, 0, 255, 0, 0, 255, 255, 255, 255, 255, 255};
    const crt_oracle_texture tex = {2, 2, texels};
    crt_oracle_material mats[2];
    memset(mats, 0, sizeof(mats));
    mats[0].type = 2; mats[0].texture = 0; mats[0].spec_texture = -1; mats[0].alpha = 16.0f;
    mats[0].coeffs[0] = 0.1f; mats[0].coeffs[1] = 0.7f; mats[0].coeffs[2] = 0.3f;
    mats[1].type = 3; mats[1].texture = -1; mats[1].spec_texture = -1;
    float uv[12 * 2];
    for (int k = 0; k < 12; ++k) { uv[2 * k] = (float)(k % 3) / 2.0f; uv[2 * k + 1] = (float)(k % 2); }
    crt_oracle_scene_set_materials(s, 2, mats, uv, nv, 1, &tex);
    crt_oracle_render_ex(s, &cam, L, 3, 2, 2, 0.0, 3, 9, 0, 1e-4, 1e-6, rgb, spp, frag);

    /* flux integral */
    const double sun[3] = {0.2, 0.1, 1.0}, center[3] = {0, 0, 0}, ref[3] = {0.1, 0, 0};
    double force[3], torque[3];
    int64_t nhit = 0;
    crt_oracle_radiation_pressure(s, sun, 4.56e-6, 0.3, 0.6, 64, center, 6.0, ref, force, torque, &nhit);
    if (nhit <= 0) return 4;

    /* the same mesh in FP64 */
    double v64[12 * 3];
    for (int k = 0; k < 36; ++k) v64[k] = v[k];
    orc_scene* s64 = crt_oracle_scene_create_f64(v64, nv, f, nt, inst, NULL, 2, colors, smooth);
    crt_oracle_trace(s64, n, o, d, NULL, NULL, 0, t, tri, u, w);
    crt_oracle_scene_destroy(s64);
    crt_oracle_scene_destroy(s);
    free(o); free(d); free(t); free(u); free(w); free(tri); free(flags); free(xyz); free(nrm); free(depth);
    free(ids); free(rgb); free(spp); free(frag);
    printf("oracle driver ok: %lld flux hits, u01 %.6f\n", (long long)nhit, crt_oracle_u01(1, 2, 3, 4));
    return 0;
}
