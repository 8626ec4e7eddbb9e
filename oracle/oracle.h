/*
 * oracle.h -- CPU oracle for the UPBP renderer hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the checker
 * or the CPU baseline; the product (libupbp.so) never links or calls it.
 *
 * A scalar FP32 restatement of the reference's per-path light and camera
 * tracing (nolmoonen/gpuupbp gpuupbp/kernel/trace_light.cu, trace_camera.cu and
 * the headers they include), with its own CPU BVH instead of OptiX, threaded
 * with std::thread. Each function cites the reference lines it follows.
 *
 * Parity pins (see tests/): range-query predicates against hand cases and
 * brute force; BSDF / phase / light sampling against closed forms; camera
 * mapping against its inverse; the renderer against an analytic direct-view
 * case; every constant-time MIS weight against the FP64 full-path balance
 * heuristic of oracle_mis.cpp (oracle_mis_check).
 */
#ifndef UPBP_ORACLE_H
#define UPBP_ORACLE_H

#include "../include/upbp_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_stats {
    uint64_t light_vertices, surface_vertices, medium_vertices, beams;
    uint64_t path_errors;
    double seconds;
} oracle_stats;

/* Renders iterations [first_iteration, first_iteration + n_iterations) and writes the
 * mean image (W*H*3, FP64) into out_rgb. fb_first (optional, W*H*3 floats) receives
 * the FP32 framebuffer of the first iteration alone. n_threads <= 0: all cores. */
int oracle_render(const upbp_scene* scene, const upbp_config* cfg, int first_iteration, int n_iterations,
                  int n_threads, double* out_rgb, float* fb_first, oracle_stats* stats);

/* Renders light paths [0, n_light_paths) and camera paths [0, n_camera_paths) of one
 * iteration (bounded CPU-baseline sample). times[0..2] receive the wall seconds of the light
 * pass, the photon-map build and the camera pass; stats (optional) the light pass's counts. */
int oracle_render_sample(const upbp_scene* scene, const upbp_config* cfg, int iteration, uint32_t n_camera_paths,
                         uint32_t n_light_paths, int n_threads, double* times, oracle_stats* stats);

/* Brute-force range query (same kinds and layout as upbp_test_range_query). */
int oracle_range_query(int kind, const float* prims, uint32_t n_prims, const float* queries, uint32_t n_queries,
                       float radius, uint32_t* offsets, uint32_t* indices, uint64_t max_indices,
                       uint64_t* n_indices);

/* Brute-force closest hit over triangles (9 floats each); rays: 8 floats each. */
int oracle_closest_hit(const float* verts, uint32_t n_tris, const float* rays, uint32_t n_rays, float* hit_t,
                       int32_t* hit_prim);

/* Analytic sphere test of every (ray, sphere) pair, same layout as upbp_test_sphere_hits. */
int oracle_sphere_hits(const float* spheres, uint32_t n_sph, const float* rays, uint32_t n_rays, float* out);

/* FP64 network of the learned proposal (oracle_mlp.cpp): parameter count, input encoding,
 * forward pass (samples n x 6: position in [0,1]^3, arrival direction -> n x 40 raw outputs)
 * and the weighted mixture NLL with its analytic gradient (samples n x 10: position, arrival
 * direction, scattered direction, weight). */
uint32_t oracle_mlp_param_count(void);
void oracle_mlp_encode(const float* x, const float* d, float* out);
int oracle_mlp_forward(const double* params, const float* samples, uint32_t n, double* out);
int oracle_mlp_loss_grad(const double* params, const float* samples, uint32_t n, double* loss, double* grads);

/* One learned-proposal record (8 x (mu.xyz, kappa, weight)) times the phase lobes of a path
 * arriving along `arrive` in a medium of mean cosine g (surface 0) or the cosine lobe about the
 * normal `arrive` (surface 1), sampled with u (3 per sample: lobe, two for the direction): out
 * 4 per sample (direction, product pdf). */
int oracle_guide_samples(const float* lobes, const float* arrive, float g, int32_t surface, const float* u, uint32_t n,
                         float* out);

/* Learned proposal for every later oracle render (same arguments as upbp_set_guide_lobes;
 * alpha 0 switches it off). */
int oracle_set_guide_lobes(const float* lobes, int32_t res, const float* lo, const float* hi, float alpha);

/* Environment-map sampling and lookup, same layout as upbp_test_env_samples. */
int oracle_env_samples(const upbp_env_map* env, const float* u, uint32_t n, float* out);

/* Camera mapping checks: raster -> primary ray direction and world -> raster, using the
 * reference's matrix construction (gpuupbp/host/camera.hpp:33-76). */
int oracle_camera_ray(const upbp_scene* scene, const upbp_config* cfg, const float* raster_xy, uint32_t n,
                      float* dirs);
int oracle_world_to_raster(const upbp_scene* scene, const upbp_config* cfg, const float* pts, uint32_t n,
                           float* raster_xy);

/* Leaf-function probes for closed-form tests. */
float oracle_hg_pdf(float mean_cosine, float cos_theta);
void oracle_hg_sample(float mean_cosine, float u0, float u1, float* cos_theta, float* pdf);
float oracle_fresnel(float cos_inc, float ior);
/* Philox4x32-10 block (Salmon et al. SC'11) used by both oracle and product streams. */
void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Renders one iteration with the MIS log on for the given camera paths and the first
 * n_lt_paths light paths; writes 6 doubles per weighted contribution: technique (see
 * oracle_mis.h), s, k, the constant-time weight, the FP64 balance-heuristic weight and the
 * FP64 weight under the reference's conventions. */
int oracle_mis_check(const upbp_scene* scene, const upbp_config* cfg, int iteration, const uint32_t* pixels,
                     uint32_t n_pixels, uint32_t n_lt_paths, int n_threads, double* out, uint32_t max_records,
                     uint32_t* n_records);

#ifdef __cplusplus
}
#endif

#endif
