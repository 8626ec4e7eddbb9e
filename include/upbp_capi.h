/*
 * upbp_capi.h -- C ABI of the MI355X-native UPBP renderer (libupbp.so).
 *
 * Plain pointers, sizes and POD structs only; no C++ or torch types cross this
 * boundary. Every entry point returns 0 on success and a negative UPBP_E_* code
 * on failure (it never calls exit()); upbp_last_error() then holds a message.
 *
 * Each declaration names the reference interface it replaces
 * (nolmoonen/gpuupbp, paths relative to the reference root).
 */
#ifndef UPBP_CAPI_H
#define UPBP_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define UPBP_OK 0
#define UPBP_E_ARG (-1)      /* invalid argument / command line */
#define UPBP_E_IO (-2)       /* file could not be read or written */
#define UPBP_E_PARSE (-3)    /* malformed OBJ / MTL / AUX content */
#define UPBP_E_HIP (-4)      /* HIP runtime failure */
#define UPBP_E_NODEVICE (-5) /* no GPU visible */
#define UPBP_E_CAPACITY (-6) /* a fixed-size buffer would overflow */
#define UPBP_HELP 1          /* parse_commandline printed help; nothing to render */

/* ---- technique / mode bits: same values as gpuupbp/shared/shared_enums.h:30-56 */
#define UPBP_BEAM_SHORT (1u << 0)
#define UPBP_BEAM_LONG (1u << 1)
#define UPBP_BPT (1u << 2)
#define UPBP_SURF (1u << 3)
#define UPBP_PP3D (1u << 4)
#define UPBP_PB2D (1u << 5)
#define UPBP_BB1D (1u << 6)
#define UPBP_BP2D (1u << 7)
#define UPBP_PREVIOUS (1u << 9)
#define UPBP_COMPATIBLE (1u << 10)
#define UPBP_SPECULAR_ONLY (1u << 12)

/* light types: gpuupbp/shared/light.h:38 */
#define UPBP_LIGHT_AREA 0
#define UPBP_LIGHT_BACKGROUND 1
#define UPBP_LIGHT_DIRECTIONAL 2
#define UPBP_LIGHT_POINT 3

/* ---- configuration: mirrors struct Config (gpuupbp/misc/config.hpp:44-120) */
typedef struct upbp_config {
    uint32_t algorithm_flags;
    int32_t iteration_count;
    float max_time;
    uint32_t res_x, res_y;
    uint32_t max_path_length;
    uint32_t min_path_length;
    float path_count_per_iter; /* resolved to an absolute count after parsing */
    uint32_t query_beam_type;  /* UPBP_BEAM_SHORT or UPBP_BEAM_LONG */
    uint32_t photon_beam_type;
    float bb1d_radius_initial, bb1d_radius_alpha;
    float bb1d_used_light_subpath_count; /* negative = relative to light paths */
    float bp2d_radius_initial, bp2d_radius_alpha;
    float pb2d_radius_initial, pb2d_radius_alpha;
    float pp3d_radius_initial, pp3d_radius_alpha;
    float surf_radius_initial, surf_radius_alpha; /* negative radius = relative to scene radius */
    int32_t continuous_output;
    int32_t ignore_fully_spec_paths;
    int32_t gpu_assert;
    uint64_t iteration_offset;
    int32_t do_log;
    int32_t use_shading_normal;
    int32_t scene_id;
    char scene_obj_file[1024];
    char output_name[1024];
    char additional_args[512];
    /* Learned scattering proposal (no counterpart in the reference; 0 = off, the default):
     * -guide <alpha> at non-specular surfaces, -guide_media <alpha> in media. The CLI then
     * calls upbp_set_guide with the library's other defaults. */
    float guide_surface_alpha;
    float guide_media_alpha;
} upbp_config;

/* ---- flattened scene: replaces struct Scene + Record (gpuupbp/shared/scene.h:41-127) */
typedef struct upbp_material {
    float diffuse[3];
    float phong[3];
    float phong_exponent;
    float mirror[3];
    float ior;        /* < 0: no Fresnel, no transmission */
    int32_t priority; /* -1 thin wall, -2 global medium */
    int32_t real;     /* 1 real geometry, 0 imaginary (medium boundary only) */
    int32_t med_idx;  /* -1 if no medium is associated */
} upbp_material;

typedef struct upbp_medium {
    float absorption[3];
    float emission[3];
    float scattering[3];
    float continuation_prob;
    float mean_cosine;
} upbp_medium;

typedef struct upbp_light {
    int32_t type;   /* UPBP_LIGHT_* */
    int32_t mat_id; /* enclosing material, -1 = global medium */
    int32_t med_id;
    float p0[3], p1[3], p2[3]; /* area light triangle */
    float position[3];         /* point light position / directional light direction */
    float intensity[3];        /* area/point/directional emission, background colour */
    float scale;               /* background scale */
    int32_t uses_env_map;      /* background: radiance from upbp_scene.env_map instead of the colour */
} upbp_light;

typedef struct upbp_camera {
    float origin[3], target[3], roll[3];
    float horizontal_fov_deg;
    float focal_dist;
    int32_t mat_id, med_id;
} upbp_camera;

typedef struct upbp_triangles {
    uint32_t count;
    const float* verts;    /* 9 floats per triangle: p0, p1, p2 */
    const float* normals;  /* 12 floats per triangle: n0, n1, n2, face normal */
    const int32_t* mat_id; /* per triangle */
    const int32_t* light_id; /* per triangle, -1 if not emissive */
} upbp_triangles;

/* Analytic spheres (predefined scenes only, gpuupbp/shared/scene.h Spheres, intersected as in
 * gpuupbp/kernel/intersector.cu:35-101). */
typedef struct upbp_spheres {
    uint32_t count;
    const float* spheres;    /* 4 floats per sphere: centre x, y, z, radius */
    const int32_t* mat_id;   /* per sphere */
    const int32_t* light_id; /* per sphere, -1 (spheres are never emissive in the reference) */
} upbp_spheres;

/* Latitude-longitude environment map of the background light (gpuupbp/shared/env_map.h Image):
 * RGB floats, rows from the top, already rotated and scaled as the AUX line asks
 * (gpuupbp/misc/tinyexr_wrapper.cpp:97-136). The luminance distribution is built from it. */
typedef struct upbp_env_map {
    uint32_t width, height; /* 0 when the scene has none */
    const float* rgb;
} upbp_env_map;

typedef struct upbp_scene {
    upbp_triangles real;
    upbp_triangles imag;
    uint32_t mat_count;
    const upbp_material* materials;
    uint32_t med_count;
    const upbp_medium* media;
    uint32_t light_count;
    const upbp_light* lights;
    uint32_t global_medium_idx;
    int32_t background_light_idx; /* -1 if none */
    upbp_camera camera;
    float scene_center[3];
    float scene_radius;
    upbp_spheres sph_real; /* zero-initialised (count 0) for OBJ scenes */
    upbp_spheres sph_imag;
    upbp_env_map env_map;
} upbp_scene;

/* Per-iteration statistics (device counters plus HIP-event times). */
typedef struct upbp_iter_stats {
    double ms_light, ms_build, ms_camera, ms_total;
    double ms_bounce;              /* camera bounce kernels (intersector, BPT, point merges) */
    double ms_merge_pb2d;          /* P-B2D segment merge kernels */
    double ms_merge_bb1d;          /* B-B1D segment merge kernels */
    double ms_dominant;            /* summed duration of the dominant merge or bounce stage in this
                                      iteration (the B-B1D stage: its bundle and per-piece kernels) */
    uint64_t dominant_launches;    /* its stage count (one per camera bounce) */
    char dominant_kernel[32];      /* its name */
    uint64_t light_vertices, surface_vertices, medium_vertices, beams, subbeams;
    uint64_t light_extend_steps, camera_extend_steps;
    uint64_t bpt_connections, point_queries, segment_queries;
    uint64_t merge_accepts, pb2d_accepts, bb1d_accepts;
    uint64_t seg_nodes, seg_candidates; /* LBVH nodes visited / leaves tested by segment queries */
    uint64_t seg_pieces;           /* pieces the segment queries were walked in */
    uint64_t seg_wave_trips;       /* merge kernels: wave-level traversal loop trips */
    uint64_t seg_cand_batches, seg_pair_batches; /* wave-level candidate-test / evaluation batches */
    uint64_t vertex_drops, beam_drops, path_errors;
    double algorithmic_bytes;      /* compulsory HBM bytes of this iteration */
    double dominant_bytes;         /* compulsory HBM bytes of the dominant kernel */
    uint64_t guide_samples;        /* learned proposal: training samples collected */
    double guide_loss;             /* mean training loss of the iteration's Adam steps */
    double ms_guide;               /* training and record refresh time */
    uint64_t merge_stack_overflows; /* segment-merge query pieces abandoned on a full traversal
                                       stack (also counted in path_errors) */
    uint64_t photon_list_resizes;   /* sizing passes of the photon lists so far (the first
                                       iteration, then one per iteration that outgrew them and
                                       was rendered again) */
    uint64_t graph_launches;        /* iterations replayed from a captured graph so far */
    uint64_t bundles;               /* B-B1D piece bundles walked as one (upbp_tuning.bundle_bounces) */
    uint64_t bundle_pieces_left;    /* pieces of wider bundles, walked per piece */
    uint64_t bundle_nodes;          /* beam-tree nodes fetched by the bundle walks */
    uint64_t bundle_leaves;         /* sub-beams staged by the bundle walks (once per bundle) */
    uint64_t bundle_tests;          /* (piece, sub-beam) exact tests of the bundle walks (leaf box hit) */
} upbp_iter_stats;

typedef struct upbp_scene_handle upbp_scene_handle;
typedef struct upbp_renderer upbp_renderer;

/* Replaces parse_commandline (gpuupbp/misc/config.cpp:735-1370): same options,
 * defaults, messages and output-name rules. Returns 0, UPBP_HELP, or UPBP_E_ARG. */
int upbp_parse_commandline(int argc, const char** argv, upbp_config* cfg);

/* Fills cfg with the parser defaults (config.cpp:738-785). */
void upbp_config_defaults(upbp_config* cfg);

/* Replaces SceneLoader::load_from_obj + load_scene (gpuupbp/misc/scene_loader_obj.cpp:926-1104,
 * gpuupbp/misc/scene_loader.hpp:44-76): reads file.obj, its mtllib and file.obj.aux. */
int upbp_scene_load_obj(const char* obj_path, uint32_t res_x, uint32_t res_y, upbp_scene_handle** out);

/* Replaces SceneLoader::load_predefined + load_scene (gpuupbp/misc/scene_loader_predef.cpp:126-603,
 * scene_loader.hpp:44-76) over the scene table of init_scene_configs (gpuupbp/misc/config.cpp:100-
 * 518): Cornell-box variants with analytic spheres, ids 0 .. upbp_predefined_scene_count() - 1.
 * UPBP_E_ARG for an id outside the table. */
int upbp_scene_load_predefined(int32_t scene_id, uint32_t res_x, uint32_t res_y, upbp_scene_handle** out);
int32_t upbp_predefined_scene_count(void);
/* Short (file-name) or long (help-text) name of a predefined scene; NULL outside the table. */
const char* upbp_predefined_scene_name(int32_t scene_id, int32_t long_name);
const upbp_scene* upbp_scene_get(const upbp_scene_handle* h);
void upbp_scene_free(upbp_scene_handle* h);

/* Replaces tinyexr_wrapper::save_exr (gpuupbp/misc/tinyexr_wrapper.cpp:34-95): scanline,
 * uncompressed, HALF, channels B,G,R. rgb is W*H*3 floats, row-major from the top row. */
int upbp_save_exr(const char* path, const float* rgb, uint32_t width, uint32_t height);

/* Replaces tinyexr's LoadEXR as tinyexr_wrapper::load_image calls it (gpuupbp/misc/
 * tinyexr_wrapper.cpp:97-136): single-part scanline OpenEXR with NONE, RLE, ZIPS or ZIP
 * compression and HALF / FLOAT / UINT channels; R, G, B (or one grey channel) into rgb as
 * width * height * 3 floats from the top row. rgb may be NULL to query the size;
 * UPBP_E_CAPACITY if capacity_floats is too small. */
int upbp_read_exr(const char* path, float* rgb, uint64_t capacity_floats, uint32_t* width, uint32_t* height);

/* Replaces save_bmp (gpuupbp/host/framebuffer.cpp:129-200): 24-bit uncompressed BMP, rows
 * bottom-up, B,G,R, gamma-corrected and clamped (rows padded to 4 bytes). */
int upbp_save_bmp(const char* path, const float* rgb, uint32_t width, uint32_t height, float gamma);

/* Replaces save_hdr (framebuffer.cpp:202-233): Radiance RGBE, flat scanlines. */
int upbp_save_hdr(const char* path, const float* rgb, uint32_t width, uint32_t height);

/* Replaces save (framebuffer.cpp:114-127): picks EXR, BMP (gamma 2.2) or HDR by the
 * extension; UPBP_E_ARG for any other. */
int upbp_save_image(const char* path, const float* rgb, uint32_t width, uint32_t height);

/* Replaces Renderer::init (gpuupbp/renderer/renderer_init.cpp:675-929). Copies the scene to the
 * GPU, builds the triangle LBVHs and sizes the SoA vertex/beam caches. The caller keeps
 * ownership of scene and cfg. */
int upbp_create(const upbp_scene* scene, const upbp_config* cfg, upbp_renderer** out);

/* Replaces Renderer::run_iteration (gpuupbp/renderer/renderer.cpp:37-418): light pass, photon
 * LBVH builds, camera pass, FP64 accumulation on the device. stats may be NULL. */
int upbp_run_iteration(upbp_renderer* r, int iteration, upbp_iter_stats* stats);

/* Replaces the D2H copy + add_framebuffer + scale (renderer.cpp:381-392, gpuupbp.cpp:132):
 * writes the FP64 mean image (W*H*3) over iters_done iterations into a host buffer. */
int upbp_read_accum(upbp_renderer* r, double* rgb, int iters_done);

/* Replaces Renderer::cleanup (renderer.cpp:420-508). Idempotent; NULL is accepted. */
void upbp_destroy(upbp_renderer* r);

/* Replaces the stderr + exit(1) error path (gpuupbp/host/exception.hpp:44-123). */
const char* upbp_last_error(const upbp_renderer* r);
const char* upbp_global_error(void);

/* Writes a seeded procedural scene as an OBJ + MTL + AUX triple (benchmark and test inputs; no
 * reference counterpart). kind: "box", "boxfog" (denser anisotropic medium, glossy back wall),
 * "blobs" or "furnace". */
int upbp_generate_scene(const char* kind, uint32_t seed, const char* obj_path);

/* Blocks until all work queued by r has finished. */
int upbp_sync(upbp_renderer* r);

/* Device memory of the current GPU in bytes (replaces optix::report_memory,
 * gpuupbp/misc/optix_helper.cpp, used by the -log CSV). */
int upbp_memory_info(uint64_t* free_bytes, uint64_t* total_bytes);

/* Launch and data-structure knobs of one renderer (no reference counterpart: the OptiX
 * build had none). They change speed only, never which photons a query accepts. */
typedef struct upbp_tuning {
    float subbeam_factor;     /* beams are cut into pieces of about subbeam_factor x the beam
                                 radius for the beam LBVH; 0 keeps whole beams (default 12) */
    float query_piece_factor; /* camera-segment queries are walked in pieces of about this
                                 factor x the merge radius, at most 32 per ray plus one per
                                 segment; 0 keeps whole segments (default 64) */
    int32_t merge_blocks_per_cu; /* resident blocks per CU of the segment merge kernels; 0 = as
                                    many as fit (default) */
    int32_t piece_order;         /* order in which the B-B1D merge takes query pieces: 0 as the
                                    camera pass wrote them (default); 1 bucketed by region, each
                                    XCD starting on its own; 2 sorted by the Morton code of the
                                    piece midpoint (one host synchronisation per bounce); 3 units
                                    of 64 consecutive pieces kept whole, sorted by the Morton code
                                    of their middle piece and cut into one range per XCD; 4 the
                                    same sorted units as one range */
    int32_t packet_bounces;      /* camera bounces below this walk the B-B1D beam tree as packets
                                    of 64 query pieces per wave (one node fetch per wave); later
                                    bounces walk it per piece (default 0: per piece everywhere,
                                    the faster of the two on the blob grid) */
    int32_t block_size;          /* threads per block of the light and camera bounce kernels: 64,
                                    128 or 256 (0: 64, the default; one wave per block lets the
                                    sparse last blocks of a bounce finish sooner) */
    int32_t beam_leaf;           /* B-B1D leaf test: 0 reads the sorted 48-byte sub-beam records
                                    (default); 1 reads 8-byte sorted references (beam slot,
                                    sub-beam index) and the beam's own record, which all its
                                    sub-beams share (a smaller working set) */
    int32_t graph;               /* 1: capture the iteration's work in a HIP graph once and replay it
                                    (iterations without statistics, online training or Morton piece
                                    order, which read back to the host); 0: launch kernel by kernel */
    int32_t bundle_bounces;      /* camera bounces below this walk the B-B1D beam tree once per
                                    bundle: a camera wave's j-th query pieces (up to 64 neighbouring
                                    rays at similar depth), one node fetch per bundle, the sub-beams
                                    reached tested by every ray of the bundle from LDS; bundles
                                    wider than bundle_factor x the beam radius, and later bounces,
                                    walk per piece (default 1: the first bounce; needs piece_order
                                    0, beam_leaf 0 and packet_bounces 0, else off) */
    float bundle_factor;         /* largest bundle radius walked as one bundle, in beam radii
                                    (default 32) */
    float bundle_piece_factor;   /* query piece length on the bounces walked in bundles, in merge
                                    radii (default 0: query_piece_factor's) */
} upbp_tuning;
void upbp_tuning_defaults(upbp_tuning* t);
int upbp_set_tuning(upbp_renderer* r, const upbp_tuning* t);
/* Batched jobs (no reference counterpart): one iteration of each of n independent renderers
 * (iterations[i] for renderer i), driven by up to `threads` host threads at once. Each renderer
 * owns a HIP stream, so the jobs' kernels overlap on the GPU; a renderer must not appear twice
 * in one batch. Returns the first job's error, if any. */
int upbp_run_batch(upbp_renderer* const* rs, const int32_t* iterations, int32_t n, int32_t threads);

/* Batched jobs in one launch per stage (no reference counterpart; the reference renders one job
 * per process, renderer.cpp:122-195). A batch groups n independent renderers (each at most once);
 * upbp_batch_run renders one iteration of each (iterations[i] for renderer i) as a round in which
 * every wavefront stage -- light generate, each light bounce, camera generate, each camera bounce
 * and its P-B2D and B-B1D segment merges, splat, accumulate -- is one kernel launch over all jobs,
 * the persistent merge kernels walking every job's query pieces. Each job's parameters, counters,
 * photon lists and LBVH builds go on its own stream between those launches, overlapping the other
 * jobs'. With use_graph, the round is captured in a HIP graph after one eager round with the same
 * settings and replayed (captured again whenever a job's settings or buffers change). A round
 * whose jobs cannot all take part -- a job still sizing its photon lists, training its proposal,
 * logging MIS weights or tracing only one subpath -- renders each job alone instead, as does a job
 * whose photon lists overflowed in the round (its batched iteration is then not accumulated).
 * The images equal the renderers' own iterations, up to the order of floating-point atomic sums;
 * the merges walk the pieces in the camera pass's order whatever the piece_order tuning. A batch
 * and its renderers are used from one host thread at a time; upbp_batch_run returns when the
 * round is complete on the GPU. */
typedef struct upbp_batch upbp_batch;
int upbp_batch_create(upbp_renderer* const* rs, int32_t n, upbp_batch** out);
int upbp_batch_run(upbp_batch* b, const int32_t* iterations, int32_t use_graph);
/* rounds run, and of those: replayed graphs, batched rounds launched eagerly, rounds of jobs alone */
int upbp_batch_counts(const upbp_batch* b, uint64_t* rounds, uint64_t* graph_rounds, uint64_t* eager_rounds,
                      uint64_t* solo_rounds);
const char* upbp_batch_last_error(const upbp_batch* b);
void upbp_batch_destroy(upbp_batch* b); /* the renderers stay; destroy the batch before them */

/* Autotuner (no reference counterpart): a coordinate search over subbeam_factor {8, 12, 16,
 * 24}, query_piece_factor {32, 64, 128}, merge_blocks_per_cu {0, 1, 2}, piece_order {0, 2},
 * packet_bounces {0, 1, 2, 64}, block_size {64, 128, 256} and bundle_bounces {0, 1, 2},
 * from the current tuning, timing `repeats` (1..64) renders of `iteration` per candidate after
 * one untimed render. The fastest tuning stays set and is returned in *chosen with its time in
 * *ms_per_iteration (either may be NULL). The accumulated image is restored afterwards and
 * online training is paused meanwhile, so the renderer's output is as if it had not run. */
int upbp_autotune(upbp_renderer* r, int32_t iteration, int32_t repeats, upbp_tuning* chosen, float* ms_per_iteration);

/* ---- learned scattering proposal (no counterpart in the reference) -------
 * Camera paths in media sample, with probability alpha, the product of a learned mixture of
 * eight von Mises-Fisher lobes (incident radiance, looked up on a res^3 grid over [lo, hi]) and
 * three vMF lobes fitted to the medium's Henyey-Greenstein phase function (mean direction along
 * the arrival for g >= 0, against it for g < 0; the table of phase_lobes.h), else the phase
 * function. At surfaces without mirror or refraction components camera paths sample, with
 * probability surface_alpha, the product of the same mixture with the vMF lobe fitted to the
 * clamped cosine about the shading normal, else the BSDF. The proposal's pdf enters every
 * camera-technique pdf (forward at camera vertices, reverse at light vertices, connections
 * and merges), so the estimator stays unbiased and MIS consistent.
 * upbp_set_guide_lobes installs explicit lobes: res^3 records, record (z * res + y) * res + x,
 * of 8 lobes x (mu.xyz, kappa, weight); mu is normalised, kappa clamped to [1e-2, 1e4],
 * weights normalised; alpha applies in media and at surfaces. alpha 0 switches the proposal
 * off (the plain renderer). */
int upbp_set_guide_lobes(upbp_renderer* r, const float* lobes, int32_t res, const float* lo, const float* hi,
                         float alpha);

/* The learned proposal trained online: each scattering of a camera path in a medium or at a
 * guidable surface is credited with everything the path collects after it (emission,
 * next-event estimation, connections, point merges and the segment merges of its later
 * segments), weighted as an estimate of incident radiance over the sampling pdf; after each
 * iteration the network takes steps_per_iteration Adam steps on random batches of those
 * samples and the grid records are re-evaluated from it. The grid covers the bounding cube of
 * the scene sphere. The first iteration samples the phase function and BSDFs only (the
 * untrained network is nearly uniform); the alphas apply from the first refresh on. Both
 * alphas 0 switch the proposal off. */
typedef struct upbp_guide_config {
    float alpha;                 /* mixture probability in media (default 0: on the slab-room
                                    workload guided medium paths attract more merge work than
                                    they save; 0.7 lowers the error per iteration) */
    int32_t grid_res;            /* record grid cells per axis (default 64, at most 64) */
    int32_t train;               /* 1: train online (default); 0: keep the network as initialised */
    int32_t steps_per_iteration; /* Adam steps after each iteration (default 10) */
    int32_t batch;               /* samples per step (default 65536) */
    float learning_rate;         /* Adam step size (default 8e-3) */
    uint32_t seed;               /* He initialisation (default 3) */
    uint32_t max_samples;        /* training samples kept per iteration (default 16M) */
    float surface_alpha;         /* mixture probability at non-specular surfaces (default 0.65) */
    uint32_t history;            /* training samples kept across iterations in a ring that the
                                    batches draw from (default 4M); 0: this iteration's only */
} upbp_guide_config;
void upbp_guide_defaults(upbp_guide_config* g);
int upbp_set_guide(upbp_renderer* r, const upbp_guide_config* g);

/* The proposal's network (BF16 MFMA forward, backward and Adam with FP32 accumulation and an
 * FP32 master copy): 51 inputs (48 Fourier features of the position in [0,1]^3 and the arrival
 * direction), 3 x 64 ReLU, 40 outputs (8 lobes x raw mean direction, raw concentration, logit;
 * kappa = softplus + 0.01, weights = softmax). Parameters flat: W1[64][51], b1, W2[64][64], b2,
 * W3[64][64], b3, W4[40][64], b4. Samples: forward n x 6 (position, arrival direction), training
 * n x 10 (position, arrival direction, scattered direction, weight W); the loss is the batch
 * mean of W * -log q(scattered). */
typedef struct upbp_mlp upbp_mlp;
uint32_t upbp_mlp_param_count(void);
int upbp_mlp_create(uint32_t seed, upbp_mlp** out); /* He-initialised */
void upbp_mlp_destroy(upbp_mlp* m);
int upbp_mlp_set_params(upbp_mlp* m, const float* params);
int upbp_mlp_get_params(upbp_mlp* m, float* params);
int upbp_mlp_forward(upbp_mlp* m, const float* samples, uint32_t n, float* out); /* n x 40 raw outputs */
int upbp_mlp_loss_grad(upbp_mlp* m, const float* samples, uint32_t n, float* loss, float* grads);
int upbp_mlp_train_step(upbp_mlp* m, const float* samples, uint32_t n, float lr, float* loss); /* one Adam step */
/* upbp_mlp_forward without any allocation (no counterpart in the reference): for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc. Same arguments and outputs as upbp_mlp_forward; its input staging and
 * outputs live in a fixed device array of 6.75 MiB per GPU (held until the process ends), so
 * n is at most upbp_mlp_noalloc_max() (32768) and a larger n returns UPBP_E_CAPACITY. Calls on
 * one GPU share that array: calls that overlap on one GPU race. */
int upbp_mlp_forward_noalloc(upbp_mlp* m, const float* samples, uint32_t n, float* out);
uint32_t upbp_mlp_noalloc_max(void);

/* ---- test hooks (parity tests and autotuning only) ----------------------
 * Range queries over a device LBVH, returning the accepted index set of every
 * query in CSR form (offsets[n_queries+1], indices). They replace the OptiX
 * photon-map traversals of gpuupbp/kernel/pde/functs_{surf,pp3d,pb2d,bb1d,bp2d}.cuh
 * with the acceptance predicates of gpuupbp/kernel/intersection.cuh:33-173.
 *   kind 0: point query,   predicate point_point(q, p, r^2)        (SURF / P-P3D)
 *   kind 1: segment query, predicate test_intersection_bre on (t0,t1) (P-B2D)
 *   kind 2: point query over beams,   point_beam(beam, q, r^2)      (B-P2D)
 *   kind 3: segment query over beams, test_intersection_beam_beam   (B-B1D)
 * points: n_prims*3 floats; beams: n_prims*7 floats (origin, unit dir, length);
 * queries: kind 0/2: 3 floats, kind 1/3: 8 floats (origin, unit dir, t0, t1).
 * Beams are cut into sub-beams of length about subbeam_len (0: whole beams) as in the
 * renderer's beam map; each accepted beam is still reported once.
 * indices has room for max_indices entries; *n_indices receives the total. */
int upbp_test_range_query(int kind, const float* prims, uint32_t n_prims, const float* queries,
                          uint32_t n_queries, float radius, float subbeam_len, uint32_t* offsets,
                          uint32_t* indices, uint64_t max_indices, uint64_t* n_indices);
/* Same, with wide != 0 walking the 4-wide collapse of the LBVH (kinds 1 and 3, as the
 * segment merge kernels do). */
int upbp_test_range_query_wide(int kind, const float* prims, uint32_t n_prims, const float* queries,
                               uint32_t n_queries, float radius, float subbeam_len, int wide, uint32_t* offsets,
                               uint32_t* indices, uint64_t max_indices, uint64_t* n_indices);

/* Closest-hit / all-hits ray queries against a triangle LBVH (replaces the optixTrace calls of
 * gpuupbp/kernel/intersector.cuh:120,214). rays: 8 floats each (origin, dir, tmin, tmax).
 * Closest: hit_t[i] (INFINITY on miss), hit_prim[i] (-1 on miss). */
int upbp_test_closest_hit(const float* verts, uint32_t n_tris, const float* rays, uint32_t n_rays,
                          float* hit_t, int32_t* hit_prim);

/* Analytic sphere test of every (ray, sphere) pair (replaces the sphere intersection program,
 * gpuupbp/kernel/intersector.cu:35-95). spheres: 4 floats each (centre, radius); rays: 8 floats
 * each. out[3 * (ray * n_sph + sphere) + {0, 1, 2}] = root count in (tmin, tmax), near root,
 * far root (INFINITY where absent). At most 2^30 pairs. */
int upbp_test_sphere_hits(const float* spheres, uint32_t n_sph, const float* rays, uint32_t n_rays, float* out);

/* Environment-map sampling and lookup (replaces sample / lookup of gpuupbp/kernel/env_map.cuh:
 * 105-137 over the distribution of gpuupbp/host/env_map.hpp:42-125). u: 2 floats per sample.
 * out: 11 floats per sample: sampled direction (3), its solid-angle pdf, radiance (3), then the
 * lookup of that direction: pdf, radiance (3). */
int upbp_test_env_samples(const upbp_env_map* env, const float* u, uint32_t n, float* out);

/* One learned-proposal record (8 x (mu.xyz, kappa, weight), packed as upbp_set_guide_lobes
 * packs it) times the phase lobes of a path arriving along `arrive` in a medium of mean cosine
 * g (surface 0), or times the cosine lobe about the normal `arrive` (surface 1), sampled with u
 * (3 per sample: lobe variate, two direction variates): out 4 per sample, the direction and
 * the product's pdf there. */
int upbp_test_guide_samples(const float* lobes, const float* arrive, float g, int32_t surface, const float* u,
                            uint32_t n, float* out);

/* MIS weight log (the MIS oracle's hook into the GPU renderer; the reference has no counterpart:
 * its weights are checked only through images). While on, every iteration logs one record per
 * MIS-weighted contribution of the camera paths whose index is a multiple of pixel_stride and of
 * light tracing from the light paths below light_paths: 4 x uint32 (technique | k << 8 | s << 16,
 * camera path (light tracing: light path), light path (next-event estimation and light hits: the
 * light), weight as FP32 bits). The full path is x_0 (on the light) .. x_k (the camera) and s its
 * light-side vertex count; techniques 0 light hit, 1 next-event estimation, 2 vertex connection,
 * 3 light tracing, 4 SURF, 5 P-P3D, 6 P-B2D, 7 B-P2D, 8 B-B1D. The weights are those of
 * gpuupbp/kernel/functs_camera.cuh:146-162 (light hit), :285-334 (next-event), :469-518
 * (connection), functs_light.cuh:358-369 (light tracing) and pde/functs_{pp,pb2d,bp2d,bb1d}.cuh.
 * Light hits and next-event estimation are logged for area lights only, connections and point
 * merges when the light vertex's predecessors were all stored, light tracing when every vertex of
 * the light path so far was stored. capacity 0 switches the log off; the log restarts with each
 * iteration. upbp_read_mis_log copies up to max_records records of the last iteration, sets
 * *n_records to the count logged and returns UPBP_E_CAPACITY if that exceeded the capacity. */
int upbp_set_mis_log(upbp_renderer* r, uint32_t capacity, uint32_t pixel_stride, uint32_t light_paths);
int upbp_read_mis_log(upbp_renderer* r, uint32_t* records, uint64_t max_records, uint64_t* n_records);

/* Training samples of the learned proposal (its oracle's hook; no reference counterpart): copies
 * up to max_samples of the samples in the training ring, 10 floats each (position in the record
 * grid's unit cube, three zero arrival inputs, scattered direction, weight), sets *n_samples to
 * the ring's fill and, with grid, the grid's low corner (3 floats) and edge length. Read after
 * the first trained iteration, the ring holds exactly that iteration's samples. Online training
 * must be on. */
int upbp_read_guide_samples(upbp_renderer* r, float* samples, uint64_t max_samples, uint64_t* n_samples, float* grid);

/* Streams `bytes` from one device buffer to another `reps` times (after one untimed copy)
 * with a plain copy kernel; *ms_per_copy receives the mean time. Reference kernel for HBM
 * traffic-counter calibration and the measured copy bandwidth. */
int upbp_test_copy(uint64_t bytes, int reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif

#endif /* UPBP_CAPI_H */
