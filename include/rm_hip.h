/*
 * rm_hip.h — C-ABI of the MI355X rendering backend (librm_hip.so).
 *
 * This is the drop-in boundary that replaces the reference's Metal device:
 * every entry point below takes plain pointers and sizes, and the record
 * layouts are the byte layouts the reference uploads to its GPU
 * (src/engine/metal_types.h:3-57; sizes asserted at src/engine/metal.mm:13-17),
 * so a caller that already marshals scenes for the reference can hand the
 * same arrays over unchanged.
 *
 * All functions return 0 on success and a nonzero RM_ERR_* code otherwise; the
 * message is available from rm_last_error(). Calls on one context are
 * serialised by an internal mutex (metal.mm:104) and are synchronous: when a
 * call returns, its outputs are in host memory.
 */
#ifndef RM_HIP_H
#define RM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RM_OK = 0,
    RM_ERR_NO_DEVICE = 1,   /* no HIP GPU (reference: "No Metal GPU found", metal.mm:47) */
    RM_ERR_INVALID = 2,     /* bad argument (reference: metal.mm:98-101) */
    RM_ERR_ALLOC = 3,       /* device allocation failed (metal.mm:185,191,267) */
    RM_ERR_DEVICE = 4,      /* a HIP call or kernel failed (metal.mm:368-371) */
    RM_ERR_NO_SCENE = 5     /* rm_render/rm_trace before rm_set_scene */
};

typedef struct { float x, y, z; } rm_vec3;

/* Triangle record, 176 bytes (metal_types.h:5-11). */
typedef struct {
    rm_vec3 v0, v1, v2, n, nn, n0, n1, n2;
    float ar, ag, ab, er, eg, eb, ior, rough, metallic, alpha;
    int32_t shader;
    float tu0, tv0, tu1, tv1, tu2, tv2;
    int32_t texture_offset, texture_width, texture_height;
} rm_triangle;

/* Binary BVH node in preorder, 36 bytes (metal_types.h:17-20): internal nodes
 * have count == 0, their left child follows them, and right is the index just
 * past the node's subtree (the escape link built at metal.mm:129-144). */
typedef struct {
    rm_vec3 mn, mx;
    int32_t right, offset, count;
} rm_bvh_node;

/* Camera, 52 bytes (metal_types.h:12-15): position, focal length in pixels and
 * the row-major camera-to-world rotation m0..m8. */
typedef struct {
    float x, y, z, focal;
    float m0, m1, m2, m3, m4, m5, m6, m7, m8;
} rm_camera;

/* Primary-hit G-buffer record, 76 bytes (metal_types.h:32-42). */
typedef struct {
    rm_vec3 pos, normal, albedo, wo, emission;
    float rough, metallic, ior;
    int32_t kind; /* 0 miss, 1 opaque, 3 dielectric */
} rm_gbuffer;

/* ReSTIR DI reservoir, 60 bytes (metal_types.h:46-56). */
typedef struct {
    rm_vec3 light_pos, light_normal, emission;
    float alpha, p_hat, sum_ws, sum_wu;
    uint32_t m, valid;
} rm_reservoir;

/* One scene as the reference uploads it (metal.mm:109-186). */
typedef struct {
    const rm_triangle* tris;   size_t ntri;
    const rm_bvh_node* nodes;  size_t nnode;   /* preorder with escape links */
    const int32_t* bvh_tri;                     /* ntri entries; leaves index it */
    const int32_t* light_tri;  const float* light_cdf; size_t nlight;
    float light_total;
    const float* texels;       size_t ntexel;  /* linear RGB triples */
} rm_scene_desc;

/* Render settings after the host-side defaults of metal.mm:146-170 (the C++
 * shim render_image_hip applies those and the environment variables). */
typedef struct {
    int32_t width, height, spp, bounces;
    float resolution, clamp_value;
    int32_t adapt_min, adapt_step;
    float adapt_rel, adapt_abs;
    int32_t nomis;
    int32_t restir_candidates, restir_spatial, restir_mcap;
    float restir_radius;
    int32_t denoise;             /* RAYMOTION_DENOISE (metal.mm:269-271): 0 off, 1 a-trous;
                                    2 learned kernel-predicting network (needs rm_kpn_set_weights) */
    int32_t denoise_iterations;  /* a-trous passes; the reference runs 2 */
    int32_t temporal;            /* RAYMOTION_RESTIR_TEMPORAL (metal.mm:251-252) */
    rm_camera camera;
} rm_frame;

/* Per-frame counters of the last rm_render (or rm_trace) call. */
typedef struct {
    int32_t path;         /* 0 megakernel (render_paths), 1 ReSTIR + denoise */
    int32_t lanes;        /* megakernel lanes per pixel: 1, 8 or 32 */
    int32_t has_history;  /* temporal reuse was active */
    int32_t refit;        /* last scene upload's tree: 0 unchanged, 1 refit, 2 rebuild from
                             the given nodes, 3 built on the GPU (rm_set_scene_gpu_tree) */
    /* Per pass: megakernel, generate, spatial, shade, probe, and the total. */
    uint64_t rays[6];        /* rays traced */
    uint64_t node_visits[6]; /* 4-wide BVH nodes fetched */
    uint64_t tri_tests[6];   /* ray/triangle tests */
    float ms[8];          /* GPU time per pass: megakernel, generate, spatial, shade,
                             temporal (or feature build for denoise = 2), atrous (all
                             passes, or the network and its filter), tonemap, whole frame */
    int32_t bvh_nodes;    /* 4-wide nodes resident */
    int32_t stack_depth;  /* traversal stack entries the resident tree needs */
    float scene_ms[2];    /* last rm_set_scene, GPU time: host-to-device copies, refit kernels */
    int32_t wave_iters;   /* wavefront schedule: trace + step iterations of the path pass */
    uint32_t wave_tail;   /* active pixels then finished by the persistent loop */
} rm_stats;

typedef struct rm_ctx rm_ctx;

/* Device context: HIP device, stream, persistent buffers and temporal history.
 * Replaces the reference's static Context (metal.mm:23-92,102). */
int rm_create(rm_ctx** out, int hip_device);
void rm_destroy(rm_ctx* ctx);
/* Message for the last failure on ctx; ctx == NULL gives the last rm_create failure. */
const char* rm_last_error(const rm_ctx* ctx);

/* Uploads a scene (metal.mm:109-186 marshalling + upload, 194-231 refit or
 * rebuild policy). Identical content is not re-uploaded, and of a changed
 * scene only the parts that changed are (geometry, materials, lights, tree).
 * When vertex positions changed and the BVH topology is the same, the
 * resident 4-wide tree is refit on the GPU, at most 31 times in a row;
 * otherwise it is rebuilt from the given nodes. */
int rm_set_scene(rm_ctx* ctx, const rm_scene_desc* scene);

/* As rm_set_scene, but the 4-wide tree is built on the GPU from the triangles
 * (Morton order, then parallel locally-ordered clustering by surface area,
 * collapsed to 4-wide nodes with one triangle per leaf slot); scene->nodes and scene->bvh_tri are not
 * used and may be empty. The tree is rebuilt when `rebuild` is nonzero, when
 * the triangle count changed or after 31 refits in a row, and otherwise refit
 * on the GPU when positions moved. Images equal those of rm_set_scene with any
 * valid tree: traversal returns the closest hit whatever the tree. No
 * counterpart in the reference, which builds its tree on the host (the
 * rebuild every 32nd geometry update of its runtime.hpp:119-130). */
int rm_set_scene_gpu_tree(rm_ctx* ctx, const rm_scene_desc* scene, int rebuild);

/* One frame (render_image_metal, metal.mm:95-414): rgb8 receives W*H*3 bytes,
 * row-major, top row first; hdr (may be NULL) receives the W*H*3 linear values
 * that were tonemapped. Either output may be NULL. */
int rm_render(rm_ctx* ctx, const rm_frame* frame, uint64_t seed, uint8_t* rgb8, float* hdr);

/* Closest-hit (any == 0) or any-hit (any != 0) queries against the resident
 * scene: the traversal of PathTracer::intersect (pathtrace.metal:119-163) as a
 * standalone probe. origins/dirs: 3*n floats; tmax may be NULL (1e30). tri
 * receives the hit triangle index (scene order) or -1; t/u/v may be NULL. */
int rm_trace(rm_ctx* ctx, const float* origins, const float* dirs, const float* tmax, size_t n, int any,
             int32_t* tri, float* t, float* u, float* v);

/* Intermediate buffers of the last ReSTIR frame in the reference layouts:
 * G-buffer, initial and merged reservoirs (W*H each) and the shaded HDR image
 * before denoising (W*H*3). Any pointer may be NULL. */
int rm_read_restir(rm_ctx* ctx, rm_gbuffer* gbuf, rm_reservoir* rinit, rm_reservoir* rmerged, float* shade);

/* The resident traversal tree as the kernels read it: counts receives the
 * number of 4-wide nodes, of triangles and the stack depth the tree needs;
 * nodes (if not NULL, room for node_cap nodes) 32 floats per node: lower x,
 * upper x, lower y, upper y, lower z, upper z bounds of the four children,
 * their references and kinds as int bits (-1 empty, 0 inner node, k > 0 leaf
 * of k triangles from leaf-order index ref); tris (if not NULL, room for
 * tri_cap triangles) 12 floats per triangle in leaf order: v0, v1, v2, the
 * scene index as int bits, two unused. For inspection and for counting a
 * traversal's fetches on the host; the reference's counterpart is its
 * acceleration structure, which it does not expose. */
int rm_read_tree(rm_ctx* ctx, float* nodes, size_t node_cap, float* tris, size_t tri_cap, int32_t* counts);

/* rm_trace on device memory, allocating nothing. No counterpart in the
 * reference: for callers that cannot allocate device memory during the call
 * (a stream being captured into a graph, a serving loop, a harness that
 * forbids allocation). origins/dirs: 3*n floats, tmax: n floats or NULL, tri:
 * n ints, tuv: 4*n floats (t, u, v, 0 per ray, 16-byte aligned) -- all device
 * memory of the context's GPU, allocated by the caller. At most 2^30 rays per
 * call. The traversal counters live in a fixed 256-byte array in device
 * memory, one per GPU, shared by every call on that GPU: calls that overlap on
 * one GPU race. Synchronous like the other entry points; rm_get_stats reports
 * the call's counters. */
int rm_trace_device(rm_ctx* ctx, const float* origins, const float* dirs, const float* tmax, size_t n, int any,
                    int32_t* tri, float* tuv);

/* The ReSTIR merge pass alone (restir_spatial, pathtrace.metal:919-1054) over
 * caller-supplied G-buffer and initial reservoirs (W*H each, the layouts
 * rm_read_restir returns), with shadow rays against the resident scene. Reads
 * the frame's width, height, camera, restir_spatial, restir_radius and
 * restir_mcap; seed is the frame seed of the spatial pass's streams. When
 * hist_gbuf, hist_res and prev_camera are all given, the temporal reuse of
 * pathtrace.metal:956-1002 runs first against them; all three NULL means no
 * history. rmerged receives W*H reservoirs. The context's rendering history
 * is not touched. No separate reference entry point: the reference runs this
 * pass only inside render_image_metal (metal.mm:299-310). */
int rm_restir_merge(rm_ctx* ctx, const rm_frame* frame, uint64_t seed, const rm_gbuffer* gbuf,
                    const rm_reservoir* rinit, const rm_gbuffer* hist_gbuf, const rm_reservoir* hist_res,
                    const rm_camera* prev_camera, rm_reservoir* rmerged);

/* The a-trous denoiser alone (denoise_temporal, then frame->denoise_iterations
 * denoise_atrous passes; pathtrace.metal:1263-1354) over a caller-supplied
 * linear image (W*H*3) and its G-buffer. Reads the frame's width, height,
 * camera and denoise_iterations (0..30). With hist_color, hist_gbuf and
 * prev_camera all given the temporal pass blends that history; all three NULL
 * passes the image through it unchanged. out receives W*H*3 linear values
 * (before tonemapping). No scene is needed and the context's rendering history
 * is not touched. The reference runs these passes only inside
 * render_image_metal (metal.mm:321-366). */
int rm_denoise_image(rm_ctx* ctx, const rm_frame* frame, const float* color, const rm_gbuffer* gbuf,
                     const float* hist_color, const rm_gbuffer* hist_gbuf, const rm_camera* prev_camera, float* out);

/* Path-replay adjoint of the path-tracing frame (the megakernel path of
 * rm_render: no ReSTIR, legacy light sampling, no adaptive stop) with the
 * same frame and seed: the samples rm_render averaged are re-traced from
 * their seeds and grad receives, per parameter group, the gradient of
 * sum_px d_hdr . hdr with respect to albedo r, g, b, roughness and emission
 * scale (5 doubles per group). tri_group gives each triangle's group (scene
 * order, -1 for none); a group's emission is scale * group_emission (3 floats
 * per group). Sampling, MIS weights and roulette are held fixed: with
 * clamp_value = 0 and nomis = 1 that is the exact derivative of the expected
 * image; with the per-sample clamp, or with the reference's MIS (whose weight
 * at a BSDF-sampled light hit uses the previous vertex's one-candidate light
 * estimate, so the weights do not sum to one), the derivatives of the
 * roulette probabilities and of the weights are left out. The
 * differentiable counterpart of the reference's forward
 * (pathtrace.metal:552-769); it has no reference counterpart. */
int rm_render_adjoint(rm_ctx* ctx, const rm_frame* frame, uint64_t seed, const float* d_hdr, const int32_t* tri_group,
                      const float* group_emission, int ngroup, double* grad);

/* Material-table update for inverse rendering: rewrites, in the resident
 * scene, the albedo (r, g, b), roughness and emission of every triangle with
 * tri_group[i] >= 0 (ntri entries, scene order, -1: unchanged) from params
 * (5 floats per group: albedo r, g, b, roughness, emission scale; a group's
 * emission is the float product group_emission[g] * scale, 3 floats per
 * group), rewrites the lights' emission the same way and installs the
 * caller's light CDF (nlight floats, CDF order) and total. Afterwards the
 * resident scene renders as rm_set_scene of the scene with those values
 * would; geometry, tree, textures and every other material field stay. The
 * first call for a tri_group builds a per-(group, material) table (O(ntri));
 * later calls with the same tri_group send only the material and light tables
 * and the CDF. ntri and nlight must equal the resident scene's. Drops the
 * rendering history. The reference re-exports the whole scene for a material
 * change (metal.mm:140-196); this entry point has no reference counterpart. */
int rm_set_group_materials(rm_ctx* ctx, const int32_t* tri_group, size_t ntri, int ngroup, const float* params,
                           const float* group_emission, const float* light_cdf, size_t nlight, float light_total);

/* ---- Learned kernel-predicting denoiser (no reference counterpart; the
 * reference's denoiser is the a-trous filter, pathtrace.metal:1263-1354, which
 * stays the default and the fallback). rm_frame.denoise = 2 runs a 6-layer
 * conv3x3 network (BF16 MFMA, FP32 accumulation) on per-pixel features of the
 * ReSTIR frame and filters the noisy radiance with the predicted 5x5 kernels. */
#define RM_KPN_FEATURES 16 /* per pixel: log(1+radiance) 3, albedo 3, normal 3,
                              log(1+depth), log(1+luminance variance), motion 2, zero 3 */
/* Parameter count (floats): per layer [cout][kp] weights, then cout biases. */
size_t rm_kpn_param_count(void);
/* Sets the weights (and restarts the optimizer state); required for denoise = 2. */
int rm_kpn_set_weights(rm_ctx* ctx, const float* params, size_t n);
int rm_kpn_get_weights(rm_ctx* ctx, float* params, size_t n);
/* Renders a ReSTIR frame (no denoiser) and returns the network's inputs:
 * features (W*H*RM_KPN_FEATURES) and noisy linear radiance (W*H*3). */
int rm_render_features(rm_ctx* ctx, const rm_frame* frame, uint64_t seed, float* features, float* noisy);
/* Device-resident training set of equally sized images with references. */
int rm_kpn_dataset_clear(rm_ctx* ctx);
int rm_kpn_dataset_add(rm_ctx* ctx, const float* features, const float* noisy, const float* reference, int width,
                       int height);
/* Adam steps on random batches of the training set; loss per step:
 * mean((log(1 + denoised) - log(1 + reference))^2). losses may be NULL. */
int rm_kpn_train(rm_ctx* ctx, int steps, int batch, float lr, uint64_t seed, float* losses);
/* Denoises one image given as features + noisy radiance; out W*H*3. */
int rm_kpn_denoise(rm_ctx* ctx, const float* features, const float* noisy, int width, int height, float* out);
/* Loss and parameter gradient of the current weights on `count` images. */
int rm_kpn_gradient(rm_ctx* ctx, const float* features, const float* noisy, const float* reference, int count,
                    int width, int height, float* grad, float* loss);

/* Schedule of the path passes (restir_shade, render_paths with one lane per
 * pixel): 0 (default) persistent lanes; 1 wavefront kernels over path states
 * kept in device memory, the last pixels finished by persistent lanes; 2
 * wavefront kernels to the end. All give identical images. No counterpart in
 * the reference, whose passes run one thread per pixel. */
int rm_set_schedule(rm_ctx* ctx, int schedule);

/* Per-pass GPU timing with HIP events (off by default). */
int rm_set_profiling(rm_ctx* ctx, int enable);
int rm_get_stats(rm_ctx* ctx, rm_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* RM_HIP_H */
