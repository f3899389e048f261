/*
 * tpd_hip.h -- C ABI of the MI355X Gaussian rasterizer (libtpd_hip.so).
 *
 * This is the drop-in boundary for torpedo's Gaussian engine (torpedo/volumetric).  The
 * reference exposes a C++ class, GaussianEngine (GaussianEngine.h:15-27), whose methods record
 * Vulkan dispatches of Slang kernels; each entry point below replaces one of its operations and
 * is cited next to its declaration.  The C++ API in include/torpedo/ (Scene, Camera,
 * PerspectiveCamera, GaussianEngine) is built on these functions; Python binds them with ctypes.
 *
 * Conventions
 *   - Every function returns TPD_OK (0) or a TPD_ERR_* code and never throws; the message of the
 *     last failure on the calling thread is available from tpd_last_error().
 *   - Pointers named d_* are device pointers (hipMalloc'd, or torch CUDA tensors); all other
 *     array pointers are host memory, borrowed for the duration of the call.
 *   - Device memory of a context is owned by the context.  One context per host thread.
 *   - stream arguments are hipStream_t values; NULL selects the context's own stream.
 *   - Matrices are 4x4 row-major float[16], as torpedo's mat4 (math/mat4.h:8-13).
 */
#ifndef TPD_HIP_H
#define TPD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum tpd_status {
    TPD_OK = 0,
    TPD_ERR_INVALID = 1,   /* bad argument */
    TPD_ERR_HIP = 2,       /* HIP runtime error */
    TPD_ERR_NO_SCENE = 3,  /* no point cloud uploaded */
    TPD_ERR_NO_CAMERA = 4, /* no camera set */
    TPD_ERR_CAPACITY = 5,  /* more than 2^30 - 1 tile/Gaussian pairs */
    TPD_ERR_STATE = 6,     /* operation needs a previous step (e.g. read before render) */
    TPD_ERR_NO_DEVICE = 7  /* no usable GPU */
};

typedef struct tpd_ctx tpd_ctx;

/* GaussianPoint record (GaussianGeometry.h:10-30): 60 floats / 240 bytes. */
#define TPD_GAUSSIAN_FLOATS 60
/* Splat record as read back (splat.slang:33-39): color[3], tiles (uint32), texel[4], copac[4]. */
#define TPD_SPLAT_FLOATS 12

typedef struct tpd_config {
    int32_t device;           /* HIP device ordinal */
    uint32_t keep_unsorted;   /* non-zero: keep a copy of the pre-sort pairs for tpd_read_pairs */
    uint32_t reserved[6];
} tpd_config;

/* Message of the last failing call on this thread ("" if none). */
const char* tpd_last_error(void);

/* Context<R>::create + bindEngine<GaussianEngine> (rendering/include/torpedo/rendering/Context.h:18-36,
 * GaussianEngine.cpp:75-170): selects the device, creates a stream and the frame resources. */
int tpd_create(tpd_ctx** out, const tpd_config* cfg);
/* GaussianEngine::destroy (GaussianEngine.cpp:877-...) */
int tpd_destroy(tpd_ctx* ctx);

/* GaussianEngine::compile (GaussianEngine.cpp:359-399): uploads n GaussianPoint records.
 * entity_idx (n entries, may be NULL = all entity 0) maps each Gaussian to one of n_entities
 * model matrices (models, 16*n_entities floats, may be NULL = identity); sh_degree is clamped to
 * 3 as in GaussianEngine.cpp:366-370.  n == 0 is accepted (renders black, as the reference
 * warns and keeps going at GaussianEngine.cpp:362-365). */
int tpd_upload(tpd_ctx* ctx, const float* gaussians, uint32_t n, const uint32_t* entity_idx, const float* models,
               uint32_t n_entities, uint32_t sh_degree);

/* Same, but the n x 60 float records already live in device memory (borrowed, not copied; they
 * must stay valid while the context renders from them).  Used by the autograd wrapper. */
int tpd_bind_device_scene(tpd_ctx* ctx, const float* d_gaussians, uint32_t n, uint32_t sh_degree);

/* TransformHost::transform (rendering/src/TransformHost.cpp:3-12): sets one entity's model matrix. */
int tpd_set_transform(tpd_ctx* ctx, uint32_t entity, const float model[16]);

/* GaussianEngine::updateCameraBuffer (GaussianEngine.cpp:764-775) + the render target size
 * (initRenderer / onFramebufferResize, GaussianEngine.cpp:206-220). */
int tpd_set_camera(tpd_ctx* ctx, const float view[16], const float proj_view[16], const float focal_ndc[2],
                   uint32_t width, uint32_t height);

/* GaussianEngine::rasterFrame (GaussianEngine.cpp:621-712): preprocess, scan + duplicate, radix
 * sort, ranges, blend.  Blocks once on the pair count, as the reference does at :665-670.
 * Produces the FP32 image (H x W x 3) and the RGBA8 image on the device. */
int tpd_render(tpd_ctx* ctx, void* stream);
/* Waits for all work of the context's stream (Engine::waitIdle, rendering/include/torpedo/rendering/Engine.h:19). */
int tpd_synchronize(tpd_ctx* ctx);

/* ----- readbacks (the reference has none: its HeadlessRenderer cannot read pixels back) ----- */
int tpd_get_pair_count(tpd_ctx* ctx, uint32_t* pairs);
int tpd_read_image_f32(tpd_ctx* ctx, float* rgb);          /* H*W*3 floats */
int tpd_read_image_rgba8(tpd_ctx* ctx, uint8_t* rgba);     /* H*W*4 bytes  */
int tpd_read_splats(tpd_ctx* ctx, float* splats);          /* N*12 floats, tiles = per-Gaussian count */
int tpd_read_tiles(tpd_ctx* ctx, uint32_t* counts, uint32_t* offsets); /* N each (offsets may be NULL) */
/* Sorted (tile|depth) keys and Gaussian indices; with unsorted != 0 the keygen order instead
 * (requires tpd_config.keep_unsorted).  capacity is the length of keys/vals. */
int tpd_read_pairs(tpd_ctx* ctx, uint64_t* keys, uint32_t* vals, uint32_t capacity, int unsorted);
int tpd_read_ranges(tpd_ctx* ctx, uint32_t* ranges);       /* 2 * tiles uint32: [start, end) per tile */
/* The blend's launch order of the tiles (no counterpart in the reference, whose blend dispatch
 * takes tiles in index order, GaussianEngine.cpp:806-863): a permutation of the tiles, classes of
 * pair count (bit length) heaviest first, empty tiles last.  Not available for an empty scene. */
int tpd_read_tile_order(tpd_ctx* ctx, uint32_t* order);    /* tiles uint32 */
/* Device pointer of the FP32 image (valid until the next render or resize). */
int tpd_device_image(tpd_ctx* ctx, float** d_rgb);
/* Device-to-device copy of the FP32 image into d_dst (H*W*3 floats), ordered on stream. */
int tpd_copy_image(tpd_ctx* ctx, float* d_dst, void* stream);

/* ----- presentation (GaussianEngine::draw + recordTargetCopy, volumetric/src/GaussianEngine.cpp:714-762,
 * 865-875: the R8G8B8A8 target, :318, is copied into the swapchain image after each frame) -----
 * tpd_present_rgba8 copies the RGBA8 image into a caller's device surface of row pitch dst_pitch
 * bytes (0 = 4 * width), ordered on stream (NULL: the context's stream).  tpd_save_png writes it
 * to a PNG file (RGBA, 8 bits per channel) for callers without a window; tpd_write_png encodes a
 * host RGBA8 array (H*W*4 bytes, rows top to bottom) the same way and needs no GPU. */
int tpd_present_rgba8(tpd_ctx* ctx, void* d_dst, uint64_t dst_pitch, void* stream);
int tpd_save_png(tpd_ctx* ctx, const char* path);
int tpd_write_png(const uint8_t* rgba, uint32_t width, uint32_t height, const char* path);

/* ----- backward (the reference's differentiable rasterizer is an empty stub:
 * differential/include/torpedo/differential/GaussianRasterizer.h:1-8, README.md:8) -----
 * Gradients of the last rendered frame with respect to the uploaded/bound records, expressed in
 * the training parameterisation: means (N x 3), log-scales (N x 3, scale = exp(log-scale) as in
 * GaussianGeometry.cpp:115), quaternions as fed to the kernel (N x 4, unnormalised xyzw),
 * opacity logits (N, opacity = sigmoid(logit) as in GaussianGeometry.cpp:112) and SH (N x 48).
 * d_dL_dimage is the upstream gradient of the FP32 image (H x W x 3, device).  All gradient
 * pointers are device memory; with accumulate != 0 the results are added to their contents.
 * The per-pixel transmittance is recomputed, not stored by the forward. */
int tpd_backward(tpd_ctx* ctx, const float* d_dL_dimage, float* d_grad_means, float* d_grad_log_scales,
                 float* d_grad_quats, float* d_grad_opacity_logits, float* d_grad_sh, int accumulate, void* stream);
/* Per-splat screen-space gradients of the last backward (N x 9 floats: image x, y, conic a, b, c,
 * colour r, g, b, opacity), for tests. */
int tpd_read_splat_grads(tpd_ctx* ctx, float* g2d);
/* Profiled backward: ms[0] splat-gradient pass, ms[1] parameter-gradient pass, ms[2] span from the
 * start of the last profiled tpd_render to the end of this backward. */
int tpd_get_backward_times(tpd_ctx* ctx, float* ms);

/* ----- device memory for callers without another GPU runtime (TransferWorker's role,
 * foundation/src/TransferWorker.cpp:111-158): kind 0 host->device, 1 device->host,
 * 2 device->device; tpd_memcpy waits for the context's stream first. */
int tpd_alloc(tpd_ctx* ctx, uint64_t bytes, void** d_ptr);
int tpd_free(tpd_ctx* ctx, void* d_ptr);
int tpd_memcpy(tpd_ctx* ctx, void* dst, const void* src, uint64_t bytes, int kind);

/* ----- stage timing (the reference has none; SurfaceRenderer only measures whole frames) -----
 * With profiling on, tpd_render records HIP events on its stream between stages; ms[0..5]
 * receive the last frame's preprocess+scan+duplicate, sort histograms, all radix passes, ranges,
 * blend and whole-frame times in milliseconds; *passes the number of radix passes. */
int tpd_set_profiling(tpd_ctx* ctx, int enable);
int tpd_get_stage_times(tpd_ctx* ctx, float* ms, uint32_t* passes);
/* Pixel-splat pairs whose power the blend of the last rendered frame evaluates (counted by blending
 * that frame once more with a counting kernel; same image). */
int tpd_get_blend_evaluations(tpd_ctx* ctx, uint64_t* evaluations);
/* Time of one plain device-to-device copy of `bytes` (bandwidth calibration), averaged over iters. */
int tpd_measure_copy(tpd_ctx* ctx, uint64_t bytes, uint32_t iters, float* ms_per_copy);
/* Achievable bandwidth of a stream of `bytes`: kind 0 copy kernel, 1 hipMemcpyAsync device to device,
 * 2 nontemporal copy kernel, 3 read-only kernel, 4 write-only kernel; `blocks` is the kernels' grid
 * (0: 8192 blocks of 256 threads).  Kinds 5 and 6: 2^25 random 8-byte / 4-byte reads from an array
 * of `bytes` (the rate of a gather by a sorted index).  *ms is the mean time per pass over iters. */
int tpd_measure_bandwidth(tpd_ctx* ctx, uint32_t kind, uint64_t bytes, uint32_t iters, uint32_t blocks, float* ms);
/* The memory pattern of one stable radix pass without its ranking (the ceiling of a pass's data
 * movement): n keys of key_bytes (4 or 8) and n 4-byte values, in blocks of keys_per_block (a
 * multiple of 512, at most 16,384) whose keys each carry one of `bins` (256 or 512) random digits,
 * are read coalesced and written into per-(digit, block) runs at a stable pass's positions,
 * consecutive lanes storing consecutive places of one run, all keys before the values (blocks of
 * 512 threads, 1,024 from 12,288 keys per block).  n is a multiple of keys_per_block.
 * blocks_per_cu > 0 caps the blocks resident per CU (by unused LDS), 0 leaves them to the
 * registers.  *ms is the mean time per pass over iters. */
int tpd_measure_run_scatter(tpd_ctx* ctx, uint32_t key_bytes, uint32_t n, uint32_t bins, uint32_t keys_per_block,
                            uint32_t blocks_per_cu, uint32_t iters, float* ms);

/* ----- stand-alone stable radix sort (radix-*.slang, GaussianEngine.cpp:819-841) -----
 * Sorts n (key, value) pairs by the key's low key_bits bits (stable); bits at and above key_bits do
 * not take part in the order and are carried with the key.  d_keys_alt /
 * d_vals_alt are scratch of the same size; *in_alt is set to 1 when the result is there. */
int tpd_sort_pairs(tpd_ctx* ctx, uint64_t* d_keys, uint32_t* d_vals, uint64_t* d_keys_alt, uint32_t* d_vals_alt,
                   uint32_t n, uint32_t key_bits, void* stream, int* in_alt);
/* As tpd_sort_pairs, allocating nothing: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The sort workspace is a fixed 8,536,064-byte device array per GPU, held until the process
 * ends; n above 4,194,304 pairs returns TPD_ERR_CAPACITY.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race and must not be issued. */
int tpd_sort_pairs_noalloc(tpd_ctx* ctx, uint64_t* d_keys, uint32_t* d_vals, uint64_t* d_keys_alt,
                           uint32_t* d_vals_alt, uint32_t n, uint32_t key_bits, void* stream, int* in_alt);

/* ----- autotuning (the reference's shapes are compile-time constants, GaussianEngine.h:121-124) -----
 * Times every radix-pass block shape on the last rendered frame's pairs and keeps the fastest
 * (results are identical for every shape).  sort_ms[0..capacity) receives the median time of a
 * full sort per shape; *n_shapes the number of shapes; *best the chosen one.  Re-renders the frame. */
int tpd_autotune(tpd_ctx* ctx, float* sort_ms, uint32_t capacity, uint32_t* n_shapes, uint32_t* best);
/* Selects a radix-pass block shape by index (0 = default) for the plain and the batched frame. */
int tpd_set_sort_shape(tpd_ctx* ctx, uint32_t shape);

/* ----- batched rendering -------------------------------------------------------------------------
 * The reference renders one camera of one compiled scene per GaussianEngine::rasterFrame
 * (GaussianEngine.h:26; one Scene per compile, GaussianEngine.cpp:359-399) and has no batched form.
 * A batch holds several scenes in the context and renders many views of them (any resolutions) with
 * one launch per stage, recorded once into a HIP graph and replayed.  Every view's pairs, ranges and
 * image are those of a plain tpd_render of that scene and camera. */

/* One view of a batch: the scene it shows and the camera uniforms of tpd_set_camera. */
typedef struct tpd_view {
    uint32_t scene;
    uint32_t width, height;
    float view[16];
    float proj_view[16];
    float focal_ndc[2];
} tpd_view;

/* One scene's parameter-gradient buffers (device; layouts as tpd_backward). */
typedef struct tpd_scene_grads {
    float* means;           /* N x 3 */
    float* log_scales;      /* N x 3 */
    float* quats;           /* N x 4 */
    float* opacity_logits;  /* N */
    float* sh;              /* N x 48 */
} tpd_scene_grads;

/* Scene::add + GaussianEngine::compile for one more scene of the batch (Scene.h:26-62,
 * GaussianEngine.cpp:359-399, transforms :442-486); arguments as tpd_upload.  *scene_id receives
 * the scene's index (0, 1, ... in call order). */
int tpd_scene_add(tpd_ctx* ctx, const void* gaussians, uint32_t n, const uint32_t* entity_idx, const float* models,
                  uint32_t n_ent, uint32_t sh_degree, uint32_t* scene_id);
/* As tpd_scene_add with records already in device memory (borrowed, 16-byte aligned, N x 60 floats;
 * they are read by every batch frame and must stay valid). */
int tpd_scene_add_device(tpd_ctx* ctx, const float* d_gaussians, uint32_t n, uint32_t sh_degree, uint32_t* scene_id);
/* TransformHost::transform (TransformHost.cpp:3-12) for an entity of a batch scene. */
int tpd_scene_set_transform(tpd_ctx* ctx, uint32_t scene, uint32_t entity, const float model[16]);
/* Removes every batch scene and view and frees the batch's buffers. */
int tpd_scene_clear(tpd_ctx* ctx);
/* Sets the batch's views (image v of the batch is views[v]).  Changing only the cameras keeps the
 * recorded frame; changing a view's scene or size re-lays the batch out.  Waits for a running frame. */
int tpd_batch_set_views(tpd_ctx* ctx, const tpd_view* views, uint32_t nviews);
/* enabled = 0 issues the batch's kernels one by one instead of replaying the recorded graph. */
int tpd_batch_set_graph(tpd_ctx* ctx, int enabled);
/* Launches a batch frame (asynchronous).  for_backward != 0 keeps the per-view state that
 * tpd_batch_backward needs.  tpd_batch_synchronize or any batch read waits for the frame and, if its
 * pairs overflowed the buffers, grows them and renders it again. */
int tpd_batch_render(tpd_ctx* ctx, int for_backward, void* stream);
int tpd_batch_synchronize(tpd_ctx* ctx);
/* Gradients of the last batch frame (rendered with for_backward) for dL/dimages d_dL_dimages
 * (all views' H x W x 3 images concatenated in view order, as tpd_batch_device_images) into the
 * per-scene buffers grads[0 .. number of scenes); accumulate != 0 adds to them. */
int tpd_batch_backward(tpd_ctx* ctx, const float* d_dL_dimages, const tpd_scene_grads* grads, int accumulate,
                       void* stream);
/* Forward + backward of the batch as one recorded graph (asynchronous, validated as tpd_batch_render). */
int tpd_batch_step(tpd_ctx* ctx, const float* d_dL_dimages, const tpd_scene_grads* grads, int accumulate, void* stream);
/* Pair count of the last batch frame (all views). */
int tpd_batch_get_pair_count(tpd_ctx* ctx, uint32_t* P);
/* Items of the last batch frame (every (Gaussian, view) of its scenes) and how many of them its
 * item sort kept: the visible ones, whose tile count is non-zero (the sort drops the others in its
 * first pass, and the later passes and the pair emission run on the kept ones).  No counterpart in
 * the reference, which renders one camera per frame. */
int tpd_batch_get_item_counts(tpd_ctx* ctx, uint32_t* items, uint32_t* kept);
/* Where view `view` lies in the batch's index spaces: its first global tile (keys carry global
 * tiles), its first item (pair values are items; item - item_base = Gaussian index) and its first
 * pixel in the concatenated images. */
int tpd_batch_view_layout(tpd_ctx* ctx, uint32_t view, uint32_t* tile_base, uint32_t* item_base, uint64_t* pix_base);
/* Readbacks of one view's image (H x W x 3 float, H x W x 4 UNORM8). */
int tpd_batch_read_image_f32(tpd_ctx* ctx, uint32_t view, float* host);
int tpd_batch_read_image_rgba8(tpd_ctx* ctx, uint32_t view, uint8_t* host);
/* Presentation of one view's RGBA8 image, as tpd_present_rgba8 / tpd_save_png (waits for the frame). */
int tpd_batch_present_rgba8(tpd_ctx* ctx, uint32_t view, void* d_dst, uint64_t dst_pitch, void* stream);
int tpd_batch_save_png(tpd_ctx* ctx, uint32_t view, const char* path);
/* Device pointer of the concatenated float images (owned by the context) and their total pixels. */
int tpd_batch_device_images(tpd_ctx* ctx, float** d_images, uint64_t* total_pixels);
/* The sorted global (key, value) pairs of the last batch frame; *P receives the count. */
int tpd_batch_read_pairs(tpd_ctx* ctx, uint64_t* keys, uint32_t* vals, uint32_t capacity, uint32_t* P);
/* With tpd_set_profiling on: ms[0] = GPU time of the last batch frame; for forward frames issued
 * without the graph also ms[1..6] = items (projection of every view), item sort, pair emission,
 * pair sort, ranges + tile order, blend (else -1).  *passes receives the radix passes per frame
 * (both sorts). */
int tpd_batch_get_times(tpd_ctx* ctx, float* ms, uint32_t* passes);
/* The last eager profiled forward frame's two sorts (items by depth, then pairs by tile): ms[4] =
 * item histogram, item radix passes, pair histogram, pair radix passes (milliseconds); counts[4]
 * = items, item passes, pairs, pair passes. */
int tpd_batch_get_sort_times(tpd_ctx* ctx, float* ms, uint32_t* counts);
/* tpd_get_blend_evaluations for the last batch frame: the pixel-splat pairs whose power its blend
 * evaluated (counted by blending that frame once more with a counting kernel; same image). */
int tpd_batch_get_blend_evaluations(tpd_ctx* ctx, uint64_t* evaluations);
/* tpd_autotune for the batch's two sorts (times whole batch frames per radix-pass block shape and
 * keeps the fastest shape for each sort); *best = the pair sort's shape. */
int tpd_batch_autotune(tpd_ctx* ctx, float* sort_ms, uint32_t capacity, uint32_t* n_shapes, uint32_t* best);
/* The radix block shapes of the batch's item sort (by depth) and pair sort (by tile). */
int tpd_batch_get_sort_shapes(tpd_ctx* ctx, uint32_t* item_shape, uint32_t* pair_shape);

/* ----- point-cloud files -----
 * GaussianPoint::fromModel (volumetric/src/GaussianGeometry.cpp:59-127): reads a 3DGS PLY file (ascii or
 * binary, any PLY property types) into *n GaussianPoint records (60 floats each) allocated with malloc;
 * release them with tpd_free_host.  Opacity logits become sigmoid(logit), rot_0..3 (w first) become a
 * normalised (x, y, z, w) quaternion, log-scales become exp(scale) with modifier 1, and the SH features
 * are f_dc_0..2 followed by the f_rest_* properties in file order (missing ones zero). */
int tpd_load_ply(const char* path, float** records, uint32_t* n);
void tpd_free_host(void* p);

/* ----- host camera math (torpedo/math, rendering/src/Camera.cpp, extension/src/PerspectiveCamera.cpp) ----- */
/* Camera::lookAt (Camera.cpp:3-16) */
void tpd_look_at(const float eye[3], const float center[3], const float up[3], float view[16]);
/* PerspectiveCamera::updateProjectionMatrix (PerspectiveCamera.cpp:9-23); fy = 1/tan(fov/2) */
void tpd_perspective(uint32_t width, uint32_t height, float fy, float znear, float zfar, float proj[16]);
/* PerspectiveCamera::setVerticalFov (PerspectiveCamera.cpp:3-7) */
float tpd_fy_from_fov(float degrees);
/* math::mul (mat4.h:256-264), compensated */
void tpd_mat4_mul(const float a[16], const float b[16], float out[16]);
/* updateCameraBuffer's host part (GaussianEngine.cpp:764-768): projView = proj * view, focal = (P00, P11) */
void tpd_camera_upload(const float view[16], const float proj[16], float proj_view[16], float focal_ndc[2]);

#ifdef __cplusplus
}
#endif

#endif /* TPD_HIP_H */
