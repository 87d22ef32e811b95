/*
 * splat.h -- C ABI of the MI355X (gfx950) splat training library `libsplat.so`.
 *
 * Drop-in boundary for the differentiable splat pipeline of harryprayiv/splatrs
 * (a Rust crate).  Every entry point below names the reference function it replaces
 * (reference file:line).  All entry points return an `int` status (SPLAT_OK = 0); on
 * failure `splat_last_error(ctx)` holds a message.  Precondition failures (the cases in
 * which the reference `assert!`s and panics) return SPLAT_ERR_PRECONDITION, so a Rust
 * wrapper can map them back to `panic!`.
 *
 * Scalars are IEEE binary32 (the reference's default `Real = f32`, src/lib.rs:18-20).
 *
 * Two tiers:
 *   - Host tier (`splat_project_all`, `splat_bin`, `splat_rasterize_forward`, ...):
 *     the reference's functions with host arrays in and out, record layouts identical
 *     to the reference structs.  They upload, compute on the GPU and download.
 *   - Trainer (`splat_trainer_*`): model, gradients, Adam moments, statistics, views and
 *     all per-view buffers stay resident in HBM; one call runs a whole training
 *     iteration (src/train.rs:196-303) on the context's stream.
 *
 * Threading: one context owns one HIP stream.  Calls on one context must not overlap;
 * distinct contexts are independent.
 */
#ifndef SPLAT_H
#define SPLAT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants ------------------------------------------------------------------ */

/* Scalars per Gaussian record: mean 3, log_scale 3, quat 4 (w,x,y,z), opacity logit 1,
 * SH 16x3 (sh[k][c] at 11 + 3k + c).  Order of tests/common/mod.rs:111-139. */
#define SPLAT_NUM_PARAMS 59
/* Tile edge in pixels, src/project.rs:18. */
#define SPLAT_TILE 16
/* Gradient slots per tile entry, src/rasterize.rs:341. */
#define SPLAT_SLOTS 11

/* ---- status codes --------------------------------------------------------------- */
enum {
  SPLAT_OK = 0,
  SPLAT_ERR_PRECONDITION = 1, /* the reference panics (assert!) on this input */
  SPLAT_ERR_HIP = 2,          /* HIP runtime error */
  SPLAT_ERR_OOM = 3,          /* device allocation failed */
  SPLAT_ERR_CAPACITY = 4,     /* caller-provided output array too small; size reported */
  SPLAT_ERR_INVALID = 5,      /* invalid argument (Error::Invalid in the reference) */
  SPLAT_ERR_IO = 6,           /* file I/O (Error::Io) */
  SPLAT_ERR_FORMAT = 7,       /* malformed file (Error::Format) */
  SPLAT_ERR_NO_DEVICE = 8     /* no usable gfx950 device */
};

/* ---- plain records (field order of the reference structs) ----------------------- */

/* Camera, src/camera.rs:13-32 (the `name` field is host-only and omitted).
 * rot is row-major world-to-camera, x_cam = rot * x_world + trans. */
typedef struct {
  uint32_t width, height;
  float fx, fy, cx, cy;
  float rot[9];
  float trans[3];
} splat_camera;

/* ProjOpts, src/project.rs:22-43 (defaults near 0.2, eps2d 0.3, antialias off, degree 3). */
typedef struct {
  float near_plane;
  float eps2d;
  int32_t antialias;
  int32_t sh_degree;
} splat_proj_opts;

/* Proj, src/project.rs:47-64 (60 bytes).  radius == 0 means culled. */
typedef struct {
  float mean2d[2];
  float conic[3];
  float depth;
  uint32_t radius;
  float opacity;
  float color[3];
  uint32_t tile_min[2];
  uint32_t tile_max[2];
} splat_proj;

/* ProjGrad, src/project.rs:84-93 (36 bytes). */
typedef struct {
  float mean2d[2];
  float conic[3];
  float opacity;
  float color[3];
} splat_proj_grad;

/* RasterSettings, src/rasterize.rs:15-22. */
typedef struct {
  uint32_t width, height;
  float background[3];
} splat_raster_settings;

/* Binning, src/rasterize.rs:26-39, as caller-owned arrays.  `sorted` is split into
 * `sorted_keys` (tile << 32 | depth bits) and `sorted_entries`.  On input `capacity`
 * is the length of entry_gaussian / sorted_keys / sorted_entries; if the binning needs
 * more, the call returns SPLAT_ERR_CAPACITY with `num_entries` set and writes nothing
 * else.  entry_offset has n+1 elements, tile_range 2 * tiles_x * tiles_y. */
typedef struct {
  uint32_t tiles_x, tiles_y;
  uint64_t num_entries;
  uint64_t capacity;
  uint32_t* entry_gaussian;
  uint32_t* entry_offset;
  uint64_t* sorted_keys;
  uint32_t* sorted_entries;
  uint32_t* tile_range;
} splat_binning;

/* LearningRates, src/adam.rs:11-24. */
typedef struct {
  float mean, log_scale, quat, opacity, sh_dc, sh_rest;
} splat_learning_rates;

/* Adam hyper-parameters and step counter, src/adam.rs:28-41 (defaults 0.9, 0.999, 1e-15). */
typedef struct {
  float beta1, beta2, eps;
  uint32_t pad_;
  uint64_t step;
} splat_adam_state;

/* DensifyConfig, src/densify.rs:69-103.  max_screen_size is Option<u32>: used when
 * has_max_screen_size != 0. */
typedef struct {
  float grad_threshold;
  float percent_dense;
  float min_opacity;
  int32_t prune_big_world;
  float max_world_size_frac;
  int32_t has_max_screen_size;
  uint32_t max_screen_size;
  uint32_t split_children;
} splat_densify_config;

/* DensifyReport, src/densify.rs:107-114. */
typedef struct {
  uint64_t cloned, split, pruned;
} splat_densify_report;

/* ---- context --------------------------------------------------------------------- */

typedef struct splat_ctx splat_ctx;

/* Create a context on HIP device `device` with its own stream.  Fails with
 * SPLAT_ERR_NO_DEVICE when the device is absent or not gfx950. */
int splat_ctx_create(int device, splat_ctx** out);
int splat_ctx_destroy(splat_ctx* ctx);
/* Message of the last failure on this context ("" if none).  Valid until the next call. */
const char* splat_last_error(const splat_ctx* ctx);
/* Block until all work queued on the context's stream has finished. */
int splat_sync(splat_ctx* ctx);
/* The context's hipStream_t, as an opaque pointer. */
void* splat_ctx_stream(splat_ctx* ctx);
/* Per-kernel timing with HIP events on the context stream (off by default).  When on,
 * every kernel launch is bracketed by events; splat_kernel_times reports, for each
 * kernel name, the launch count and the summed duration in milliseconds since the last
 * reset.  `names` receives up to `cap` NUL-separated names in one buffer of `names_len`
 * bytes.  Returns the number of kernels in *count. */
int splat_set_timing(splat_ctx* ctx, int enabled);
/* Algorithm of splat_bin (no counterpart in the reference, whose bin() sorts all keys with
 * par_sort_unstable): 0 (default) counts entries per tile and sorts each tile's depth keys,
 * 1 sorts the rows by depth and then the entries by tile with stable radix sorts (the
 * batched engine always bins this way).  Both give the reference's (tile, depth, entry)
 * order. */
int splat_set_bin_method(splat_ctx* ctx, int method);
int splat_reset_timing(splat_ctx* ctx);
int splat_kernel_times(splat_ctx* ctx, char* names, size_t names_len, uint64_t* launches,
                       double* total_ms, uint32_t cap, uint32_t* count);

/* ---- host tier: the reference functions ----------------------------------------- */

/* project_all, src/project.rs:320-330.  gaussians: n records of SPLAT_NUM_PARAMS
 * floats.  opts->sh_degree is the model's active degree (the reference overrides the
 * option with `model.sh_degree`). */
int splat_project_all(splat_ctx* ctx, const float* gaussians, uint64_t n,
                      const splat_camera* cam, const splat_proj_opts* opts,
                      splat_proj* out);

/* project_backward_all, src/project.rs:477-493.  out: n records of 59 floats. */
int splat_project_backward_all(splat_ctx* ctx, const float* gaussians, uint64_t n,
                               const splat_camera* cam, const splat_proj_opts* opts,
                               const splat_proj_grad* grads, float* out);

/* bin, src/rasterize.rs:60-128. */
int splat_bin(splat_ctx* ctx, const splat_proj* projs, uint64_t n, uint32_t width,
              uint32_t height, splat_binning* out);

/* rasterize_forward, src/rasterize.rs:305-337.  image: width*height*3, final_t and
 * n_contrib: width*height.  The binning arrays are those splat_bin produced. */
int splat_rasterize_forward(splat_ctx* ctx, const splat_proj* projs, uint64_t n,
                            const splat_binning* binning,
                            const splat_raster_settings* settings, float* image,
                            float* final_t, uint32_t* n_contrib);

/* rasterize_backward, src/rasterize.rs:459-505.  grads: n ProjGrad, abs_grads: 2n.
 * `ordered` != 0 sums each entry's per-pixel terms in the reference's row-major order
 * (bitwise-identical sums); 0 uses a tree reduction (same values up to rounding). */
int splat_rasterize_backward(splat_ctx* ctx, const splat_proj* projs, uint64_t n,
                             const splat_binning* binning,
                             const splat_raster_settings* settings, const float* final_t,
                             const uint32_t* n_contrib, const float* dimage, int ordered,
                             splat_proj_grad* grads, float* abs_grads);

/* render, src/render.rs:14-23.  image: width*height*3. */
int splat_render(splat_ctx* ctx, const float* gaussians, uint64_t n,
                 const splat_camera* cam, const splat_proj_opts* opts,
                 const float background[3], float* image);

/* to_rgb8, src/render.rs:26-31 (host arithmetic). */
int splat_to_rgb8(const float* image, uint64_t len, uint8_t* out);

/* photometric_loss, src/loss.rs:137-171.  pred, target, grad: width*height*3. */
int splat_photometric_loss(splat_ctx* ctx, const float* pred, const float* target,
                           uint32_t width, uint32_t height, float lambda_dssim,
                           float* loss, float* grad);

/* No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight evaluation loops, harnesses that forbid hipMalloc).
 * splat_photometric_loss on device buffers: pred, target, grad are device pointers of
 * width*height*3 floats; loss (device, one float) may be NULL.  Runs on the context's
 * stream and returns without waiting.  Nothing is allocated: the kernels' scratch is one
 * static device array of 73 MiB per GPU, held until the process ends, which limits images
 * to width*height <= 2097152 pixels (1920x1080 fits; larger: SPLAT_ERR_INVALID).  Every call
 * on one GPU shares that array, so calls that overlap on one GPU (from any context or
 * stream) race: serialise them.  Results equal splat_photometric_loss's bitwise. */
int splat_photometric_loss_device_noalloc(splat_ctx* ctx, const float* pred,
                                          const float* target, uint32_t width,
                                          uint32_t height, float lambda_dssim, float* loss,
                                          float* grad);

/* ssim_with_grad, src/loss.rs:75-134. */
int splat_ssim_with_grad(splat_ctx* ctx, const float* x, const float* y, uint32_t width,
                         uint32_t height, float* ssim, float* grad);

/* blur, src/loss.rs:31-72 (11-tap sigma 1.5 zero-padded separable blur). */
int splat_blur(splat_ctx* ctx, const float* src, uint32_t width, uint32_t height,
               uint32_t channels, float* dst);

/* psnr, src/loss.rs:175-190 (host, f64 sum). */
int splat_psnr(const float* pred, const float* target, uint64_t len, float* out);

/* Adam::step, src/adam.rs:77-161.  params, m, v updated in place; n records of 59.
 * state->step is incremented. */
int splat_adam_step(splat_ctx* ctx, float* params, const float* grads, float* m, float* v,
                    uint64_t n, splat_adam_state* state, const splat_learning_rates* lr);

/* DensifyStats::accumulate, src/densify.rs:36-60.  grad_accum, denom, max_radii: n,
 * updated in place.  abs_grads may be NULL (then ProjGrad.mean2d is used). */
int splat_stats_accumulate(splat_ctx* ctx, const splat_proj* projs,
                           const splat_proj_grad* grads, const float* abs_grads, uint64_t n,
                           uint32_t width, uint32_t height, float* grad_accum,
                           uint32_t* denom, uint32_t* max_radii);

/* densify_and_prune, src/densify.rs:125-221.  Inputs: n records of params, m, v and the
 * stats.  `normals` holds the standard normal draws the reference takes from its RNG for
 * split children (src/densify.rs:116-121,174-179), in draw order: split parent (index
 * order), child, axis x,y,z; its length must be at least 3 * split_children * (number of
 * split parents) -- pass `normals_len` and the call fails with SPLAT_ERR_CAPACITY (and
 * report->split set) when it is too short.  Outputs are written to the *_out arrays
 * (capacity `out_capacity` records); *n_out receives the new count.  The statistics of
 * the new set are zero (the reference resets them). */
int splat_densify_and_prune(splat_ctx* ctx, const float* params, const float* m,
                            const float* v, const float* grad_accum, const uint32_t* denom,
                            const uint32_t* max_radii, uint64_t n,
                            const splat_densify_config* cfg, float extent,
                            const float* normals, uint64_t normals_len, float* params_out,
                            float* m_out, float* v_out, uint64_t out_capacity,
                            uint64_t* n_out, splat_densify_report* report);

/* reset_opacity, src/densify.rs:224-230 with Adam::reset_opacity_state
 * (src/adam.rs:176-183).  In place on n records. */
int splat_reset_opacity(splat_ctx* ctx, float* params, float* m, float* v, uint64_t n,
                        float max_alpha);

/* ---- initialisation from points ----------------------------------------------------- */

/* mean_nn_dist2, src/model.rs:128-186: for n points (xyz binary32 triples) the mean
 * squared distance (binary64) to the three nearest other points; coincident points count
 * as zero-distance neighbours, fewer than three others use what exists, a lone point 0. */
int splat_mean_nn_dist2(splat_ctx* ctx, const float* xyz, uint64_t n, double* out);

/* Model::from_points, src/model.rs:77-104: n records (59 floats) with log-scale
 * ln(sqrt(max(d2, 1e-7))) on all axes, identity quaternion, opacity logit(0.1) and the
 * RGB8 colour in the DC coefficients (the model's SH degree starts at 0). */
int splat_from_points(splat_ctx* ctx, const float* xyz, const uint8_t* rgb, uint64_t n,
                      float* gaussians);

/* ---- PLY codec (host only, no context) ------------------------------------------- */

/* ply::write, src/ply.rs:19-69: the 3DGS binary little-endian float32 layout.  Writes
 * n records of SPLAT_NUM_PARAMS floats; SH coefficients above max_sh_degree (clamped to
 * 3) are not written.  On failure returns SPLAT_ERR_IO and copies a message into err. */
int splat_ply_write(const char* path, const float* gaussians, uint64_t n, int32_t max_sh_degree,
                    char* err, size_t err_len);

/* ply::read, src/ply.rs:73-181.  Property order is taken from the header.  *n and
 * *sh_degree receive the vertex count and the degree implied by the f_rest count; the
 * records are written only if capacity >= *n (otherwise SPLAT_ERR_CAPACITY), so a call
 * with gaussians = NULL and capacity = 0 sizes the output.  Malformed files give
 * SPLAT_ERR_FORMAT, unreadable ones SPLAT_ERR_IO, with the message in err. */
int splat_ply_read(const char* path, float* gaussians, uint64_t capacity, uint64_t* n,
                   int32_t* sh_degree, char* err, size_t err_len);

/* ---- COLMAP sparse models (src/colmap/{mod,binary,text}.rs, src/dataset.rs:83-129,213-237) --
 * Binary (.bin, preferred) or text (.txt) cameras / images / points3D in one directory;
 * 2D observations and point tracks are parsed past and not kept. */
typedef struct {
  uint32_t id;
  int32_t model;        /* COLMAP model id: 0 SIMPLE_PINHOLE, 1 PINHOLE, ... 11
                           RAD_TAN_THIN_PRISM_FISHEYE (src/colmap/mod.rs:17-63) */
  uint64_t width, height;
  uint32_t num_params;  /* the model's parameter count */
  uint32_t pad;
  double params[16];
} splat_colmap_camera;

typedef struct {
  uint32_t id;
  uint32_t camera_id;
  double qvec[4];       /* world-to-camera rotation (w, x, y, z) */
  double tvec[3];       /* world-to-camera translation */
  char name[256];       /* file name relative to the images directory, NUL-terminated */
} splat_colmap_image;

typedef struct {
  uint64_t id;
  double xyz[3];
  uint8_t rgb[3];
  uint8_t pad[5];
  double error;
} splat_colmap_point;

/* read_model, src/colmap/mod.rs:198-238: counts are always set; records are written only if
 * every capacity suffices (else SPLAT_ERR_CAPACITY).  Cameras and images in file order; a
 * camera id that occurs more than once is one record, with the values read last (the
 * reference's map by id), at the place of its first occurrence.  Missing files:
 * SPLAT_ERR_INVALID; unreadable: SPLAT_ERR_IO; malformed (unknown model, wrong parameter
 * count, truncated binary, implausible counts, bad tokens): SPLAT_ERR_FORMAT.  Image names
 * are limited to 255 bytes (the record's `name`; the reference has no limit): a longer
 * name is SPLAT_ERR_FORMAT. */
int splat_colmap_read(const char* dir, splat_colmap_camera* cameras, uint64_t camera_capacity,
                      uint64_t* num_cameras, splat_colmap_image* images, uint64_t image_capacity,
                      uint64_t* num_images, splat_colmap_point* points, uint64_t point_capacity,
                      uint64_t* num_points, char* err, size_t err_len);
/* write_model_bin / write_model_txt (src/colmap/binary.rs:165-224, text.rs:132-190):
 * cameras sorted by id, images without observations, points without tracks; text numbers
 * in the shortest form that reads back exactly, without exponent. */
int splat_colmap_write(const char* dir, int32_t binary, const splat_colmap_camera* cameras,
                       uint64_t num_cameras, const splat_colmap_image* images,
                       uint64_t num_images, const splat_colmap_point* points,
                       uint64_t num_points, char* err, size_t err_len);
/* camera_from_colmap, src/dataset.rs:83-129: PINHOLE / SIMPLE_PINHOLE intrinsics and the
 * normalised quaternion's rotation, in binary32; other models or an unknown camera id are
 * SPLAT_ERR_INVALID. */
int splat_camera_from_colmap(const splat_colmap_camera* cameras, uint64_t num_cameras,
                             const splat_colmap_image* image, splat_camera* out, char* err,
                             size_t err_len);
/* match_image_size, src/dataset.rs:213-237: the camera unchanged for its own size, scaled
 * (Camera::resized, src/camera.rs:48-62) for a uniformly scaled image (both axis factors
 * within one pixel), SPLAT_ERR_INVALID otherwise. */
int splat_camera_match_image_size(const splat_camera* camera, uint32_t width, uint32_t height,
                                  splat_camera* out);

/* ---- dataset loading (src/dataset.rs:131-205, 239-250) ---------------------------------
 * The reference decodes each image file with the `image` crate; here the caller decodes
 * and passes RGB8 pixels, and the library does the rest of dataset::load on the GPU. */

/* image::imageops::resize(.., FilterType::Lanczos3) (image 0.25) of an RGB8 image
 * (height x width x 3) to out_width x out_height, as dataset::load downscales views
 * (src/dataset.rs:176-180): separable, binary32 taps and sums, clamped and rounded to
 * nearest.  Host buffers; the resampling runs on the GPU. */
int splat_resize_rgb8(splat_ctx* ctx, const uint8_t* src, uint32_t width, uint32_t height,
                      uint32_t out_width, uint32_t out_height, uint8_t* dst);
/* The view size of a downscaled image, src/dataset.rs:172-175: each side divided by
 * `downscale` and rounded, at least 1.  downscale 0: SPLAT_ERR_INVALID (:136-138). */
int splat_downscaled_size(uint32_t width, uint32_t height, uint32_t downscale,
                          uint32_t* out_width, uint32_t* out_height);
/* split_eval, src/dataset.rs:239-250: is_test[i] = 1 for every hold-th view (i % hold ==
 * 0, none when hold is 0), the training set being the others, in order. */
int splat_split_eval(uint64_t n, uint32_t hold, uint8_t* is_test);
/* The view order of dataset::load (src/dataset.rs:141): images sorted by name, byte-wise
 * and stable. */
int splat_colmap_sort_images(splat_colmap_image* images, uint64_t n);

/* ---- checkpoints ---------------------------------------------------------------------
 * No counterpart in the reference, whose train loop saves only the model (.ply,
 * src/ply.rs:19-69, called from src/train.rs:275-279): a checkpoint adds the Adam moments
 * and step count and the densification statistics, so that a run resumed with
 * splat_trainer_restore continues exactly as the uninterrupted run would.  Layout (little
 * endian): "SPLATCKP", u32 version 1, i32 max_sh_degree, u64 n, u64 adam_step, then params,
 * m, v (n x 59 f32 each, the canonical row order), grad_accum (n f32), denom, max_radii
 * (n u32 each).  Null arrays are written as zeros (write) or skipped (read).  read: *n,
 * *adam_step and *max_sh_degree are set from the header; the arrays are filled only if
 * capacity >= *n (otherwise SPLAT_ERR_CAPACITY), so a call with null arrays and capacity 0
 * sizes the output.  A foreign file, an unknown version or a length that does not match
 * the header is SPLAT_ERR_FORMAT. */
int splat_checkpoint_write(const char* path, const float* gaussians, const float* m,
                           const float* v, const float* grad_accum, const uint32_t* denom,
                           const uint32_t* max_radii, uint64_t n, uint64_t adam_step,
                           int32_t max_sh_degree, char* err, size_t err_len);
int splat_checkpoint_read(const char* path, float* gaussians, float* m, float* v,
                          float* grad_accum, uint32_t* denom, uint32_t* max_radii,
                          uint64_t capacity, uint64_t* n, uint64_t* adam_step,
                          int32_t* max_sh_degree, char* err, size_t err_len);

/* ---- trainer: device-resident training (src/train.rs) ---------------------------- */

typedef struct splat_trainer splat_trainer;

/* Per-iteration inputs of one training step (the values train.rs computes on the host
 * each iteration, src/train.rs:196-247). */
typedef struct {
  int32_t sh_degree;            /* model.sh_degree (train.rs:197-202) */
  int32_t accumulate_stats;     /* densify && it < densify_until_iter (train.rs:228) */
  int32_t absgrad;              /* TrainConfig::absgrad */
  int32_t ordered_backward;     /* 1: the reference's summation order (bitwise), 0: fast */
  int32_t want_loss;            /* 1: compute the loss value (synchronises the stream) */
  int32_t antialias;            /* ProjOpts::antialias */
  float near_plane;             /* ProjOpts::near (0.2) */
  float lambda_dssim;           /* 0.2 */
  float background[3];
  splat_learning_rates lr;      /* with lr.mean = position_lr(it, extent) */
} splat_step_params;

typedef struct {
  float loss;              /* (1 - lambda) L1 + lambda (1 - SSIM), if want_loss */
  uint64_t num_entries;    /* tile entries of this view */
  uint64_t num_visible;    /* Gaussians with radius > 0 (if want_loss) */
} splat_step_result;

/* Create a trainer holding n Gaussians (AoS, 59 floats each) with room for `capacity`
 * (grown on demand by densification).  Adam state starts at zero, step 0
 * (Adam::new, src/adam.rs:65-74); statistics at zero (DensifyStats::new). */
int splat_trainer_create(splat_ctx* ctx, const float* gaussians, uint64_t n, uint64_t capacity,
                         splat_trainer** out);
int splat_trainer_destroy(splat_trainer* t);
/* Add a training view: camera and its RGB8 image (height x width x 3), kept in HBM as
 * bytes and converted in the loss kernel (View::to_real, src/dataset.rs:23-25). */
int splat_trainer_add_view(splat_trainer* t, const splat_camera* cam, const uint8_t* rgb,
                           uint32_t* index);
/* The per-image step of dataset::load on a trainer (src/dataset.rs:146-185): the image's
 * camera (camera_from_colmap at the model's resolution) matched to the decoded size
 * (match_image_size), the view size for `downscale`, the camera resized to it
 * (Camera::resized) and written to *used (optional), and the RGB8 pixels resized on the
 * GPU (Lanczos3, as splat_resize_rgb8; copied when downscale is 1) into the new view. */
int splat_trainer_add_view_downscaled(splat_trainer* t, const splat_camera* cam,
                                      const uint8_t* rgb, uint32_t width, uint32_t height,
                                      uint32_t downscale, splat_camera* used,
                                      uint32_t* index);
/* No counterpart in the reference (its views are plain host vectors): a view's camera and,
 * if rgb is given, its RGB8 pixels (height x width x 3; capacity in bytes, else
 * SPLAT_ERR_CAPACITY). */
int splat_trainer_view(splat_trainer* t, uint32_t view, splat_camera* cam, uint8_t* rgb,
                       uint64_t capacity);
/* One iteration on view `view`: project -> bin -> composite -> loss -> backward ->
 * statistics -> Adam (src/train.rs:211-247).  In graph mode (the default) the step is a
 * replay of a captured HIP graph and returns without waiting for the GPU when r is NULL
 * (at most two steps stay unchecked in flight); with r it waits and fills r.  Every other
 * trainer call waits for the steps in flight first. */
int splat_trainer_step(splat_trainer* t, uint32_t view, const splat_step_params* p,
                       splat_step_result* r);
/* No counterpart in the reference (its steps are plain function calls).  enabled = 1:
 * graph-captured steps (default); 0: kernels launched one by one, with a host read of the
 * entry count inside each step (the mode kernel timing uses).  Results are identical. */
int splat_trainer_set_graph(splat_trainer* t, int32_t enabled);
/* No counterpart in the reference.  Wait for the steps in flight (replaying any that
 * outgrew the entry arrays). */
int splat_trainer_sync(splat_trainer* t);
/* No counterpart in the reference.  Capacity of the tile-entry arrays that graph-captured
 * steps preallocate (default max(4 n, 65536), grown by 1.25x on demand); a caller that
 * knows its scenes can size them up front. */
int splat_trainer_reserve_entries(splat_trainer* t, uint64_t entries);
/* No counterpart in the reference.  Graph captures, graph replays, and steps replayed
 * because their tile entries outgrew the preallocated arrays. */
int splat_trainer_graph_stats(splat_trainer* t, uint64_t* captures, uint64_t* replays,
                              uint64_t* aborts);
/* densify_and_prune on the resident model (src/densify.rs:125-221); the normal draws
 * for split children come from a counter-based generator keyed by `seed`.  The new
 * parameter and moment planes are written into a second set, which is swapped in; the
 * old set is kept for the next call, so the trainer holds two sets (3 x 59 binary32 planes
 * each) at its grown capacity. */
int splat_trainer_densify(splat_trainer* t, const splat_densify_config* cfg, float extent,
                          uint64_t seed, splat_densify_report* report);
/* reset_opacity (src/densify.rs:224-230). */
int splat_trainer_reset_opacity(splat_trainer* t, float max_alpha);
/* Forward render of the resident model through view `view`'s camera (render.rs:14). */
int splat_trainer_render(splat_trainer* t, uint32_t view, int32_t sh_degree, int32_t antialias,
                         const float background[3], float* image);
/* TrainConfig, src/train.rs:20-109 (defaults from splat_train_config_default). */
typedef struct {
  uint32_t iterations;
  float position_lr_init, position_lr_final;
  uint32_t position_lr_max_steps;
  float feature_lr, opacity_lr, scaling_lr, rotation_lr;
  float lambda_dssim;
  int32_t densify;
  uint32_t densify_from_iter, densify_until_iter, densification_interval;
  uint32_t opacity_reset_interval;
  float densify_grad_threshold, percent_dense, min_opacity;
  int32_t has_max_screen_size;
  uint32_t max_screen_size;
  int32_t absgrad, antialias, white_background;
  int32_t sh_degree;
  uint32_t sh_degree_interval;
  float near_plane;
  uint64_t seed;
  const uint32_t* save_iterations; /* may be NULL */
  uint32_t num_save_iterations;
  uint32_t log_interval;
  int32_t ordered_backward; /* 1: reference summation order in the backward pass */
  /* No counterpart in the reference (its train() always starts at iteration 1).  0 or 1:
   * start at iteration 1.  K + 1: continue a run whose state after iteration K was saved
   * (splat_checkpoint_write from on_save) and restored (splat_trainer_restore with Adam
   * step K): the SH degree, view order, densification windows and opacity resets follow
   * the uninterrupted run's, so the result equals it bitwise.  The loss EMA of the
   * progress records restarts at the first resumed iteration.  Not available with an
   * attached scorer (its logged samples are not part of a checkpoint). */
  uint32_t start_iteration;
} splat_train_config;

/* Progress, src/train.rs:140-157. */
typedef struct {
  uint32_t iteration;
  float loss, ema_loss, psnr;
  uint64_t num_gaussians;
  double ms_per_iter;
  uint64_t cloned, split, pruned;
} splat_progress;

typedef void (*splat_progress_fn)(const splat_progress* p, void* user);
/* Return 0 to continue; non-zero aborts training with SPLAT_ERR_IO (on_save's Err). */
typedef int (*splat_save_fn)(uint32_t iteration, splat_trainer* t, void* user);

void splat_train_config_default(splat_train_config* cfg);

/* train, src/train.rs:161-306, on the trainer's views (all of them are training views).
 * The view order is a per-epoch Fisher-Yates shuffle drawn from SplitMix64(seed) and
 * popped from the back; split-child normals are drawn as in splat_trainer_densify with a
 * seed derived from (seed, iteration).  Returns SPLAT_ERR_INVALID for no views, zero
 * iterations or sh_degree > 3 (before any save). */
int splat_train(splat_trainer* t, float extent, const splat_train_config* cfg,
                splat_progress_fn on_progress, splat_save_fn on_save, void* user);

/* The view index splat_train uses at (1-based) iteration `it` for `num_views` views and
 * `seed` (the order is independent of everything else). */
uint32_t splat_train_view_at(uint64_t seed, uint32_t num_views, uint32_t it);

/* Current Gaussian count, Adam step count. */
int splat_trainer_info(splat_trainer* t, uint64_t* n, uint64_t* adam_step);
/* Copy the model (and optionally the Adam moments and statistics) to the host. */
int splat_trainer_download(splat_trainer* t, float* gaussians, float* m, float* v,
                           float* grad_accum, uint32_t* denom, uint32_t* max_radii,
                           uint64_t capacity);
/* evaluate, src/train.rs:309-321, on the trainer's views: each listed view rendered with
 * the given SH degree, antialias flag and background, its PSNR against the view's target
 * (psnr, src/loss.rs:175-190: binary64 mean squared error of the clamped values, binary32
 * result; +inf for a perfect match) into psnr[q] (optional) and their binary32 mean, summed
 * in list order, into *mean (NaN for an empty list). */
int splat_trainer_evaluate(splat_trainer* t, const uint32_t* views, uint32_t count,
                           int32_t sh_degree, int32_t antialias, const float* background,
                           float* psnr, float* mean);
/* Replace the trainer's model and training state with a saved one (the inverse of
 * splat_trainer_download; no counterpart in the reference): n rows of parameters, Adam
 * moments m, v (null: zeros), densification statistics (null: zeros) and the Adam step
 * count.  Capacity grows to n as needed; views and launch shapes are kept.  With learned
 * densification attached, its extra statistics restart from zero and every row counts as
 * an original (no lineage). */
int splat_trainer_restore(splat_trainer* t, const float* gaussians, const float* m,
                          const float* v, const float* grad_accum, const uint32_t* denom,
                          const uint32_t* max_radii, uint64_t n, uint64_t adam_step);

/* ---- batched training ----------------------------------------------------------------
 * No counterpart in the reference (one view of one scene per iteration, train.rs:211-247).
 * Many independent scenes and many views of uneven resolution train in one step: each
 * stage (projection, binning, compositing, loss, backward, Adam) is one launch over every
 * (scene, view) pair, and the step is captured in a HIP graph.  A step renders each listed
 * view of its scene with the reference's forward pass, takes the reference's loss and
 * gradient per view, sums each scene's parameter gradients over its views in list order,
 * and applies one Adam step to every scene (its own position learning rate).  One view of
 * one scene is exactly the reference's iteration.  At most 64 scenes; scenes are added
 * before the first step. */
typedef struct splat_batch splat_batch;

typedef struct {
  int32_t sh_degree;
  int32_t ordered_backward;  /* 1: the reference's summation order (bitwise) */
  int32_t antialias;
  float near_plane;          /* 0.2 */
  float lambda_dssim;        /* 0.2 */
  float background[3];
  splat_learning_rates lr;   /* lr.mean unused: per-scene rates are passed to the step */
} splat_batch_step_params;

/* Launch shapes of the compositing kernels (results do not depend on them): entries per
 * LDS batch (forward 32/64/128/256, tree-reduced backward 16/32), persistent launches
 * (a fixed grid of per_cu workgroups per compute unit taking tiles from a counter), and
 * whether each 16-lane row of a tree-reduced backward wave walks its own entry list
 * (bwd_row_lists = 1) or the wave walks every entry that reaches one of its rows (0). */
typedef struct {
  int32_t fwd_batch, fwd_persistent, fwd_per_cu;
  int32_t bwd_batch, bwd_persistent, bwd_per_cu;
  float fwd_ms, bwd_ms;                  /* best measured times */
  float default_fwd_ms, default_bwd_ms;  /* times of the default shape (64 / 32, one
                                            workgroup per tile, row lists) */
  uint32_t candidates;
  int32_t bwd_row_lists;
} splat_tune_result;

/* No counterpart in the reference.  Autotuner of the single-view trainer: times every
 * candidate launch shape of the compositing kernels on view `view` (the model is not
 * changed) and keeps the fastest for later steps; set_shapes pins given shapes. */
int splat_trainer_tune(splat_trainer* t, uint32_t view, const splat_step_params* p,
                       splat_tune_result* out);
int splat_trainer_set_shapes(splat_trainer* t, const splat_tune_result* shapes);
/* No counterpart in the reference (its design notes quote per-pair costs, not counts).  The
 * compositing half of a step on view `view` (projection, binning, forward, loss and the
 * tree-reduced backward; the model is not changed) through counting instantiations of the
 * compositing kernels.  counts[0..4]: forward (entry, wave) visits, visits in which some
 * pixel evaluates the Gaussian, evaluated (pixel, Gaussian) pairs (the exponential is
 * computed: 0 <= sigma <= sigma_max), passing pairs (alpha >= 1/255 on a pixel still
 * compositing), composited pairs; counts[5..8]: the backward walk's visits, active visits,
 * evaluated and passing pairs.  Synchronises. */
int splat_trainer_pair_counts(splat_trainer* t, uint32_t view, const splat_step_params* p,
                              uint64_t counts[9]);

int splat_batch_create(splat_ctx* ctx, splat_batch** out);
int splat_batch_destroy(splat_batch* b);
int splat_batch_add_scene(splat_batch* b, const float* gaussians, uint64_t n, uint32_t* scene);
int splat_batch_add_view(splat_batch* b, uint32_t scene, const splat_camera* cam,
                         const uint8_t* rgb, uint32_t* view);
/* One step over views[0..nviews) (each at most once); lr_mean[s] is scene s's position
 * learning rate; losses (optional, host) receives each listed view's loss. */
int splat_batch_step(splat_batch* b, const uint32_t* views, uint32_t nviews,
                     const splat_batch_step_params* p, const float* lr_mean, float* losses);
/* Autotuner: times every candidate launch shape of the compositing kernels on this batch
 * (the model is not changed) and keeps the fastest for later steps. */
int splat_batch_tune(splat_batch* b, const uint32_t* views, uint32_t nviews,
                     const splat_batch_step_params* p, splat_tune_result* out);
/* Use the given launch shapes (the shape fields of a tune result). */
int splat_batch_set_shapes(splat_batch* b, const splat_tune_result* shapes);
int splat_batch_download(splat_batch* b, uint32_t scene, float* gaussians, float* m, float* v,
                         uint64_t capacity);
/* Parameter rows (all scenes, padded), Adam steps, graph captures / replays, and steps
 * run again because their tile entries outgrew the preallocated arrays. */
int splat_batch_info(splat_batch* b, uint64_t* rows, uint64_t* adam_step, uint64_t* captures,
                     uint64_t* replays, uint64_t* aborts);

/* ---- learned densification scorer --------------------------------------------------
 * No counterpart in the reference, whose densification is the threshold heuristic
 * (src/densify.rs:125-221).  A small MLP scores each Gaussian from 12 statistics:
 * 12 -> 64 -> 64 -> 64 (tanh GELU) -> 3 logits (clone, split, prune), bfloat16 weights and
 * activations on the matrix cores with binary32 accumulation, binary32 master weights.
 * Parameters (SPLAT_SCORER_PARAMS floats): W1[64][12] b1[64] W2[64][64] b2[64]
 * W3[64][64] b3[64] W4[3][64] b4[3].  Labels are bit masks (1 clone, 2 split, 4 prune);
 * the training loss is the weighted binary cross-entropy of the three logits, averaged over
 * the minibatch.  For callers that want a learned densification policy. */
#define SPLAT_SCORER_PARAMS 9347
#define SPLAT_SCORER_FEATURES 12

typedef struct splat_scorer splat_scorer;

typedef struct {
  uint32_t epochs;   /* passes over the samples */
  uint32_t batch;    /* minibatch size (the last partial minibatch of an epoch is skipped) */
  float lr;          /* Adam learning rate (betas 0.9, 0.999, eps 1e-8) */
  uint32_t pad_;
  uint64_t seed;     /* per-epoch sample permutation */
} splat_scorer_fit_config;

/* params: SPLAT_SCORER_PARAMS floats, or NULL for a seeded N(0, 1/fan_in) initialisation. */
int splat_scorer_create(splat_ctx* ctx, const float* params, uint64_t seed, splat_scorer** out);
int splat_scorer_destroy(splat_scorer* s);
int splat_scorer_get_params(splat_scorer* s, float* params);
/* Replaces the weights and resets the scorer's Adam state. */
int splat_scorer_set_params(splat_scorer* s, const float* params);
/* logits[n][3] of features[n][12] (host arrays). */
int splat_scorer_forward(splat_scorer* s, const float* features, uint64_t n, float* logits);
/* Gradient (SPLAT_SCORER_PARAMS floats) and value of the loss over the n samples. */
int splat_scorer_loss_grad(splat_scorer* s, const float* features, const uint8_t* labels,
                           const float* weights, uint64_t n, float* grad, float* loss);
void splat_scorer_fit_config_default(splat_scorer_fit_config* cfg);
/* Minibatch Adam over the samples; *loss = mean minibatch loss of the last epoch. */
int splat_scorer_fit(splat_scorer* s, const float* features, const uint8_t* labels,
                     const float* weights, uint64_t n, const splat_scorer_fit_config* cfg,
                     float* loss);

/* Learned densification inside splat_train (no counterpart in the reference).
 * mode 1 (learn): the reference's decisions are applied unchanged; a sample of them is
 *   logged at each densification event with the 12 features, labelled `outcome_delay`
 *   iterations later, weighted by the PSNR change over that window, and the scorer is fit
 *   online on every labelled sample.  A clone/split is positive, with credit_views > 0,
 *   when the Gaussians it added (a clone's copy, a split's children, and their
 *   descendants) hold at least credit_ratio times the median Gaussian's share of the
 *   images of credit_views evenly spaced views (sum of alpha T over the pixels); with
 *   credit_views = 0, while any descendant is alive with an opacity of at least
 *   useful_ratio times its parent's at the decision and at least useful_opacity.
 * mode 2 (learned): the scorer's logits decide clone / split / prune at each event; the
 *   reference's prune rules still apply.
 * Features (binary32, per Gaussian): log2(mean grad / thr + 1e-6) / 8, log2(max grad / thr
 * + 1e-6) / 8, strong steps / max(visible steps, 1), (log-scale_k - ln(percent_dense *
 * extent)) / 4 for k = 0..2, sigmoid(opacity), log2(1 + max radius) / 8, log2(1 + visible
 * steps) / 8, DC luma (0.299 r + 0.587 g + 0.114 b of SH_C0 dc + 0.5), log2(1 + age) / 12,
 * (max log-scale - ln(percent_dense * extent)) / 4; "strong" steps are those whose
 * screen-space gradient norm reached the densify threshold. */
typedef struct {
  int32_t mode;                /* 1 learn, 2 learned */
  uint32_t samples_per_event;  /* logged decisions per densification event */
  uint32_t max_events;         /* sample store capacity, in events */
  uint32_t outcome_delay;      /* iterations before a logged decision is labelled */
  float threshold;             /* mode 2: clone/split probability threshold */
  float prune_threshold;       /* mode 2: prune probability threshold */
  uint32_t eval_views;         /* views whose mean PSNR weights the samples */
  float useful_opacity;        /* floor of the opacity a descendant needs to count as useful
                                  (any survivor counts after an opacity reset in the
                                  window) */
  float useful_ratio;          /* ... and at least this fraction of its parent's opacity at
                                  the decision */
  uint32_t credit_views;       /* > 0: label by credit instead: the image share (sum of
                                  alpha T over the pixels of this many evenly spaced views)
                                  of the Gaussians the decision added, at the outcome */
  float credit_ratio;          /* a clone / split is useful when its credit reaches this
                                  fraction of the median Gaussian's share at the outcome */
  float share_prune;           /* > 0, mode 1: label as prune the logged rows whose share of
                                  all views' images is below this fraction of the median
                                  Gaussian's (events from share_prune_from on); mode 2:
                                  prune those rows (a measured rule, events from
                                  share_prune_from on) */
  uint32_t share_prune_from;   /* mode 2: the scorer's prune head (and the rule above)
                                  applies at the events from this iteration on */
  uint32_t pad_;
  splat_scorer_fit_config fit; /* each online update */
} splat_learn_config;

typedef struct {
  uint64_t logged, resolved, updates, positives, events;
  float last_loss;
  uint32_t pad_;
} splat_learn_stats;

void splat_learn_config_default(splat_learn_config* cfg);
/* Attach a scorer (s = NULL detaches).  Turns on the extra per-Gaussian statistics and
 * lineage planes; splat_train then densifies through the scorer's mode. */
int splat_trainer_attach_scorer(splat_trainer* t, splat_scorer* s, const splat_learn_config* cfg);
int splat_trainer_learn_stats(splat_trainer* t, splat_learn_stats* out);
/* The scorer's features of the resident model ([n][12]) at `iteration`. */
int splat_trainer_features(splat_trainer* t, float extent, float grad_threshold,
                           float percent_dense, uint32_t iteration, float* out);
/* The extra statistics (largest per-step gradient norm, strong steps) and lineage planes
 * (birth iteration, logged decision the row descends from or 0xffffffff). */
int splat_trainer_download_lineage(splat_trainer* t, float* max_grad, uint32_t* strong,
                                   uint32_t* birth, uint32_t* origin, uint64_t capacity);
/* The labelled samples of mode 1 (features [k][12], label bits, weights); *count = k. */
int splat_trainer_learn_samples(splat_trainer* t, float* features, uint8_t* labels,
                                float* weights, uint64_t capacity, uint64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* SPLAT_H */
