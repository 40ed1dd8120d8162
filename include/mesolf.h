/*
 * mesolf.h -- C ABI of libmesolf, an MI355X (gfx950) implementation of
 * MesoLF's phase-space Richardson-Lucy light-field deconvolution and of its
 * CUDA MEX operator.
 *
 * Arrays cross this boundary as plain pointers in MATLAB column-major order
 * (the first index is contiguous).  No entry point takes or returns a
 * framework type.  Every call is thread safe; calls on one handle are
 * serialised by that handle's mutex.
 *
 * Drop-in targets in the reference (vazirilab/MesoLF):
 *   mlf_recon_phase_space       <- src/recon/recon_theory_phase_space_LFM_with_filter.m:10
 *   mlf_lfm_convolution_vec     <- src/cuda/mex_lfm_convolution_vec.cu:276 (mexFunction)
 * The handle API (context / psf / patch / batch / server) has no single
 * counterpart: it exposes the stages of recon:10-151 so that the PSF and the
 * volumes stay resident on the device between calls.
 */
#ifndef MESOLF_H
#define MESOLF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MLF_ABI_VERSION 2

typedef enum {
  MLF_OK = 0,
  MLF_ERR_INVALID_ARGUMENT = 1, /* maps to the MEX error id parallel:gpu:mexGPUExample:InvalidInput */
  MLF_ERR_UNSUPPORTED = 2,      /* valid in the reference, not supported here (message says why) */
  MLF_ERR_DEVICE = 3,           /* HIP runtime error (message carries hipGetErrorString) */
  MLF_ERR_OUT_OF_MEMORY = 4,
  MLF_ERR_BUDGET = 5,           /* request refused by the HBM admission budget */
  MLF_ERR_INTERNAL = 6,
  MLF_ERR_CANCELLED = 7,        /* queued request dropped before it ran (server destroyed) */
  MLF_ERR_IO = 8                /* file missing, unreadable, or not a well-formed MAT-file */
} mlf_status;

typedef enum {
  MLF_PRECISION_FP32 = 0, /* FP32 operands, FP32 accumulation (the drop-in)            */
  MLF_PRECISION_BF16 = 1, /* BF16 MFMA operands, FP32 accumulation, FP32 update        */
  MLF_PRECISION_FP16 = 2  /* FP16 MFMA operands (PSF pre-scaled per view and plane)    */
} mlf_precision;

typedef struct mlf_context_s* mlf_context;
typedef struct mlf_psf_s* mlf_psf;
typedef struct mlf_patch_s* mlf_patch;

/* Per-call statistics of a reconstruction (recon:99,144-145 prints the
 * per-iteration time; this struct returns it). */
typedef struct {
  int32_t iterations;
  int32_t views_per_iteration;
  double total_seconds;
  double iteration_seconds[64];   /* first 64 iterations */
  double relative_change[64];     /* ||X_k - X_{k-1}|| / ||X_{k-1}||, FP64, if requested */
} mlf_recon_stats;

/* ------------------------------------------------------------------ errors */
/* Thread-local message of the last failing call on this thread. */
const char* mlf_last_error(void);
const char* mlf_status_string(mlf_status s);
int32_t mlf_abi_version(void);

/* ----------------------------------------------------------------- context */
/* One context per device (device is 0-based; the MATLAB gateway converts the
 * 1-based gpu_id of recon:76).  Owns one HIP stream. */
mlf_status mlf_context_create(int32_t device, mlf_context* out);
mlf_status mlf_context_destroy(mlf_context ctx);
mlf_status mlf_context_synchronize(mlf_context ctx);
/* Free/total device memory as seen by the context. */
mlf_status mlf_context_memory(mlf_context ctx, uint64_t* free_bytes, uint64_t* total_bytes);

/* --------------------------------------------------------------------- PSF */
/* From the raw LFM PSFs exactly as recon:13-40 consumes them:
 * H, Ht are [S, S, N, N, Z] single (column-major).  Performs the per-depth
 * normalisation (recon:13-16), gen_phasespace_PSF_theory (crop + de-shift +
 * phase-space permutation), the view weights / aperture (recon:30-40) and the
 * spiral visiting order (recon:95-105), then uploads compact tables.
 * N must be 15 (the phase count of the MEX and of MesoLF's optics). */
mlf_status mlf_psf_create_dense(mlf_context ctx, const float* H, const float* Ht,
                                int64_t S, int64_t N, int64_t Z, mlf_psf* out);

/* From compact phase-space tables (already normalised, views already in
 * visiting order).  For view k and plane z:
 *   fwd block: (2*rf[z]+1)^2 * N*N floats, index ((d2+rf)*(2rf+1) + (d1+rf))*N*N + s2*N + s1
 *              y[i] = sum_z sum_{L,s} X[N*L+s, z] * fwd[i-L+rf, s]
 *   bwd block: same with rb[z]:   B[N*a+s, z] = sum_i R[i] * bwd[a-i+rb, s]
 * Blocks are concatenated view-major, then plane.  bwd may be NULL: then the
 * exact adjoint of fwd is used (rb = rf, bwd[d] = fwd[-d]). */
mlf_status mlf_psf_create_compact(mlf_context ctx, int64_t N, int64_t Z, int64_t nviews,
                                  const int32_t* view_uv, const double* view_weight,
                                  const int32_t* rf, const float* fwd,
                                  const int32_t* rb, const float* bwd, mlf_psf* out);
/* ---------------------------------------------------- MAT-file (-v7.3) reader */
/* MesoLF saves its PSFs with save(..., '-v7.3') and loads them whole with
 * load(SI.psffile) (mesolf.m:20-21): HDF5 behind a 512-byte MATLAB user
 * block.  These entry points read the real numeric and logical variables of
 * such a file (double, single, [u]int8..64, logical, char codes; any byte
 * order; contiguous, compact or chunked with deflate / shuffle / fletcher32)
 * in MATLAB's column-major order.  Complex, sparse, cell, struct and object
 * variables are reported by mlf_mat_info and refused by the readers with
 * MLF_ERR_UNSUPPORTED.  A malformed file gives MLF_ERR_IO. */
typedef struct mlf_mat_s* mlf_mat;
#define MLF_MAT_MAX_DIMS 32
typedef struct {
  char matlab_class[32];           /* MATLAB_class attribute ("double", "single", "struct", ...) */
  int32_t ndim;                    /* MATLAB rank (>= 2) */
  int64_t dims[MLF_MAT_MAX_DIMS];  /* MATLAB dims, first index contiguous */
  int64_t numel;                   /* 0 for an empty variable */
  int32_t numeric;                 /* 1 when mlf_mat_read_f32 / _f64 can read it */
  int32_t stored_bytes;            /* bytes per element in the file */
} mlf_mat_var;
mlf_status mlf_mat_open(const char* path, mlf_mat* out);
mlf_status mlf_mat_close(mlf_mat m);
/* Top-level variables: count, and the name of variable i (cap bytes incl. NUL). */
mlf_status mlf_mat_count(mlf_mat m, int32_t* n);
mlf_status mlf_mat_name(mlf_mat m, int32_t i, char* name, size_t cap);
/* name may be a '/'-separated path into a struct (a group). */
mlf_status mlf_mat_info(mlf_mat m, const char* name, mlf_mat_var* info);
/* Convert the variable to float / double; numel must equal its element count. */
mlf_status mlf_mat_read_f32(mlf_mat m, const char* name, float* out, int64_t numel);
mlf_status mlf_mat_read_f64(mlf_mat m, const char* name, double* out, int64_t numel);

typedef struct {
  int64_t S, N, Z;     /* H is [S, S, N, N, Z] */
  double x_pixel_um;   /* |x1objspace(2) - x1objspace(1)| * 1e6 (mesolf.m:231), NaN if absent */
  double z_pixel_um;   /* |x3objspace(2) - x3objspace(1)| * 1e6 (mesolf.m:232), NaN if absent */
} mlf_psf_file_info;
/* load(SI.psffile) (mesolf.m:20-21) followed by mlf_psf_create_dense on its
 * H and Ht: the PSF is read once and stays resident on the device, where the
 * reference copies H(:,:,u,v,:) and Ht(:,:,u,v,:) to the GPU for every view
 * of every iteration (recon:109, 121, 123).  Nnum, when present, must equal
 * N.  meta may be NULL. */
mlf_status mlf_psf_load_mat(mlf_context ctx, const char* path, mlf_psf* out, mlf_psf_file_info* meta);

/* nviews, Z, and per-plane radii (arrays of length Z may be NULL). */
mlf_status mlf_psf_info(mlf_psf psf, int64_t* nviews, int64_t* Z, int32_t* rf, int32_t* rb);
/* Active views in visiting order: uv[2*k], uv[2*k+1] (0-based), weight[k]. */
mlf_status mlf_psf_views(mlf_psf psf, int32_t* uv, double* weight);
mlf_status mlf_psf_destroy(mlf_psf psf);

/* ------------------------------------------------------------------- patch */
/* One reconstruction problem: n1 x n2 lenslets (image 15*n1 x 15*n2 px),
 * all planes of the PSF.  Holds the volume, the image and work buffers. */
mlf_status mlf_patch_create(mlf_psf psf, int64_t n1, int64_t n2, mlf_precision prec, mlf_patch* out);
mlf_status mlf_patch_destroy(mlf_patch p);
/* LFIMG [15*n1, 15*n2] single, host memory. */
mlf_status mlf_patch_set_image(mlf_patch p, const float* lfimg);
/* X = sum(LFIMG) / (I1*I2*Z) everywhere (recon:72-73). */
mlf_status mlf_patch_init_volume(mlf_patch p);
/* X [15*n1, 15*n2, Z] single, host memory (warm start). */
mlf_status mlf_patch_set_volume(mlf_patch p, const float* X);
mlf_status mlf_patch_get_volume(mlf_patch p, float* X);
mlf_status mlf_patch_get_volume_f64(mlf_patch p, double* X);
/* Run RL iterations (recon:98-146) asynchronously on the context stream.
 * views_limit <= 0 visits every active view; otherwise the first views_limit.
 * filter != 0 applies sample_filter at the end of each iteration (recon:142). */
mlf_status mlf_patch_run(mlf_patch p, int32_t iterations, int32_t views_limit, int32_t filter);
/* Forward projection of the current volume for active view k (recon:109,
 * 243-249): y [n1, n2] single, host.  Synchronous. */
mlf_status mlf_patch_forward_view(mlf_patch p, int32_t view, float* y);
/* Forward projection of every active view interleaved into a sensor image
 * LF[u + 15*i1, v + 15*i2] (zero at inactive views), host [15*n1, 15*n2]. */
mlf_status mlf_patch_project_all(mlf_patch p, float* sensor);
/* sample_filter (recon:222-240) on the current volume, FP64 plane energies. */
mlf_status mlf_patch_sample_filter(mlf_patch p);
/* Replaces the post-reconstruction lines of LFM_reconstruction_module.m:45-53
 * (top-layer balance: plane 1 times max(plane 2) / max(plane 1); each plane
 * divided by its max) and :60-62 (the stack divided by its max; the TIFF
 * image im2uint16(stack / max(stack))), then mesolf.m:101-102 (keep planes
 * ceil(Z/2) - floor(valid_range/2) .. ceil(Z/2) + floor(valid_range/2),
 * 1-based).  FP64 arithmetic on the patch's volume in HBM, in MATLAB's
 * evaluation order; maxima omit NaN.  The volume itself is not changed.
 *   valid_range: positive and odd (mesolf_set_params.m:138), the planes within
 *                1..Z; 0 keeps all Z planes (the module's output before the crop)
 *   out:  double [15*n1, 15*n2, valid_range or Z] column-major, host, or NULL
 *   tiff: uint16 [15*n1, 15*n2, Z] column-major, host, or NULL
 * Needs Z >= 2 (plane 2 is read).  The optional vessel mask
 * (LFM_reconstruction_module.m:56-58, off by default) is not applied.
 * Synchronous. */
mlf_status mlf_patch_postprocess(mlf_patch p, int32_t valid_range, double* out, uint16_t* tiff);
/* Time of the last mlf_patch_run on the device, milliseconds (HIP events). */
mlf_status mlf_patch_last_run_ms(mlf_patch p, double* ms);
/* Per-kernel timing: while enabled, the launches of mlf_patch_run are
 * bracketed by HIP events on the context stream: enable = 1 times every view
 * step, enable = N > 1 every N-th view step (fewer events in the stream; the
 * sample filter is always timed), 0 turns timing off.  mlf_patch_kernel_times
 * returns, per kernel class, the summed milliseconds and the launch count
 * since the previous read, then resets them.  Classes (MLF_KERNEL_CLASSES):
 * 0 forward projection, 1 ratio, 2 backprojection + update, 3 sample_filter,
 * 4 post-processing (mlf_patch_postprocess, one event pair per call). */
#define MLF_KERNEL_CLASSES 5
mlf_status mlf_patch_set_timing(mlf_patch p, int32_t enable);
mlf_status mlf_patch_kernel_times(mlf_patch p, double* total_ms, int64_t* launches);

/* ----------------------------------------------------------------- batches */
/* No counterpart in the reference, which reconstructs one patch per MATLAB
 * worker (LFM_reconstruction_module.m:26-37 inside mesolf.m:55-61's parfor).
 * A batch runs RL iterations (recon:98-146) for up to `capacity` patches of
 * one PSF and precision in shared launches: each view-step kernel covers the
 * work entries of every attached patch, whatever its lenslet grid, and one
 * iteration's 3 x nviews launches are replayed from one captured hipGraph.
 * A patch's result does not depend on the batch it ran in (the same work
 * entries produce the same bits for the same tile shape; the first attached
 * patch sets the batch's shape and later ones are switched to it).  Patches
 * keep their own image and volume; attach them after set_image and
 * init_volume / set_volume.  A destroyed
 * patch is detached first.  Calls on one batch are serialized; running a
 * batch and one of its patches from two threads at once is not supported. */
typedef struct mlf_batch_s* mlf_batch;
mlf_status mlf_batch_create(mlf_psf psf, mlf_precision prec, int32_t capacity, mlf_batch* out);
mlf_status mlf_batch_destroy(mlf_batch b);
mlf_status mlf_batch_attach(mlf_batch b, mlf_patch p);
mlf_status mlf_batch_detach(mlf_batch b, mlf_patch p);
mlf_status mlf_batch_size(mlf_batch b, int32_t* attached);
/* enable != 0 (default): replay a captured graph per iteration. */
mlf_status mlf_batch_set_graph(mlf_batch b, int32_t enable);
/* `iterations` RL iterations of every attached patch, asynchronously on the
 * context stream; filter != 0 applies sample_filter after each (recon:142). */
mlf_status mlf_batch_run(mlf_batch b, int32_t iterations, int32_t filter);
mlf_status mlf_batch_last_run_ms(mlf_batch b, double* ms);

/* -------------------------------------------------------------- autotuning */
/* No counterpart in the reference.  The view-step kernels are built in
 * several tile shapes: the forward projection's input columns per wave
 * (BF16: 4 or 2, FP32: 2 or 1; a workgroup covers 16 x 4*columns lenslets),
 * the backprojection's 16-phase tiles per phase group (5 or 3, and 4 for
 * BF16: 3, 5 or 4 groups of 80, 48 or 64 phases cover the 225 phases) and its
 * phase groups per work entry (0: all groups
 * of a lenslet tile in one entry, which builds the tile's ratio window once;
 * 1: one group per entry).
 * Autotuning times every shape on `views` view steps of the patch (or of all
 * patches of the batch, in its shared launches), restores the volumes, and
 * keeps the fastest.  Results are cached per precision, lenslet grid(s) and
 * PSF radii for the process and, if the environment variable MLF_TUNE_CACHE
 * names a file, across processes: patches created later with a cached key
 * start with the tuned shape.  Image and volume must be set. */
typedef struct {
  int32_t fwd_columns_per_wave;
  int32_t bwd_phase_tiles;
  int32_t bwd_groups_per_entry;
  int32_t bwd_waves_per_simd;  /* 16-bit backprojection: 2, or 3 with bwd_phase_tiles = 3 */
  double ms_per_view;  /* measured view step (forward + ratio + backprojection) */
} mlf_tile_shape;
mlf_status mlf_patch_autotune(mlf_patch p, int32_t views, mlf_tile_shape* best);
mlf_status mlf_patch_tile_shape(mlf_patch p, mlf_tile_shape* current);
/* Select a shape explicitly (one of the built ones; ms_per_view is ignored). */
mlf_status mlf_patch_set_tile_shape(mlf_patch p, const mlf_tile_shape* shape);
/* Tunes the shape shared by every attached patch (key: all their grids). */
mlf_status mlf_batch_autotune(mlf_batch b, int32_t views, mlf_tile_shape* best);

/* ----------------------------------------------------------------- serving */
/* No counterpart in the reference.  A request loop for frame x patch
 * deconvolutions: requests are queued as they arrive, admitted into batch
 * slots at RL-iteration boundaries while other requests are mid-flight
 * (continuous batching), and run `iterations` RL iterations each.  A patch id
 * keeps its volume resident between requests, so the next frame of that
 * patch starts warm from the previous result; its first request (or the first
 * after an eviction) starts from recon:72-73.  Resident per-patch state is
 * capped by `hbm_budget_bytes`: a request whose patch is not resident is
 * admitted only when its state fits, evicting idle patches (least recently
 * used) if needed.  Frames of one patch run in submission order. */
typedef struct mlf_server_s* mlf_server;
typedef struct {
  int32_t max_slots;          /* patches in flight at once (batch capacity) */
  int32_t iterations;         /* RL iterations per request */
  int32_t filter;             /* sample_filter after each iteration */
  int32_t use_graph;          /* replay captured graphs */
  uint64_t hbm_budget_bytes;  /* cap on resident per-patch device memory */
} mlf_server_config;
typedef struct {
  int64_t submitted, completed, pending, in_flight;
  int64_t resident_patches, evictions, cold_starts, iterations_run;
  uint64_t resident_bytes, peak_resident_bytes;
  double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> volume ready, completed requests */
  double volumes_per_s;                   /* completed / (last completion - first submit) */
  double mean_slots;                      /* attached patches per iteration, averaged */
} mlf_server_stats;
mlf_status mlf_server_create(mlf_psf psf, mlf_precision prec, const mlf_server_config* cfg,
                             mlf_server* out);
/* Queue one frame of patch `patch_id` (grid n1 x n2, fixed per id): lfimg
 * [15*n1, 15*n2] single is copied before return.  volume_out, if not NULL,
 * receives the volume [15*n1, 15*n2, Z] single when the request completes and
 * must stay valid until then.  Returns a ticket for mlf_server_wait. */
mlf_status mlf_server_submit(mlf_server s, int32_t patch_id, int64_t n1, int64_t n2,
                             const float* lfimg, float* volume_out, uint64_t* ticket);
/* Block until the request completes; latency_ms (may be NULL) = submit to
 * completion. */
mlf_status mlf_server_wait(mlf_server s, uint64_t ticket, double* latency_ms);
/* Block until every submitted request has completed. */
mlf_status mlf_server_drain(mlf_server s);
mlf_status mlf_server_get_stats(mlf_server s, mlf_server_stats* st);
/* Requests still queued are cancelled (their wait returns MLF_ERR_CANCELLED);
 * those in flight complete first. */
mlf_status mlf_server_destroy(mlf_server s);
/* Failure handling.  mlf_server_submit refuses (MLF_ERR_BUDGET) a patch whose
 * state alone exceeds hbm_budget_bytes.  A request that fails on the device
 * completes with that error, its patch's resident state is dropped (the next
 * frame of the patch starts cold) and the loop keeps serving the others.
 * mlf_server_inject_fault forces such a failure from the host, for tests of
 * the failure paths: MLF_FAULT_ADMIT fails the admission of the next request
 * of patch_id; MLF_FAULT_ITERATION fails the next batched iteration as a
 * device error there would (every request in flight fails; patch_id is
 * ignored).  No kernel is made to fault. */
#define MLF_FAULT_ADMIT 1
#define MLF_FAULT_ITERATION 2
mlf_status mlf_server_inject_fault(mlf_server s, int32_t kind, int32_t patch_id);

/* ------------------------------------------------------------ drop-ins */
/* recon_theory_phase_space_LFM_with_filter(LFIMG,H,Ht,maxIter,gpuFlag,gpu_id)
 * (recon:10).  LFIMG [I1, I2] single; H, Ht [S,S,N,N,Z] single; X_out
 * [I1, I2, Z] double (the reference returns a gathered double array, recon:147).
 * device is 0-based.  gpu_flag == 0 returns MLF_ERR_UNSUPPORTED: the
 * reference CPU branch cannot run as written (undefined Xguessv at recon:112).
 * stats may be NULL. */
mlf_status mlf_recon_phase_space(const float* lfimg, int64_t I1, int64_t I2,
                                 const float* H, const float* Ht, int64_t S, int64_t N, int64_t Z,
                                 int32_t max_iter, int32_t gpu_flag, int32_t device,
                                 mlf_precision prec, double* X_out, mlf_recon_stats* stats);

/* mex_lfm_convolution_vec(bitmap, kernels) (mex:276-421):
 *   out[x,y,n] = sum_{kx,ky<kw} bitmap[x-kw/2+kx, y-kw/2+ky, n] * K[kx,ky,x%15,y%15]
 * with zero outside the bitmap (taps clipped as mex:55-78 clips them).
 * bitmap [width, height, n] single, kernels [kw, kw, 15, 15] single, out like
 * bitmap.  Host pointers, synchronous.  No kernel-width limit (the reference
 * stops at 341, mex:41). */
mlf_status mlf_lfm_convolution_vec(const float* bitmap, int64_t width, int64_t height, int64_t n,
                                   const float* kernels, int64_t kw, int64_t nphase,
                                   float* out, int32_t device);
/* Same on device pointers, asynchronous on ctx's stream. */
mlf_status mlf_lfm_convolution_vec_device(mlf_context ctx, const float* d_bitmap, int64_t width,
                                          int64_t height, int64_t n, const float* d_kernels,
                                          int64_t kw, int64_t nphase, float* d_out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mlf_lfm_convolution_vec_device with no allocation of
 * any kind.  Its GEMM operand lives in a 64 MiB device array of the library
 * that every GPU it runs on holds until the process ends; odd kernels are
 * limited to kw <= 257 by that array (MLF_ERR_INVALID_ARGUMENT beyond), even
 * kernels need no scratch.  Results equal mlf_lfm_convolution_vec_device's
 * bit for bit.  Calls on one context run in its stream's order; calls that
 * overlap on one GPU from different contexts race on the shared array.
 * ctx must have been created beforehand (creation resolves the array's
 * device address). */
mlf_status mlf_lfm_convolution_vec_noalloc(mlf_context ctx, const float* d_bitmap, int64_t width,
                                           int64_t height, int64_t n, const float* d_kernels,
                                           int64_t kw, int64_t nphase, float* d_out);

/* Replaces img_rectify_mod (image_load_module_aug.m:124-164): frames
 * [H, W] single (column-major, n = frames of them) resampled at a grid of
 * step dx/M px through (x_center, y_center) (1-based pixels; x along the
 * second, y along the first axis), started at a whole lenslet and cut to
 * whole lenslets, by griddedInterpolant's 'linear' interpolation with linear
 * extrapolation past the image.  FP64 arithmetic; out single
 * [out_h, out_w, frames].  mlf_rectify_size gives out_h, out_w (host only).
 * H, W >= 2; the centre inside the image; out_w <= 65535.  Synchronous. */
mlf_status mlf_rectify_size(int64_t H, int64_t W, double x_center, double y_center, int32_t M,
                            double dx, int64_t* out_h, int64_t* out_w);
mlf_status mlf_rectify(const float* in, int64_t H, int64_t W, int64_t frames, double x_center,
                       double y_center, int32_t M, double dx, float* out, int32_t device);
mlf_status mlf_rectify_device(mlf_context ctx, const float* d_in, int64_t H, int64_t W, int64_t frames,
                              double x_center, double y_center, int32_t M, double dx, float* d_out);

/* Replaces the 'carpet' branch of determine_patch_info_aug.m:103-143: an
 * image of global_h x global_w px in n_h x n_w patches of floor(size / n)
 * rounded up to a multiple of M (the last row and column end at the image
 * edge); each patch's height and width are cut down to a multiple of M and a
 * patch below min_size (50 in the reference) in either is discarded.
 * corners: int64 [n_h * n_w][4] = top, left, bottom, right (1-based,
 * inclusive) in MATLAB's ind2sub order (n_h index fastest); *count = patches
 * kept.  Host only. */
mlf_status mlf_carpet_patches(int64_t global_h, int64_t global_w, int32_t n_h, int32_t n_w,
                              int32_t M, int32_t min_size, int64_t* corners, int32_t* count);
/* Replaces determine_patch_info_aug.m:12-281 for a raw frame of raw_h x raw_w
 * px (the reference reads the first TIFF of the recording for its size; the
 * pixel values only feed a figure).  The whole frame's resample grid steps
 * dx/M through (x_center, y_center) = (SI.frames_x_offset, SI.frames_y_offset)
 * from pixel 1 to the frame size, starts at a whole lenslet and is cut to
 * whole lenslets (:51-84); crop_border = SI.frames_crop_border_microlenses
 * [left, right, up, down] microlenses is cut off (:87-103); the patches tile
 * the result: mode MLF_PATCH_CARPET with n_patches [2] (SI.n_patches, as
 * mlf_carpet_patches), or MLF_PATCH_ROI with roi_list [n_roi][4]
 * (SI.ROI_list rows: top, left, bottom, right) and roi_patch_size
 * (SI.ROI_patch_size, cut down to a multiple of M; :144-218).  For every
 * patch, raw_cut is the raw-frame region to read (rows raw_cut[0]..[1],
 * columns [2]..[3], 1-based inclusive) and (raw_cut_xCenter,
 * raw_cut_yCenter) the microlens centre in it that mlf_rectify takes to give
 * exactly the patch's size (:221-275, image_load_module_aug.m:88-104).
 * out NULL: *count only.  Host only. */
#define MLF_PATCH_CARPET 0
#define MLF_PATCH_ROI 1
typedef struct {
  int64_t location[4];   /* top, left, bottom, right in the cropped rectified image (1-based) */
  int64_t size[2];       /* height, width (multiples of M) */
  int64_t raw_cut[4];    /* first row, last row, first column, last column of the raw frame */
  double raw_cut_xCenter, raw_cut_yCenter;
} mlf_patch_info;
mlf_status mlf_determine_patch_info(int64_t raw_h, int64_t raw_w, double x_center, double y_center,
                                   double dx, int32_t M, const int32_t* crop_border, int32_t mode,
                                   const int32_t* n_patches, const int64_t* roi_list, int32_t n_roi,
                                   int32_t roi_patch_size, mlf_patch_info* out, int32_t capacity,
                                   int32_t* count);
/* sensor_movie(top : top+h-1, left : left+w-1, :) of a single movie
 * [H, W, frames] on the device (column-major, 1-based corner) into d_out
 * [h, w, frames]: the patch movie of mesolf.m's per-patch loop, the input of
 * mlf_summary_image_device.  Asynchronous on ctx's stream; w <= 65535. */
mlf_status mlf_crop_movie_device(mlf_context ctx, const float* d_movie, int64_t H, int64_t W,
                                 int64_t frames, int64_t top, int64_t left, int64_t h, int64_t w,
                                 float* d_out);

/* Replaces summary_image_generation_module.m:62-96 and :130 for the default
 * detrend mode 'pixelwise' (mesolf_set_params.m:114): the rank-1 background
 * of rank_1_factorization.m (bg_iter rounds, divisor max(movie)) and the
 * summary image: x = movie / max(movie); d = x - movmean(x, window_1);
 * y = movmean(d, window_2) .^ poly_index over frames; compute_std_image.m
 * with zero factors, sqrt(sum(y.^2)/T - mean(y).^2); divided by its max.
 * movmean is MATLAB's (window [j-a, j+b], a = b = (k-1)/2 for odd k, a = k/2,
 * b = k/2-1 for even k, 'shrink' endpoints).  FP64 arithmetic on the single
 * samples; the detrended movie is held in FP32.
 *   movie:       single [height*width, frames] column-major (pixel fastest), host
 *   summary:     double [height, width]
 *   bg_spatial:  double [height, width] or NULL;  bg_temporal: double [frames] or NULL
 * Defaults of the reference: window_1 = 100, window_2 = 20, poly_index = 3,
 * bg_iter = 5 (mesolf_set_params.m:120-132).  Synchronous. */
mlf_status mlf_summary_image(const float* movie, int64_t height, int64_t width, int64_t frames,
                             int32_t window_1, int32_t window_2, int32_t poly_index,
                             int32_t bg_iter, double* summary, double* bg_spatial,
                             double* bg_temporal, int32_t device);
/* Same on device pointers (bg_spatial / bg_temporal may be NULL), on ctx's
 * stream; returns when done (its scratch, about one FP32 movie, is released). */
mlf_status mlf_summary_image_device(mlf_context ctx, const float* d_movie, int64_t height,
                                    int64_t width, int64_t frames, int32_t window_1,
                                    int32_t window_2, int32_t poly_index, int32_t bg_iter,
                                    double* d_summary, double* d_bg_spatial, double* d_bg_temporal);

/* Same for detrend mode 'global' (summary_image_generation_module.m:43-60,
 * :130): the movie divided by its max, its rank-1 background (bg_iter
 * rounds, rank_1_factorization.m with divisor 1), and compute_std_image.m as
 * written with that background, sqrt((A - s.^2)/T - B.^2 + 2 B.*s C - s.^2 C^2),
 * A = sum(x.^2, 2), B = mean(x, 2), C = mean(bg_temporal); divided by its max.
 * A negative argument gives NaN (MATLAB would return a complex value).  The
 * Savitzky-Golay baseline of :45-48 only feeds a plot and is not computed. */
mlf_status mlf_summary_image_global(const float* movie, int64_t height, int64_t width,
                                    int64_t frames, int32_t bg_iter, double* summary,
                                    double* bg_spatial, double* bg_temporal, int32_t device);
mlf_status mlf_summary_image_global_device(mlf_context ctx, const float* d_movie, int64_t height,
                                          int64_t width, int64_t frames, int32_t bg_iter,
                                          double* d_summary, double* d_bg_spatial,
                                          double* d_bg_temporal);

/* All MEX projections of the seed-library builder in one call
 * (gen_seed_library_local_with_ring_bg_same_size.m:263-318 calls the MEX 4
 * times per plane and accumulates per segment on the host side):
 *   out[:,:,s] = sum over slices i in [seg_ptr[s], seg_ptr[s+1]) of
 *                mex_lfm_convolution_vec(slices[:,:,i], kernels[:,:,:,:,plane[i]])
 * slices [width, height, nslices] single; plane [nslices] 0-based plane of
 * each slice; seg_ptr [nseg+1] non-decreasing with seg_ptr[0] = 0 and
 * seg_ptr[nseg] = nslices; kernels [kw, kw, 15, 15, Z] single; out
 * [width, height, nseg] single (a segment without slices gets zeros).  Each
 * slice's projection is formed on its own and then added to its segment in
 * ascending plane order (slices of one plane in slice order), the order in
 * which the reference's loop over planes adds them (:263-318).  Odd kw runs
 * on the MFMA form of the operator, one GEMM-operand build per plane.  Host
 * pointers, synchronous. */
mlf_status mlf_lfm_convolution_segments(const float* slices, int64_t width, int64_t height,
                                        int64_t nslices, const int32_t* plane,
                                        const int64_t* seg_ptr, int64_t nseg,
                                        const float* kernels, int64_t kw, int64_t nphase,
                                        int64_t Z, float* out, int32_t device);
/* Same with slices, kernels and out in device memory (plane and seg_ptr stay
 * host arrays), on ctx's stream; returns when done. */
mlf_status mlf_lfm_convolution_segments_device(mlf_context ctx, const float* d_slices, int64_t width,
                                               int64_t height, int64_t nslices, const int32_t* plane,
                                               const int64_t* seg_ptr, int64_t nseg,
                                               const float* d_kernels, int64_t kw, int64_t nphase,
                                               int64_t Z, float* d_out);

#ifdef __cplusplus
}
#endif

#endif /* MESOLF_H */
