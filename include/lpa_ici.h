/*
 * lpa_ici.h — C ABI of the MI355X LPA-ICI 2D denoiser.
 *
 * Plain C: pointers, sizes and status codes only.  Device pointers are HIP
 * device memory owned by the caller; `stream` is a hipStream_t passed as
 * void* (NULL = the default stream).  Calls are stream-ordered and
 * asynchronous unless stated otherwise.  A kernel set (lpa_kset_t) is
 * immutable after creation and may be shared between threads; a plan
 * (lpa_plan_t) belongs to one stream at a time.
 *
 * Reference boundary these entry points replace: the L3 orchestrator
 *   void anisotropic_lpa_ici_gpu_fused(float* host_img_noisy,
 *        float* host_img_denoised, int width, int height, int channels,
 *        float* host_global_weights_buffer, int num_dirs, int num_scales,
 *        float* host_kernel_weights, int host_kernel_weights_size,
 *        int* host_offsets, int* host_rows, int* host_cols,
 *        float* host_kernel_variances, float ici_gamma)
 * declared at src/v4_1_lpa_ici_2D_sparse_shared.cu:297-307 and defined at
 * :1122-1244, with its CPU twin anisotropic_lpa_ici_cpu_naive
 * (src/v0_0_lpa_ici_2D_naive.cpp:230-239, 817-929).  The reference exits the
 * process on any error (src/v4_1_lpa_ici_2D_sparse_shared.cu:24-40, 1185,
 * 1193); here every failure is a status code plus lpa_last_error().
 */
#ifndef LPA_ICI_H
#define LPA_ICI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum lpa_status {
    LPA_OK = 0,
    LPA_EINVAL = 1,  /* bad argument or malformed kernel file */
    LPA_ETOOBIG = 2, /* a size beyond what the kernels support (radius > 31, ...) */
    LPA_ENOMEM = 3,  /* host or device allocation failed */
    LPA_EHIP = 4,    /* a HIP runtime call failed */
    LPA_EIO = 5      /* a file could not be read or written */
} lpa_status;

/* Thread-local text of the last failure on this thread ("" if none). */
const char* lpa_last_error(void);

/* ---------------------------------------------------------------------- */
/* Kernel sets                                                              */
/* ---------------------------------------------------------------------- */

typedef struct lpa_kset* lpa_kset_t;

/* Dense cropped kernels in, device-resident run-length layout out.
 * dense/offsets/rows/cols are host arrays laid out as the reference's
 * loader leaves them: kernel k = d*S + s starts at dense[offsets[k]] and is
 * rows[k] x cols[k], row-major.  Runs are taken per kernel row from the first
 * to the last nonzero (interior zeros kept), with offsets from the centre
 * (rows/2, cols/2); there is no cap on weights or runs.
 * Replaces build_sparse_kernels + the constant uploads,
 * src/v4_1_lpa_ici_2D_sparse_shared.cu:826-876 and :1175-1213.
 * Errors: LPA_EINVAL (bad sizes), LPA_ETOOBIG (a kernel radius above 31,
 * :1189-1195), LPA_ENOMEM, LPA_EHIP. */
lpa_status lpa_kset_create(int num_dirs, int num_scales, const float* dense, const int* offsets,
                           const int* rows, const int* cols, lpa_kset_t* out);
void lpa_kset_destroy(lpa_kset_t ks);

typedef struct lpa_kset_info {
    int num_dirs, num_scales;
    int num_runs;      /* row runs over all kernels */
    int num_weights;   /* run weights over all kernels (interior zeros included) */
    int num_nonzero;   /* nonzero taps over all kernels */
    int max_radius;    /* largest |row| or |column| offset of a nonzero tap */
    int ring_form;     /* 1: every kernel has uniform weights and each scale's support
                          contains the previous one's, so the fast walk builds each
                          scale's estimate from the previous sum plus the new ring */
    int num_ring_taps; /* taps over all rings (ring form only) */
    int ring_frac_bits; /* ring form: the default walk sums tile values in fixed
                           point, q(v) = rint(v * 2^ring_frac_bits), exactly in
                           32-bit integers; yhat = (float)sum * (w * 2^-F) */
    float ring_range;   /* ring form: a tile window whose values all lie strictly
                           within +-ring_range is walked in fixed point; any
                           other window (or a non-finite value) in float */
} lpa_kset_info;
lpa_status lpa_kset_get_info(lpa_kset_t ks, lpa_kset_info* info);

/* Host copies of the run-length layout (any pointer may be NULL):
 * run_row/run_col/run_len/run_woff [num_runs], kernel_first_run/num_runs
 * [D*S], weights [num_weights], nonzero taps per kernel nnz [D*S] and
 * sum of squared weights per kernel (float, dense row-major order) [D*S]. */
lpa_status lpa_kset_get_runs(lpa_kset_t ks, int* run_row, int* run_col, int* run_len, int* run_woff,
                             int* kernel_first_run, int* kernel_num_runs, float* weights, int* nnz,
                             float* sumsq);

/* ---------------------------------------------------------------------- */
/* Images and the forward pass                                              */
/* ---------------------------------------------------------------------- */

typedef struct lpa_image {
    const float* x;      /* device: noisy image, row-major, `pitch` floats per row */
    float* y;            /* device: denoised image, same layout as x */
    uint8_t* scale_map;  /* device, may be NULL: [D][height][width] selected scale
                            index per direction and pixel (dense, pitch = width) */
    int width, height;
    int pitch;           /* floats per row of x and y; 0 means width */
    float sigma;         /* noise standard deviation on the [0,1] scale */
    float gamma;         /* ICI threshold */
    const float* variances; /* host, may be NULL: D*S variances used instead of
                               sigma^2 * sum(w^2) (the reference's
                               host_kernel_variances argument) */
    float* estimates;    /* device, may be NULL: [D][height][width] (dense, pitch =
                            width), each direction's estimate at its selected
                            scale.  lpa_ici_forward writes it (the values it
                            aggregated, bit for bit); lpa_ici_backward reads it,
                            for ring-form sets, instead of recomputing them.
                            D*H*W floats of HBM traded for the recomputation. */
} lpa_image;

enum {
    LPA_FWD_PLAIN = 1u,  /* one launch per image (the plain form) */
    LPA_FWD_EXACT = 2u,  /* separately rounded mul and add in the reference's tap
                            order: the same bits as the CPU v0_0 path (slower) */
    LPA_FWD_GRAPH = 4u,  /* plans only: capture the launch into a hipGraph and replay it */
    LPA_FWD_WEIGHTED = 8u, /* use the weighted run walk even for a ring-form kernel set */
    LPA_FWD_RING_FLOAT = 16u /* ring-form sets: float ring sums for every window
                                (the default takes fixed point where it fits) */
};

/* Denoise n images.  Without LPA_FWD_PLAIN all n go through one persistent
 * launch fed by a tile work queue; with it each image gets its own launch.
 * Per-image tables (variance, threshold, inverse variance) are uploaded
 * synchronously before the launch, so the host arrays may be freed on
 * return.  The images' x/y/scale_map buffers must stay valid until the
 * stream reaches the launch.  Errors: LPA_EINVAL, LPA_ETOOBIG (width or
 * height above 65535, or more than 2^31 pixels), LPA_ENOMEM, LPA_EHIP. */
lpa_status lpa_ici_forward(lpa_kset_t ks, const lpa_image* imgs, int n, unsigned flags, void* stream);

/* lpa_ici_forward with no device allocation at all.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors as lpa_ici_forward; the
 * per-call workspace (work counter, image descriptors, per-image tables and
 * the tile queue) lives in a static 4 MiB array that each device holds until
 * the process ends, instead of a stream-ordered allocation.
 * Limit: 256 + 56 n + 12 K n + 8 T bytes (each term rounded up to 256; K =
 * D*S kernels, T = tiles of the set's tile configuration over the batch,
 * 64x64 px by default) must fit in 4 MiB: for instance one 32768 x 32768
 * image, or about 2,900 small images of the default set; beyond that
 * LPA_ETOOBIG.  The kernel set must already be resident on the current
 * device (its first lpa_ici_forward, lpa_plan_create, lpa_ici_estimates or
 * lpa_ici_backward call makes it so); otherwise LPA_EINVAL.  Every call on
 * one device shares the array, which keeps the last call's contents: calls
 * on one stream are ordered, calls that overlap on one device (other
 * streams, other threads) race.  Like lpa_ici_forward it copies the tables
 * from host memory and waits for that copy. */
lpa_status lpa_ici_forward_noalloc(lpa_kset_t ks, const lpa_image* imgs, int n, unsigned flags, void* stream);

/* ---------------------------------------------------------------------- */
/* Plans: descriptors, tables and work queue uploaded once, replayed often  */
/* ---------------------------------------------------------------------- */

typedef struct lpa_plan* lpa_plan_t;

/* Build the device state for a fixed batch (descriptors, per-image tables,
 * tile queue, work counter).  flags as for lpa_ici_forward.  With
 * LPA_FWD_GRAPH the first lpa_plan_run captures the launch sequence on its
 * stream into a hipGraph and every later run replays it.  The plan keeps
 * no host pointer from imgs. */
lpa_status lpa_plan_create(lpa_kset_t ks, const lpa_image* imgs, int n, unsigned flags, lpa_plan_t* out);
lpa_status lpa_plan_run(lpa_plan_t plan, void* stream);
void lpa_plan_destroy(lpa_plan_t plan);

/* Tile configuration the plan uses (after autotuning, if any). */
typedef struct lpa_tile_config {
    int tile_w, tile_h;   /* output pixels per work-group tile */
    int px_per_lane;      /* adjacent output pixels one lane computes */
    int body;             /* inner loop: 2 = one chain of unrolled taps, left after a run's last, 0 = sliding-window loop */
    int halo;             /* LDS halo radius, >= the kernel set's max radius */
    int blocks_per_cu;    /* persistent work-groups per CU (batched form) at the full halo; 0 without a device */
} lpa_tile_config;

/* Number of compiled tile configurations, and the i-th of them. */
int lpa_num_tile_configs(void);
lpa_status lpa_tile_config_get(int index, lpa_tile_config* cfg);
/* Make `index` the configuration used by later launches of this kernel set
 * (the autotuner's output); -1 restores the built-in choice: configuration
 * 9 for the fixed-point ring walk, 7 for the other walks.  The getter
 * returns the configuration of the set's default walk. */
lpa_status lpa_kset_set_tile_config(lpa_kset_t ks, int index);
int lpa_kset_get_tile_config(lpa_kset_t ks);

/* Per-direction estimates at the scales a forward selected: for each image,
 * est[i] (device, [D][height][width] floats, dense) receives, per direction
 * d and pixel p, the estimate of d at scale scale_map[d][p] (the map is
 * required), computed with the forward's walk for the same flags
 * (LPA_FWD_EXACT, LPA_FWD_WEIGHTED, LPA_FWD_RING_FLOAT) and, for the
 * fixed-point ring walk, on the kernel set's tile configuration (that walk
 * chooses fixed point or float ring sums per tile window), so it is the value
 * a forward with the same flags and configuration aggregated, bit for bit.
 * A forward can also keep these values as it walks (lpa_image.estimates).
 * These are the reference's per
 * direction debugging outputs (the commented-out writes of img_denoised_d and
 * scale_map_d, src/v0_0_lpa_ici_2D_naive.cpp:848, 901, 906-914).  Errors:
 * LPA_EINVAL (missing scale map or est, LPA_FWD_GRAPH), LPA_ETOOBIG, LPA_ENOMEM,
 * LPA_EHIP. */
lpa_status lpa_ici_estimates(lpa_kset_t ks, const lpa_image* imgs, float* const* est, int n, unsigned flags,
                             void* stream);

/* ---------------------------------------------------------------------- */
/* Adjoint (selection frozen at the forward's scale maps)                   */
/* ---------------------------------------------------------------------- */

/* For each image i: given dy[i] (device, same layout as imgs[i].y) and the
 * scale map the forward wrote (imgs[i].scale_map, required), write
 *   dx[i]   (device, same layout as x) = dL/dx, overwritten, and
 *   dkernel (device, num_weights floats in run-weight order, may be NULL)
 *           += dL/d(run weights), summed over the n images.
 * The estimates are recomputed from the scale maps; nothing else of the
 * forward is stored.  The threshold plays no part once the scales are
 * frozen; the variances sigma^2*sum(w^2) are differentiated through sum(w^2).
 * Per-call workspace is allocated and freed inside the call (stream-ordered).
 * Errors: LPA_EINVAL (missing scale map, ...), LPA_ENOMEM, LPA_EHIP. */
lpa_status lpa_ici_backward(lpa_kset_t ks, const lpa_image* imgs, const float* const* dy, float* const* dx,
                            float* dkernel, int n, void* stream);

/* Peak device memory the library's per-call workspaces (lpa_ici_forward,
 * lpa_ici_estimates, lpa_ici_backward: descriptors, tables, tile queues, the
 * adjoint's q map and partial rows) have held at once since the last reset,
 * in bytes: the used-memory high-water mark of the current device's default
 * stream-ordered pool, which those workspaces come from.  reset != 0 clears
 * the mark after reading it.  No counterpart in the reference (which
 * allocates with cudaMalloc per call, src/v4_1_lpa_ici_2D_sparse_shared.cu:
 * 1142-1157); it lets a caller state peak HBM use, inputs and outputs
 * (caller-owned) added.  Errors: LPA_EINVAL, LPA_EHIP. */
lpa_status lpa_workspace_peak(unsigned long long* bytes, int reset);

/* ---------------------------------------------------------------------- */
/* Host substrate of the driver (all host memory, synchronous)              */
/* ---------------------------------------------------------------------- */

/* load_LPA_kernels_standardized, src/v0_0_lpa_ici_2D_naive.cpp:450-535:
 * header "D S", then per (d,s) "rows cols" and rows*cols weights; each
 * kernel is centre-cropped to direction 0's size for its scale.  Files
 * ending in ".gz" are read through zlib.  Two calls: first with NULL arrays
 * to get D, S and the weight count, then with arrays of that size
 * (weights [n_weights], offsets/rows/cols [D*S]). */
lpa_status lpa_load_kernels(const char* path, int* num_dirs, int* num_scales, int* n_weights, float* weights,
                            int* offsets, int* rows, int* cols);

/* Kernel-set generation (no counterpart in the reference's code: its sets
 * come precomputed from a modified LASIP demo_CreateLPAKernels.m that is not
 * in the repository, readme.md:87-92,239).  Uniform wedges: kernel (d, s) is
 * 1/n (six decimals, as the files store it) on the n lattice points (x, y)
 * (y = -row) that are the centre or lie within pi/D of direction
 * phi_d = 2 pi d / D and project onto it below h[s] - 1/2; (2h-1)^2 dense,
 * centred.  D = 16 with h = 1,2,4,8,16,24,32 gives the shipped default set
 * (lpa_kernels_m_1_0_d_16_h12_1_2_4_8_16_24_32_symmetric.txt) weight for
 * weight.  Needs D >= 4 and 1 <= h <= 32.  Same two-call protocol and
 * layout as lpa_load_kernels. */
lpa_status lpa_wedge_kernels(int num_dirs, int num_scales, const int* h, int* n_weights, float* weights,
                             int* offsets, int* rows, int* cols);

/* Directional LPA kernel sets of polynomial order `order` (no counterpart in
 * the reference's code: its sets come precomputed from a modified LASIP
 * demo_CreateLPAKernels.m, readme.md:87-92, 239).  Kernel (d, s) is the
 * zero-order LPA estimator: a degree-min(order, h-1) polynomial P fitted by
 * weighted least squares on the samples u = 0..h[s]-1 along direction
 * phi_d = 2 pi d / D, then k = w(u) w(v) P(u) on the sector wedge of
 * lpa_wedge_kernels ((u, v) the point's coordinates along and across the
 * direction), divided by its sum.  window: LPA_WINDOW_UNIFORM (w = 1),
 * LPA_WINDOW_GAUSSIAN (w(t) = exp(-t^2 / (2 (0.3 h)^2))), or
 * LPA_WINDOW_SQUARE (D = 4 only: the h x h quadrant to the right of the
 * direction, fitted on all its points).  Weights with six decimals, as the
 * files store them.  order 0 with the uniform window is lpa_wedge_kernels.
 * The square and symmetric sets, and the axis directions of the m = 1 and
 * m = 2 sets, come out equal to the shipped files; oblique directions of
 * m >= 1 sets differ from them (LASIP's discretization there is not known).
 * Needs 0 <= order <= 4, 1 <= h <= 32, D >= 4.  Same two-call protocol and
 * layout as lpa_load_kernels. */
enum { LPA_WINDOW_UNIFORM = 0, LPA_WINDOW_GAUSSIAN = 1, LPA_WINDOW_SQUARE = 2 };
lpa_status lpa_make_kernels(int num_dirs, int num_scales, const int* h, int order, int window, int* n_weights,
                            float* weights, int* offsets, int* rows, int* cols);

/* Write a kernel set in the reference's text format (the format
 * lpa_load_kernels reads): "D S", then per kernel "rows cols" and its rows
 * of "%f" weights.  Errors: LPA_EINVAL, LPA_EIO. */
lpa_status lpa_write_kernels(const char* path, int num_dirs, int num_scales, const float* weights,
                             const int* offsets, const int* rows, const int* cols);

/* compute_kernel_variances, src/v0_0_lpa_ici_2D_naive.cpp:675-701:
 * var[k] = (sigma*sigma) * sum(w^2), float accumulation in dense order. */
void lpa_kernel_variances(int num_kernels, const float* weights, const int* offsets, const int* rows,
                          const int* cols, float sigma, float* variances);

/* create_noisy_image, src/v0_0_lpa_ici_2D_naive.cpp:743-760:
 * noisy[i] = clean[i] + N(0, sigma), drawn in order from std::mt19937(seed)
 * through std::normal_distribution<float>. */
void lpa_make_noisy(const float* clean, float* noisy, long long n, float sigma, unsigned seed);

/* calculate_PSNR, src/v0_0_lpa_ici_2D_naive.cpp:762-783 (peak 1.0). */
double lpa_psnr(const float* a, const float* b, long long n);

/* Synthetic test image in [0,1]: piecewise-constant polygons and disks over
 * smooth linear ramps, with oriented sinusoidal stripe patches (periods
 * 3-12 px).  Deterministic in (width, height, seed). */
void lpa_synth_image(float* out, int width, int height, unsigned seed);

/* stbi_loadf(path, &w, &h, &c, 1) semantics for 8-bit PNG: RGB folded to
 * gray as (77R+150G+29B)>>8, then powf(v/255, 2.2).  Call with out=NULL to
 * get the size.  Errors: LPA_EIO, LPA_EINVAL (unsupported PNG). */
lpa_status lpa_load_image(const char* path, int* width, int* height, float* out, long long capacity);

/* Radiance RGBE .hdr writer, one channel replicated to RGB; values below
 * 1e-32 (negatives included) are written as 0. */
lpa_status lpa_write_hdr(const char* path, int width, int height, const float* data);

/* Multichannel input and output (the reference loads with one channel only,
 * src/v0_0_lpa_ici_2D_naive.cpp:254-256).  lpa_load_image_channels:
 * stbi_loadf(path, &w, &h, &c, channels) semantics for channels 1 or 3
 * (interleaved; gray PNGs replicate to RGB, palettes expand, alpha is
 * dropped, each channel goes through powf(v/255, 2.2)).
 * lpa_write_hdr_channels: RGBE with a shared exponent per pixel, as
 * stbi_write_hdr (stb_image_write.h:639-654); a pixel whose largest component
 * is below 1e-32 is written as 0. */
lpa_status lpa_load_image_channels(const char* path, int channels, int* width, int* height, float* out,
                                   long long capacity);
lpa_status lpa_write_hdr_channels(const char* path, int width, int height, int channels, const float* data);

#ifdef __cplusplus
}
#endif

#endif /* LPA_ICI_H */
