/*
 * gaussmle.h -- C ABI of libgaussmle_hip.so, the gfx950 (MI355X) drop-in for
 * GaussMLE.jl's per-ROI maximum-likelihood Gaussian fit.
 *
 * The reference launches one KernelAbstractions kernel per batch,
 *   unified_gaussian_mle_kernel!(backend)(results, uncertainties, log_likelihoods,
 *       data, psf_pixels, Val(scmos), variance_map, x_corners, y_corners,
 *       constraints, iterations; ndrange = n)        (src/unified_kernel.jl:285-297)
 * from six call sites in src/interface.jl (:211-216, :249-256, :328-333,
 * :358-364, :415-420, :445-451).  gmle_fit() replaces one such launch together
 * with the GPU branch's batch loop (src/interface.jl:220-262, :336-369,
 * :423-456): same arrays, same layout, same outputs.
 *
 * Conventions (all follow the reference's Julia arrays, passed without copies):
 *   data         Float32 (box, box, n_rois), column-major; ROI k starts at
 *                box*box*k; pixel (row i, column j) at (i-1) + box*(j-1).
 *   params, crlb Float32 (P, n_rois); ROI k's P values are contiguous.
 *   llr          Float32 (n_rois), log-likelihood ratio at the estimate.
 *   var_map      Float32 (ny, nx), column-major, indexed [row = y, col = x]
 *                with row = y_corner + i - 1, col = x_corner + j - 1.
 *   x/y_corners  Int32, 1-based camera coordinates of each ROI's first pixel.
 *   psf          8 floats in pixel units (after to_pixel_units,
 *                src/psf_models.jl:311-331):
 *                  GMLE_XYNB  : psf[0] = sigma
 *                  GMLE_ASTIG : sigma_x0, sigma_y0, Ax, Ay, Bx, By, gamma, d
 *                  others     : ignored (may be NULL)
 *   lower, upper, max_step   P floats each (ParameterConstraints{P},
 *                src/constraints.jl:6-54).
 * Parameter order: XYNB [x, y, N, bg]; XYNBS [x, y, N, bg, sigma];
 * XYNBSXSY [x, y, N, bg, sigma_x, sigma_y]; ASTIG [x, y, z, N, bg].
 *
 * Errors: every entry point returns GMLE_OK (0) or a negative code and never
 * throws.  gmle_last_error() returns the calling thread's last message.
 * Numerical failure is in-band, exactly as in the reference: a singular Fisher
 * matrix gives CRLB = +Inf with the estimate still written
 * (src/unified_kernel.jl:425-430).
 */
#ifndef GAUSSMLE_H
#define GAUSSMLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GMLE_OK 0
#define GMLE_ERR_ARG (-1)         /* invalid argument (the reference's ArgumentError) */
#define GMLE_ERR_BOUNDS (-2)      /* sCMOS ROI outside the variance map */
#define GMLE_ERR_DATA (-3)        /* NaN or Inf in the input data */
#define GMLE_ERR_HIP (-4)         /* HIP runtime error */
#define GMLE_ERR_NOMEM (-5)       /* device allocation failed */
#define GMLE_ERR_NODEVICE (-6)    /* no usable GPU */
#define GMLE_ERR_UNSUPPORTED (-7) /* box size / lanes-per-ROI not built */

enum gmle_model { GMLE_XYNB = 0, GMLE_XYNBS = 1, GMLE_XYNBSXSY = 2, GMLE_ASTIG = 3 };
enum gmle_camera { GMLE_IDEAL = 0, GMLE_SCMOS = 1 };
enum gmle_precision {
  GMLE_FP32 = 0,  /* everything Float32, as the reference */
  GMLE_MIXED = 1  /* Float32 Newton iterations; Float64 LLR, Fisher, inversion, CRLB */
};

/* Library version string. */
const char* gmle_version(void);

/* Number of usable GPUs; 0 when none.  Replaces the CUDA.functional() probe of
 * src/devices.jl:11-35 and src/interface.jl:123-137. */
int gmle_device_count(void);

/* Copies the calling thread's last error message into buf (NUL-terminated,
 * truncated to len); returns the full message length. */
size_t gmle_last_error(char* buf, size_t len);

/* Number of fitted parameters of a model (length(psf_model), src/psf_models.jl:160-161),
 * or GMLE_ERR_ARG. */
int gmle_n_params(int model);

/* default_constraints(psf_model, box_size) (src/constraints.jl:16-47): fills P
 * floats of each array. */
int gmle_default_constraints(int model, int box, float* lower, float* upper, float* max_step);

/* Largest box size supported by the kernels (smallest is 3). */
int gmle_max_box(void);

/* 1 when a kernel with this many lanes per ROI exists for the box size:
 * 2, 4, 8 and 16 for every box, 1 up to 12 px. */
int gmle_lanes_supported(int box, int lanes);

/* Lanes per ROI chosen when a fit call passes lanes = 0: the library's tuned
 * choice for this shape and batch size (more lanes for smaller batches, so
 * that the launch still fills the chip). */
int gmle_default_lanes(int model, int camera, int box, int precision, int64_t n_rois);

/* Replaces one reference kernel launch plus the GPU batch loop
 * (src/unified_kernel.jl:285-297; src/interface.jl:220-262).  All pointers are
 * host memory; the call is synchronous and keeps no pointer after it returns.
 * Validates like validate_fit_input (src/interface.jl:10-40): empty input,
 * box < 3 and NaN/Inf data are errors; sCMOS ROIs must lie inside var_map
 * (the reference reads it unchecked).  var_map / corners may be NULL for
 * GMLE_IDEAL.  crlb64 / llr64 (optional, GMLE_MIXED only) receive the Float64
 * CRLB (P, n_rois) and LLR (n_rois).  lanes = 0 selects automatically. */
int gmle_fit(int model, const float* psf, int box, int64_t n_rois, const float* data, int camera,
             const float* var_map, int ny, int nx, const int32_t* x_corners, const int32_t* y_corners,
             const float* lower, const float* upper, const float* max_step, int iterations, int precision,
             int lanes, float* params, float* crlb, float* llr, double* crlb64, double* llr64);

/* Same computation with every array already in device memory (HBM), enqueued
 * on `stream` (a hipStream_t; NULL = the null stream) and returning without
 * waiting.  psf / lower / upper / max_step are host arrays.  Inputs are not
 * scanned for NaN; an sCMOS ROI whose corner lies outside var_map yields NaN
 * outputs for that ROI instead of an out-of-bounds read. */
int gmle_fit_device(int model, const float* psf, int box, int64_t n_rois, const float* d_data, int camera,
                    const float* d_var_map, int ny, int nx, const int32_t* d_x_corners,
                    const int32_t* d_y_corners, const float* lower, const float* upper, const float* max_step,
                    int iterations, int precision, int lanes, float* d_params, float* d_crlb, float* d_llr,
                    double* d_crlb64, double* d_llr64, void* stream);

/* ---------------------------------------------------------------------------
 * Mixed-shape batches: ROIs of different box sizes and PSF models in one call.
 *
 * The reference fits one box size and one model per call (the kernel reads a
 * single box size, src/unified_kernel.jl:301, and the model is a type
 * parameter of the launch, :285-297), so a mixed set of ROIs takes one `fit`
 * call (src/interface.jl:177-291) per (model, box size).  These entry points
 * take the whole set at once:
 *   models, boxes  Int8 (n_rois): ROI k's model (enum gmle_model) and box size.
 *   data           ROI k's box*box Float32 pixels, column-major, start at
 *                  data + offsets[k] (in floats).  The host entry point packs
 *                  ROIs back to back in ROI order and needs no offsets.
 *   psf            4 x 8 floats: row m is the psf argument of gmle_fit for
 *                  model m (rows of models absent from the batch are ignored).
 *   lower, upper, max_step   4 x 6 floats: row m holds model m's P values.
 *   params, crlb   Float32 (6, n_rois): slot k holds ROI k's P values, then NaN.
 *   llr            Float32 (n_rois).  crlb64 (6, n_rois) / llr64 (n_rois) as in
 *                  gmle_fit (GMLE_MIXED only, optional).
 *   lanes          GMLE_NUM_SHAPES ints indexed by gmle_shape_index(model, box):
 *                  lanes per ROI of each shape (0 = automatic), or NULL.
 *   streams        1-4: the per-shape launches are spread over this many HIP
 *                  streams so that shapes overlap on the GPU (1 = in order).
 *                  0: one launch for every shape, each with its lanes from
 *                  the table where that kernel fits 256 VGPRs (else 4 lanes);
 *                  shapes whose 4-lane kernel needs more than 256 VGPRs
 *                  (astigmatic sCMOS ROIs of 14 px and up) keep their own
 *                  launch.  GMLE_FP32 only.
 * Every ROI's result is bit-identical to a gmle_fit call on that ROI alone with
 * the same model, box, lanes and precision.
 * ------------------------------------------------------------------------- */
#define GMLE_NUM_SHAPES 56 /* 4 models x box sizes 3..16 */

/* Index of (model, box) in the lanes / shape_counts tables, or -1. */
int gmle_shape_index(int model, int box);

/* Lanes per ROI a shape runs with in the one-launch mode (streams = 0) when
 * `lanes` are requested (0 = automatic), or 0 when the shape gets a launch of
 * its own there. */
int gmle_one_launch_lanes(int model, int camera, int box, int lanes);

/* Device workspace gmle_fit_batch_device needs for n_rois ROIs. */
size_t gmle_batch_workspace_bytes(int64_t n_rois);

/* Host arrays, synchronous.  data holds sum(boxes[k]^2) floats.  Validation
 * as gmle_fit, per ROI (box sizes 3..16, known models, finite data, sCMOS
 * ROIs inside var_map). */
int gmle_fit_batch(int64_t n_rois, const int8_t* models, const int8_t* boxes, const float* data,
                   const float* psf, const float* lower, const float* upper, const float* max_step,
                   int camera, const float* var_map, int ny, int nx, const int32_t* x_corners,
                   const int32_t* y_corners, int iterations, int precision, const int32_t* lanes, int streams,
                   float* params, float* crlb, float* llr, double* crlb64, double* llr64);

/* Device arrays (models, boxes, offsets, data, corners, outputs, workspace),
 * enqueued on `stream`.  shape_counts (host, GMLE_NUM_SHAPES entries, or NULL)
 * is the number of ROIs of each shape; with it the call returns without
 * waiting, and ROIs beyond their shape's count get NaN outputs.  With NULL
 * the library counts on the device and waits for the counts.  ROIs of an
 * unknown model or box size get NaN outputs. */
int gmle_fit_batch_device(int64_t n_rois, const int8_t* d_models, const int8_t* d_boxes,
                          const int64_t* d_offsets, const float* d_data, const float* psf, const float* lower,
                          const float* upper, const float* max_step, int camera, const float* d_var_map, int ny,
                          int nx, const int32_t* d_x_corners, const int32_t* d_y_corners, int iterations,
                          int precision, const int32_t* lanes, int streams, const int64_t* shape_counts,
                          float* d_params, float* d_crlb, float* d_llr, double* d_crlb64, double* d_llr64,
                          void* d_workspace, size_t workspace_bytes, void* stream);

/* gmle_fit_batch_device with no allocation of any kind: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The grouping
 * workspace is a static device array of GMLE_NOALLOC_SCRATCH_BYTES (16 MiB)
 * that every GPU the library runs on holds until the process ends; its address
 * is looked up on the first call on each device.  Limits: n_rois up to the
 * largest n with gmle_batch_workspace_bytes(n) <= GMLE_NOALLOC_SCRATCH_BYTES
 * (4,194,176 ROIs); streams 0 (one launch) or 1 (shapes in order on `stream`);
 * larger inputs or stream counts return GMLE_ERR_ARG.  Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU race: serialise
 * them (one stream).  With shape_counts NULL the call waits for the device's
 * shape count, as gmle_fit_batch_device; pass shape_counts for a call that
 * never waits.  Results are bit-identical to gmle_fit_batch_device. */
#define GMLE_NOALLOC_SCRATCH_BYTES (16u << 20)
int gmle_fit_batch_device_noalloc(int64_t n_rois, const int8_t* d_models, const int8_t* d_boxes,
                                  const int64_t* d_offsets, const float* d_data, const float* psf,
                                  const float* lower, const float* upper, const float* max_step, int camera,
                                  const float* d_var_map, int ny, int nx, const int32_t* d_x_corners,
                                  const int32_t* d_y_corners, int iterations, int precision,
                                  const int32_t* lanes, int streams, const int64_t* shape_counts,
                                  float* d_params, float* d_crlb, float* d_llr, double* d_crlb64,
                                  double* d_llr64, void* stream);

/* ---------------------------------------------------------------------------
 * Result epilogue on the device: what GaussMLE.jl does per emitter on the host
 * after the fit (src/interface.jl:265-275, src/roi_batch.jl:29-34, 125-302).
 * For each ROI k, out[10*k + q] holds, in Float32:
 *   0, 1  x, y in microns: ((corner + roi coordinate - 1) - 1) * pixel size
 *   2, 3  their CRLB in microns
 *   4     p-value: 1 - Float32(cdf(Chisq(box^2 - P), -2 LLR))
 *   5, 6  fitted widths in microns (XYNBS: sigma in both; XYNBSXSY: sigma_x,
 *         sigma_y; NaN otherwise)
 *   7, 8  their CRLB in microns (NaN where 5, 6 are)
 *   9     chi^2 = -2 LLR
 * Photons, background, z and their CRLB are used as fitted (unscaled).
 * model / box apply to every ROI unless d_models / d_boxes (per ROI) are
 * given; stride is the params / crlb values per ROI (P, or 6 for mixed
 * batches).  Without corners, ROI k sits at x = 1 + k*box, y = 1, as in
 * fit(fitter, Array) (src/interface.jl:203-205).  Enqueued on `stream`.
 * ------------------------------------------------------------------------- */
int gmle_epilogue_device(int64_t n_rois, int model, int box, const int8_t* d_models, const int8_t* d_boxes,
                         int stride, const float* d_params, const float* d_crlb, const float* d_llr,
                         const int32_t* d_x_corners, const int32_t* d_y_corners, double pixel_size_x,
                         double pixel_size_y, float* d_out, void* stream);

/* The epilogue's p-value on the host: 1 - Float32(cdf(Chisq(df), chi2)). */
float gmle_chi2_pvalue(float chi2, int df);

/* ---------------------------------------------------------------------------
 * ROI extraction on the device: the step before the fit, SMLMBoxer.getboxes
 * in GaussMLE.jl's pipeline (examples/full_pipeline_boxer.jl:70-80,
 * examples/full_pipeline.jl:73-81; SMLMBoxer is a separate package, listed
 * unversioned in examples/Project.toml:9).  Frames are Float32 (ny, nx,
 * n_frames), column-major (row index fastest), in photons or electrons.
 *   D = G(sigma_small) * frame - G(sigma_large) * frame, separable, replicated
 *       edges, 1-D weights exp(-k^2 / 2 sigma^2) for |k| <= floor(3 sigma +
 *       0.5) normalised to sum 1 (gmle_boxer_weights), columns then rows;
 *   a detection is a pixel with D > minval that no pixel within Chebyshev
 *       distance ceil(boxsize - overlap) - 1 (at least 1) beats (larger D, or
 *       equal D and earlier in column-major order): kept detections are at
 *       least boxsize - overlap pixels apart;
 *   its box's top-left corner is the detection minus (boxsize - 1) / 2,
 *       clamped into the frame; ROIs hold the raw frame pixels.
 * Detections are ordered by frame, then column-major pixel index.  Outputs
 * (capacity max_boxes): rois (boxsize, boxsize, n) column-major, 1-based
 * x (column) and y (row) corners and frame indices (a ROIBatch), and the
 * optional strength D at each detection.  The count of all detections goes
 * to *d_count (device) / *n_found (host); only the first max_boxes are
 * written.  boxsize in [3, 16], overlap in [0, boxsize), sigmas up to 4 px.
 * ------------------------------------------------------------------------- */
int64_t gmle_boxer_workspace_bytes(int ny, int nx, int n_frames);
int gmle_find_boxes_device(const float* d_frames, int ny, int nx, int n_frames, int boxsize, float overlap,
                           float sigma_small, float sigma_large, float minval, int64_t max_boxes, float* d_rois,
                           int32_t* d_x_corners, int32_t* d_y_corners, int32_t* d_frame_indices,
                           float* d_strength, int64_t* d_count, void* d_workspace, void* stream);
int gmle_find_boxes(const float* frames, int ny, int nx, int n_frames, int boxsize, float overlap,
                    float sigma_small, float sigma_large, float minval, int64_t max_boxes, float* rois,
                    int32_t* x_corners, int32_t* y_corners, int32_t* frame_indices, float* strength,
                    int64_t* n_found);
/* The filter's half-kernel w[0..R] for one sigma; returns R (-1 if sigma is
 * not positive, R > 12, or capacity < R + 1). */
int gmle_boxer_weights(float sigma, float* w, int capacity);

/* ---------------------------------------------------------------------------
 * Simulator on the device: noisy ROI stacks as GaussMLE.jl's
 * generate_roi_batch draws them (src/simulator.jl:35-345), for tests and
 * benchmarks.  Per pixel, Float32: expected = bg + photons * Phi_x * Phi_y;
 * ideal camera: Poisson(expected); sCMOS: ADU = (Poisson(expected * qe) +
 * N(0,1) * readnoise) / gain + offset, with qe / readnoise / gain / offset
 * from (ny, nx) maps when given (indexed as the fitting kernel indexes the
 * variance map: [y_corner + i - 1, x_corner + j - 1]) or scalars.  Random
 * numbers: Philox4x32-10 keyed by `seed`, counted by (ROI, pixel, purpose);
 * Poisson by multiplication (mean < 10) or PTRS (Hoermann 1993); normal by
 * Box-Muller.  d_params holds `stride` values per ROI in the model's order;
 * d_out is (box, box, n) in the fitting layout.  Enqueued on `stream`.
 * ------------------------------------------------------------------------- */
int gmle_simulate_device(int model, const float* psf, int box, int64_t n_rois, const float* d_params, int stride,
                         int camera, const int32_t* d_x_corners, const int32_t* d_y_corners, const float* d_qe,
                         const float* d_readnoise, const float* d_gain, const float* d_offset, int ny, int nx,
                         float qe, float readnoise, float gain, float offset, uint64_t seed, float* d_out,
                         void* stream);
/* The generator itself, on the host: out = Philox4x32-10(counter[4], key[2]). */
void gmle_philox4x32_10(const uint32_t* counter, const uint32_t* key, uint32_t* out);

/* ---------------------------------------------------------------------------
 * Serving: a resident fitting service with continuous batching.
 *
 * The reference is a library call per batch: upload, launch, wait, download
 * (src/interface.jl:220-262, GPU branch) with the sCMOS variance map copied
 * again for every batch (:232-233) and ADU -> electrons done on the host
 * (src/camera_models.jl:110-113, src/interface.jl:391-392).  A service keeps
 * the sensor's offset / gain / variance maps and per-slot device arenas
 * resident, and takes fit requests (one camera frame's ROIs in raw uint16
 * counts) as they arrive:
 *   - a dispatcher thread packs every pending request that fits into a free
 *     slot's arena (admission by the HBM budget fixed at creation) and
 *     launches the slot's HIP graph: ingest (brings the requests' raw counts
 *     to the GPU, converts ADU to electrons with the resident maps),
 *     grouping by shape, the fit (Float32: one multi-shape kernel for every
 *     served shape; mixed precision: one kernel per shape and its Float64
 *     final pass), result scatter (writes straight into each request's
 *     pinned output buffers);
 *   - up to `slots` batches are in flight at once, each on its own stream, so
 *     requests arriving while others run are admitted as slots free up and
 *     their work overlaps on the GPU;
 *   - a request's results are bit-identical whichever requests share its
 *     batch (each ROI is fitted alone, by its shape's kernel).
 * Request buffers must come from gmle_host_alloc (pinned, device-visible) and
 * stay untouched until the request completes.
 * Raw counts reach the GPU by DMA for every run of a batch's requests whose
 * counts lie back to back in host memory and add up to at least 1 MiB (one
 * copy per run, at the copy engine's rate), and are read in place by the
 * ingest kernel otherwise.  A caller that carves consecutive frames' counts
 * out of one gmle_host_alloc block (a frame ring) gets one copy per batch.
 * GMLE_SERVICE_INGEST=direct|dma|staged|auto overrides the choice (dma: every
 * run by DMA; staged: packed on the host, then one DMA; direct: in place).
 * ------------------------------------------------------------------------- */
typedef struct gmle_service gmle_service;

typedef struct gmle_request {
  int64_t n_rois;
  const int8_t* models;       /* (n_rois) enum gmle_model */
  const int8_t* boxes;        /* (n_rois) box sizes */
  const int32_t* x_corners;   /* (n_rois) 1-based sensor column of each ROI's first pixel */
  const int32_t* y_corners;   /* (n_rois) 1-based sensor row */
  const uint16_t* adu;        /* raw counts, ROIs back to back, each box*box column-major */
  float* params;              /* out (6, n_rois), as gmle_fit_batch */
  float* crlb;                /* out (6, n_rois) */
  float* llr;                 /* out (n_rois) */
} gmle_request;

/* Pinned, device-visible host memory for request buffers (NULL on failure). */
void* gmle_host_alloc(size_t bytes);
int gmle_host_free(void* p);

/* A service on `device`.  hbm_budget: device bytes it may hold (maps plus
 * `slots` arenas of equal size); psf / lower / upper / max_step as
 * gmle_fit_batch; lanes (GMLE_NUM_SHAPES): lanes per ROI of each served shape,
 * 0 for shapes the service does not serve (with GMLE_FP32 the served shapes
 * run in one multi-shape kernel, as gmle_fit_batch with streams = 0). */
int gmle_service_create(int device, size_t hbm_budget, int slots, int precision, int iterations,
                        const float* psf, const float* lower, const float* upper, const float* max_step,
                        const int32_t* lanes, gmle_service** out);

/* Upload the sensor's maps, Float32 (ny, nx) column-major [row = y, col = x]:
 * electrons = (adu - offset) * gain; var = read-noise variance (electrons^2). */
int gmle_service_calibrate(gmle_service* s, const float* offset, const float* gain, const float* var, int ny,
                           int nx);

/* Queue a request; returns a ticket (>= 0) or a negative error code.  Checks:
 * pinned buffers, served shapes, ROIs inside the sensor, size within one
 * slot's capacity. */
int64_t gmle_service_submit(gmle_service* s, const gmle_request* r);

/* Wait for a ticket: 0 done (outputs written), 1 timed out, < 0 error. */
int gmle_service_wait(gmle_service* s, int64_t ticket, double timeout_s);

/* Benchmark driver: submits reqs[i] at arrival_s[i] seconds after the call
 * starts (ascending), waits for all, and returns each request's latency from
 * its scheduled arrival to completion, and the total elapsed time. */
int gmle_service_replay(gmle_service* s, int64_t n, const gmle_request* reqs, const double* arrival_s,
                        double* latency_s, double* elapsed_s);

/* Capacity of one slot (ROIs, pixels), number of slots, batches launched. */
int gmle_service_info(gmle_service* s, int64_t* slot_rois, int64_t* slot_pixels, int* slots, int64_t* batches);

/* Per-batch averages since creation, up to n of: batches, ROIs per batch,
 * requests per batch, host time to form and launch a batch (us), GPU time of
 * a batch (us), time from launch to the completion being seen (us), time from
 * launch to the batch's start on the GPU being seen (us), largest GPU time of
 * a batch (us), largest launch-to-completion time (us), batches whose raw
 * counts were staged (GMLE_SERVICE_INGEST=staged), the fraction of raw-count
 * bytes that reached the GPU by DMA, and host-to-device copies per batch.  A
 * batch stops growing at 65,536 ROIs (a larger request forms a batch of its
 * own). */
int gmle_service_stats(gmle_service* s, double* out, int n);

/* Waits for in-flight work, then frees everything. */
int gmle_service_destroy(gmle_service* s);

#ifdef __cplusplus
}
#endif

#endif /* GAUSSMLE_H */
