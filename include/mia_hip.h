/*
 * mia_hip.h -- C ABI of the MI355X (gfx950) implementation of MIA's B-spline
 * registration inner loop: cost + gradient of `image:cost=ssd` under the
 * `spline` transformation, for one or many (batched) 2D/3D registrations.
 *
 * Plain pointers and sizes only.  Host buffers are owned by the caller; device
 * buffers are owned by the opaque context.  One context = one HIP stream; a
 * context must not be used from two threads at once (the reference's objects
 * have the same restriction, mia/3d/transform/spline.cc:230,549).
 *
 * Every entry point returns a status: MIA_HIP_OK (0) or a negative code; the
 * message of the last failure on the calling thread is mia_hip_last_error().
 * The C++ layer (mia_amd/host) maps the codes back to the reference's
 * exception types (invalid_argument, runtime_error).
 *
 * Layouts (the reference's, SURVEY of mia/3d/datafield.cxx:57-58):
 *   images   : float, x fastest, idx = x + nx*(y + ny*z)
 *   params   : double, interleaved per coefficient [cx,cy(,cz)], coefficients
 *              x fastest (mia/3d/transform/spline.cc:289-358,
 *              mia/2d/transform/spline.cc:292-316); a batch concatenates the
 *              instances' parameter vectors in instance order.
 */
#ifndef MIA_HIP_H
#define MIA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MIA_HIP_OK 0
#define MIA_HIP_ERR_ARG -1      /* bad argument: reference throws invalid_argument */
#define MIA_HIP_ERR_SIZE -2     /* image size mismatch: reference throws runtime_error */
#define MIA_HIP_ERR_STATE -3    /* call order (images not set, no setup, ...) */
#define MIA_HIP_ERR_DEVICE -4   /* HIP runtime failure */
#define MIA_HIP_ERR_LIMIT -5    /* size beyond a documented kernel limit */

/* cost flags (mia_hip_set_cost) */
#define MIA_HIP_GRAD_REFERENCE 0 /* gradient of the warped image (what MIA computes) */
#define MIA_HIP_GRAD_EXACT 1     /* analytic derivative of the cost (what FD measures) */

typedef struct mia_hip_ctx mia_hip_ctx;

/* Thread-local message of the last failed call. */
const char *mia_hip_last_error(void);

/* Coefficient grid size along one axis for `spline:rate=<rate>` with the
 * default cubic kernel; replaces the size computation in the
 * C2D/C3DSplineTransformation constructors (mia/3d/transform/spline.cc:130-143,
 * mia/2d/transform/spline.cc:103-116).  Returns -1 where the reference throws
 * invalid_argument (rate too large for the image). No device use. */
int mia_hip_spline_csize(uint32_t n, float rate, int dim);

/* Create / destroy a context bound to HIP device `device`. */
int mia_hip_ctx_create(int device, mia_hip_ctx **out);
int mia_hip_ctx_destroy(mia_hip_ctx *ctx);

/* Describe the registration problem(s): `n_inst` independent instances of
 * dimension `dim` (2 or 3) with image sizes sizes[3*i..3*i+2] (x,y,z; z = 1 in
 * 2D) and coefficient grid sizes csizes[3*i..] (including the spline's
 * 2-coefficient enlargement, as get_coeff_size() returns them,
 * mia/3d/transform/spline.hh:38-101).  `degree` is the image interpolation
 * kernel, 3 (`imgkernel=bspline:d=3`, default) or 1 (`bspline:d=1`); boundary
 * is mirror (mia/template/transformfactory.cxx:58-59).  Replaces the grid
 * set-up of C?DSplineTransformation::reinit (3D spline.cc:224-275, 2D
 * :170-203).  Allocates all device buffers.  MIA_HIP_ERR_LIMIT for an x size
 * beyond the row limit or a 3D image of 2^29 voxels or more (the kernels
 * address an image with 32-bit byte offsets). */
int mia_hip_setup(mia_hip_ctx *ctx, int dim, int n_inst, const uint32_t *sizes,
                  const uint32_t *csizes, int degree);

/* New coefficient grids csizes[3*i..] for the instances of the last
 * mia_hip_setup, their image sizes, dimension and interpolation kernel
 * unchanged: the same as mia_hip_setup with those sizes, except that the
 * images already set (the prefiltered moving image, the fixed image, the
 * clamp range) are kept and need not be set again.  For the refinement of a
 * spline grid at one level (mia/3d/transform/spline.cc:360-397 followed by
 * the next minimisation).  Penalties are detached as by mia_hip_setup. */
int mia_hip_setup_grid(mia_hip_ctx *ctx, const uint32_t *csizes);

/* Upload instance `inst`'s moving (source) and fixed (reference) images from
 * host memory, prefilter the moving image and record its clamp range.
 * Replaces C?DImageFullCost::do_reinit/set_reference + the interpolator
 * construction that the reference repeats on every warp
 * (mia/3d/fullcost/image.cc:112-128, mia/3d/interpolator.cxx:98-164). */
int mia_hip_set_images(mia_hip_ctx *ctx, int inst, const float *moving, const float *fixed);

/* Same with device pointers (float, HBM-resident, same layout). */
int mia_hip_set_images_device(mia_hip_ctx *ctx, int inst, const void *d_moving, const void *d_fixed);

/* Which reference-gradient stage-1 layout instance `inst` runs with after
 * mia_hip_setup: 1 = per (z knot cell, y knot band) with the slices' W in
 * registers (3D instances with x rows up to 256 voxels and y bands of up to
 * 9 rows, where the tuned or forced layout asks for it:
 * MIA_HIP_TUNE_STAGE1, mia_hip_set_stage1_path), 0 = per (slice, y band)
 * (2D, exact mode, larger rows, or chosen).  The choice depends only on the
 * instance's own sizes, so an instance computes the same in any batch.
 * Counterpart of the reference's translate path
 * (mia/3d/transform/spline.cc:731-791); introspection only. */
int mia_hip_get_stage1_path(mia_hip_ctx *ctx, int inst, int *path);

/* Cost options: ssd `norm` (default false), `autothresh` (0 = off), the
 * full-cost `weight` (mia/3d/fullcost/image.cc:155-165,
 * mia/template/fullcost.cxx:36-45) and the gradient mode
 * (MIA_HIP_GRAD_REFERENCE / MIA_HIP_GRAD_EXACT).  Plugin parameters of
 * mia/template/ssd.hh:217-231. */
int mia_hip_set_cost(mia_hip_ctx *ctx, int norm, float autothresh, double weight, int grad_mode);

/* Number of parameters of instance `inst` and its offset in the concatenated
 * parameter vector (degrees_of_freedom(), spline.cc:414-418). */
int mia_hip_n_params(mia_hip_ctx *ctx, int inst, size_t *n, size_t *offset);

/* Cost and gradient of every instance for host parameter vector `params`
 * (all instances concatenated).  Writes grad (same layout, OVERWRITTEN, i.e.
 * the weighted gradient of this cost alone) and costs[n_inst].  Replaces
 * C?DImageFullCost::do_evaluate + TFullCost::evaluate
 * (mia/3d/fullcost/image.cc:67-78, mia/2d/fullcost/image.cc:81-105,
 * mia/template/fullcost.cxx:36-45).  Synchronous. */
int mia_hip_cost_grad(mia_hip_ctx *ctx, const double *params, double *grad, double *costs);

/* Same with device pointers; enqueued on the context stream, asynchronous. */
int mia_hip_cost_grad_device(mia_hip_ctx *ctx, const void *d_params, void *d_grad, void *d_costs);

/* The context's own device buffers for parameters, gradient (n_params
 * doubles each) and costs (n_inst doubles), for callers that keep the
 * optimizer state in HBM and call mia_hip_cost_grad_device on them.  Valid
 * until the next mia_hip_setup or mia_hip_ctx_destroy. */
int mia_hip_device_buffers(mia_hip_ctx *ctx, void **d_params, void **d_grad, void **d_costs);

/* Warped moving image of instance `inst` for its parameters `params_inst`
 * (transformation operator()(image), mia/core/transformation.hh:151-160,
 * mia/3d/transform.cc:280-323).  Synchronous, host output. */
int mia_hip_warp(mia_hip_ctx *ctx, int inst, const double *params_inst, float *out);

/* Capture one cost+gradient iteration (host params -> device -> kernels ->
 * host grad/costs, all through pinned staging buffers) as a HIP graph.  After
 * this, mia_hip_graph_params() returns the pinned parameter buffer to fill and
 * mia_hip_graph_run() replays the iteration and waits; the results appear in
 * the pinned buffers returned by mia_hip_graph_outputs(). */
int mia_hip_graph_capture(mia_hip_ctx *ctx);
int mia_hip_graph_params(mia_hip_ctx *ctx, double **params);
int mia_hip_graph_outputs(mia_hip_ctx *ctx, double **grad, double **costs);
int mia_hip_graph_run(mia_hip_ctx *ctx);

/* Minimise every instance's cost with L-BFGS on the device, all instances in
 * lockstep: each iteration is the batched evaluation followed by one
 * optimizer step per instance (one thread block each for instances of up to
 * 2048 parameters; larger ones in segments of 512 parameters over three
 * launches, so that their pair history streams through many CUs), queued in
 * chunks of `iters_per_graph` iterations, each chunk ending with a copy of
 * the instances' done flags that the host checks (see
 * mia_hip_set_optimize_launch for how a chunk is queued).  Same algorithm
 * as the host `lbfgs` minimizer: memory `memory` (<= 16), `maxeval`
 * evaluations per instance (NLopt's maxeval, which MIA's `maxiter` maps to,
 * addons/nlopt/nlopt.cc:237), first step moving no parameter by more than
 * `first_step`; the direction is the two-loop recursion's in the compact
 * representation, so trajectories agree with the host to rounding, and an
 * instance's trajectory does not depend on the batch.  Penalties attached
 * with mia_hip_set_penalty are part of every evaluation.  x0: starting
 * parameters of all instances (host); outputs (host, any may be NULL): best
 * parameters, best costs, evaluations per instance, wall time of the
 * iteration loop in seconds (graph capture and instantiation excluded).
 * Replaces the
 * minimizer loop of TNonrigidRegister::run for the level
 * (mia/template/nonrigidregister.cxx:261-269). */
int mia_hip_optimize(mia_hip_ctx *ctx, const double *x0, int maxeval, int memory, double first_step,
                     int iters_per_graph, double *x_out, double *costs_out, int *evals_out, double *seconds);

/* mia_hip_optimize on n_ctx contexts of one device at once: context k
 * starts from x0[k] and writes x_out[k], costs_out[k], evals_out[k] (any of
 * the three arrays, or an entry, may be NULL).  Each context's iterations
 * are queued by a host thread of its own on one of three streams that the
 * library keeps per device, created with the device's first context so that
 * HIP gives them hardware queues of their own: one context's optimizer step,
 * which occupies a block per instance, runs beside another context's
 * evaluation, which spreads over the whole GPU.  Every context's results are
 * bitwise those mia_hip_optimize gives it alone (an instance's trajectory
 * depends on nothing outside it).  Chunks are queued as stream launches
 * whatever mia_hip_set_optimize_launch chose; with more than three contexts,
 * contexts k and k + 3 share a stream.  seconds: wall time of the joint
 * iteration loop.  No counterpart in the reference (its minimizers run on
 * the host, one registration at a time: src/2dmany2one-nonrigid.cc:185). */
int mia_hip_optimize_many(mia_hip_ctx *const *ctx, int n_ctx, const double *const *x0, int maxeval, int memory,
                          double first_step, int iters_per_graph, double *const *x_out, double *const *costs_out,
                          int *const *evals_out, double *seconds);

/* How mia_hip_optimize queues its chunks of iterations.
 * MIA_HIP_LAUNCH_STREAM (the default): the kernels of each chunk are launched
 * on the context's stream, the next chunk queued while the host waits for the
 * previous one's done flags (only chunks that start before the maxeval-th
 * iteration are queued ahead, so no evaluation is spent past maxeval).
 * MIA_HIP_LAUNCH_GRAPH: one chunk captured as a HIP graph, instantiated once
 * and launched until every instance is done.  Both run the same kernels in
 * the same order and give bitwise the same results; on ROCm a graph launch
 * submits its nodes one by one from the host, so it costs more host time per
 * kernel than direct launches.  No counterpart in the reference (its
 * minimizer runs on the host). */
#define MIA_HIP_LAUNCH_STREAM 0
#define MIA_HIP_LAUNCH_GRAPH 1
int mia_hip_set_optimize_launch(mia_hip_ctx *ctx, int mode);

/* Wait for all work on the context stream. */
int mia_hip_synchronize(mia_hip_ctx *ctx);

/* The context's hipStream_t (as void*), for callers that enqueue around it. */
int mia_hip_get_stream(mia_hip_ctx *ctx, void **stream);

/* Per-kernel timing with HIP events on the context stream.  While enabled,
 * each kernel launch is bracketed by events; mia_hip_profile_read returns, per
 * stage id (see mia_hip_stage_name), the summed milliseconds and the number of
 * launches since the last read, then resets the counters. */
int mia_hip_profile_enable(mia_hip_ctx *ctx, int enable);
int mia_hip_profile_read(mia_hip_ctx *ctx, double *ms, int *launches, int max_stages, int *n_stages);
const char *mia_hip_stage_name(int stage);

/* ---- tile variants, stage-1 layouts and their autotuner ---------------- */

/* Kinds of tuned choices, each made per (interpolation degree, size bucket
 * ceil(log2(voxels or pixels))) from a table keyed by the GPU architecture. */
#define MIA_HIP_TUNE_WARP3D 0 /* 3D warp tile variant (mia_hip_warp_variants) */
#define MIA_HIP_TUNE_WARP2D 1 /* 2D warp tile variant (mia_hip_warp2d_variants) */
#define MIA_HIP_TUNE_STAGE1 2 /* 3D reference-gradient stage-1 layout: 0 per (slice, y band) with LDS
                                 accumulators, 1 per (z knot cell, y band) with the band in registers
                                 (only where x size <= 256 and a band has at most 9 rows) */

/* The GPU architecture of `device` ("gfx950"; the name the tuning table is
 * keyed by) into buf.  No counterpart in the reference. */
int mia_hip_device_arch(int device, char *buf, int len);

/* The tile shapes (x, y, z voxels per block) of the 3D warp kernel's
 * variants; tile_dims receives 3 ints per variant (up to max_variants). */
int mia_hip_warp_variants(int *tile_dims, int max_variants, int *n_variants);
/* The tile shapes of the 2D warp kernel's variants: tile_dims receives 3
 * ints per variant (up to max_variants): x, y pixels per block and 1 when the
 * tile's taps go through an LDS box, 0 when each pixel gathers them from the
 * image. */
int mia_hip_warp2d_variants(int *tile_dims, int max_variants, int *n_variants);

/* Force tile variant `variant` (0..n-1) for the instances of the next
 * mia_hip_setup on this context (3D or 2D variants by the setup's dim; a
 * number beyond the dim's variants makes that setup fail), or
 * -1 (default) for the tuned choice per size bucket.  Results do not depend
 * on the variant beyond FP32 rounding of the displacement contraction;
 * every instance always uses the same variant for a given size, so batched
 * and per-instance runs stay bitwise equal. */
int mia_hip_set_warp_variant(mia_hip_ctx *ctx, int variant);
/* The variant instance `inst` runs with after mia_hip_setup. */
int mia_hip_get_warp_variant(mia_hip_ctx *ctx, int inst, int *variant);
/* Force the stage-1 layout of 3D reference gradients for the next
 * mia_hip_setup: 0 per slice, 1 per z knot cell where the instance allows
 * it, -1 (default) the tuned choice per size bucket.  The layouts sum the
 * gradient in different orders (equal to FP rounding); the choice depends
 * on the instance's size only, so batches stay bitwise equal to single
 * instances.  mia_hip_get_stage1_path reports the layout taken. */
int mia_hip_set_stage1_path(mia_hip_ctx *ctx, int path);

/* Time every choice of `kind` (MIA_HIP_TUNE_*) on synthetic instances of 2^b
 * voxels or pixels (cubes / squares), b in [b_min, b_max]: 3D warp at spline
 * rate 6 (b in [12, 26]), 2D warp at rate 8 (b in [10, 24]), stage-1 layout
 * at rate 8 (b in [12, 24]: x sizes up to 256); displacements N(0, 1.5)
 * clipped to 4 voxels; for interpolation degree 3, 1 or both (0).  The
 * fastest choice per (degree, b) replaces the entry of that kind, degree
 * and bucket in the table json_path (may be NULL: nothing written; the
 * file's other entries are kept when it holds a table of this
 * architecture), which is used from then on.  us_out (may be NULL): mean microseconds per launch (stage 1 plus
 * reduce for the layout), n_degrees x n_buckets x n_choices. */
int mia_hip_autotune(int device, int kind, int degree, int b_min, int b_max, const char *json_path,
                     double *us_out);
/* mia_hip_autotune(device, MIA_HIP_TUNE_WARP3D, ...). */
int mia_hip_autotune_warp(int device, int degree, int b_min, int b_max, const char *json_path, double *us_out);

/* (Re)load the tuning table: json_path, or NULL for $MIA_HIP_TUNING, else
 * <library dir>/../tuning/<arch>.json.  A table for another architecture is
 * ignored (every kind's default: 3D and 2D variant 0, z-cell stage 1). */
int mia_hip_load_warp_tuning(int device, const char *json_path);
/* The file the table came from (empty if none) into buf; returns the number
 * of (kind, degree, bucket) entries. */
int mia_hip_warp_tuning_source(char *buf, int len);

/* ---- per-level work of the multiresolution driver ---------------------- */

/* An image kept in device memory (float, x fastest), so that the levels of a
 * registration resample it and set it into a context without going through
 * host memory.  No counterpart in the reference (its images are host data,
 * mia/core/datapool.hh): the device-resident form of the images the full cost
 * reads from the data pool (mia/3d/fullcost/image.cc:80-97,111-127). */
typedef struct mia_hip_image mia_hip_image;
/* Upload `data` (dim 2 or 3, sizes x, y[, z]) to device `device`; synchronous. */
int mia_hip_image_create(int device, int dim, const uint32_t *size, const float *data, mia_hip_image **out);
/* A new device image: `in` resampled to out_size with mia_hip_scale_image's
 * arithmetic (bitwise the same values); synchronous. */
int mia_hip_image_scale(const mia_hip_image *in, const uint32_t *out_size, mia_hip_image **out);
/* Copy the image to host memory (its element count of floats). */
int mia_hip_image_read(const mia_hip_image *img, float *out);
int mia_hip_image_destroy(mia_hip_image *img);
/* mia_hip_set_images_device from two device images of the instance's size
 * (MIA_HIP_ERR_SIZE otherwise) on the context's device. */
int mia_hip_set_images_image(mia_hip_ctx *ctx, int inst, const mia_hip_image *moving, const mia_hip_image *fixed);

/* Resample a float image (dim 2 or 3, sizes x,y[,z]) to out_size with the
 * `scale` filter's semantics, interp=bspline:d=3: per axis that changes,
 * downscaling applies a binomial filter of half width round(1/scale), then
 * the cubic B-spline prefilter with mirror boundaries, then samples at
 * i * (n_in - 1) / (n_out - 1) with cubic weights and mirrored indices;
 * double precision between the axes.  Host buffers, synchronous.  Replaces
 * C3DScale / C2DScale::operator() (mia/3d/filter/scale.cc:38-104,
 * mia/2d/filter/scale.cc) with C1DScalar (mia/core/scaler1d.cc:26-230). */
int mia_hip_scale_image(int device, int dim, const uint32_t *in_size, const float *in, const uint32_t *out_size,
                        float *out);

/* The 3D divcurl spline penalty: value c^T P c and gradient of the banded
 * quadratic form over the coefficient grid csize (x,y,z) for an image of
 * extent range (x,y,z) with divergence / curl weights.  Replaces
 * C3DPPDivcurlMatrix (mia/3d/ppmatrix.cc: reset :178-262, evaluate
 * :300-520).  coeffs: host, 3 floats per coefficient, x fastest.  value:
 * the form; gradient (may be NULL): its derivative, 3 doubles per
 * coefficient, unweighted and with the sign of the derivative (the caller
 * applies -weight, mia/3d/splinetransformpenalty.cc:60-71). */
typedef struct mia_hip_divcurl mia_hip_divcurl;
int mia_hip_divcurl_create(int device, const uint32_t *csize, const float *range, double div_weight,
                           double curl_weight, mia_hip_divcurl **out);
int mia_hip_divcurl_evaluate(mia_hip_divcurl *p, const float *coeffs, double *value, double *gradient);
int mia_hip_divcurl_destroy(mia_hip_divcurl *p);

/* Attach penalty p (a divcurl object created for the coefficient grid of 3D
 * instance `inst`) to a context, or detach with p = NULL.  From then on every
 * evaluation of the context (mia_hip_cost_grad, _device, the captured graph,
 * mia_hip_optimize) adds weight * c^T P c to the instance's cost and
 * -weight * d(c^T P c)/dc to its gradient on the device, c being the
 * parameters rounded to float: the penalty term of
 * TNonrigRegGradientProblem::evaluate_fdf (mia/template/nonrigidregister.cxx:
 * 422-429) with C3DSplineTransformPenalty::value_and_gradient's weight and
 * sign (mia/3d/splinetransformpenalty.cc:52-71; weight already divided by the
 * range's volume for normalize=1).  The context does not own p: it must
 * outlive the attachment, and evaluations of one penalty object must not
 * overlap (its partial sums live in the object). */
int mia_hip_set_penalty(mia_hip_ctx *ctx, int inst, mia_hip_divcurl *p, double weight);

/* ---- allocation-free evaluation ----------------------------------------- */

/* Size of the scratch memory of mia_hip_cost_grad_noalloc: one array of this
 * many bytes per GPU the library runs on, held until the process ends. */
#define MIA_HIP_NOALLOC_SCRATCH_BYTES (64u << 20)

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Cost and gradient of ONE instance, as mia_hip_setup +
 * mia_hip_set_images_device + mia_hip_set_cost + mia_hip_cost_grad_device
 * compute them for a batch of that one instance (same kernels, same
 * results), with every buffer beyond the inputs and outputs (the instance's
 * tables, the prefiltered moving image, the warped image, the stage-1 partial
 * sums) in the library's file-scope scratch memory of
 * MIA_HIP_NOALLOC_SCRATCH_BYTES on `device`.  Calls no hipMalloc, hipFree or
 * allocating library routine, and does not synchronise: everything is
 * enqueued on `stream` (a hipStream_t, NULL for the null stream).
 *   size, csize : x, y[, z] image and coefficient grid sizes (dim entries)
 *   d_moving, d_fixed : device float images (x fastest), not modified
 *   d_params : device doubles, 2 or 3 per coefficient (interleaved)
 *   d_grad   : device doubles, same layout, overwritten
 *   d_cost   : one device double
 * MIA_HIP_ERR_LIMIT when the instance needs more scratch memory than that
 * (about 17 bytes per voxel plus the partial sums: a 3D volume of 3.5 M
 * voxels at spline rate 8, a 2D frame of 1800 x 1800 pixels).  Every call on
 * one GPU uses that GPU's one array and it keeps the last call's contents
 * (nothing in it needs to start at zero): calls that overlap on one GPU
 * (different streams, or host threads) race and must be serialised by the
 * caller. */
int mia_hip_cost_grad_noalloc(int device, int dim, const uint32_t *size, const uint32_t *csize, int degree,
                              int norm, float autothresh, double weight, int grad_mode, const float *d_moving,
                              const float *d_fixed, const double *d_params, double *d_grad, double *d_cost,
                              void *stream);

#ifdef __cplusplus
}
#endif

#endif
