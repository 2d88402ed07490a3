/*
 * mia_host.h -- C ABI of the host-side registration layer (libmia_host.so)
 * built on libmia_hip: batched series registration of independent frames,
 * the MI355X counterpart of `mia-2dmany2one-nonrigid`'s pfor over frames
 * (src/2dmany2one-nonrigid.cc:102-118,185) with the nonrigid driver's
 * single-level loop (mia/template/nonrigidregister.cxx:192-296).
 *
 * Status codes are those of mia_hip.h; mia_host_last_error() gives the
 * message of the last failure on the calling thread.
 */
#ifndef MIA_HOST_H
#define MIA_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* run modes */
#define MIA_HOST_BATCHED_GRAPH 0 /* one context, one captured graph per iteration for all frames */
#define MIA_HOST_BATCHED 1       /* one context, eager launches per iteration for all frames */
#define MIA_HOST_PER_FRAME 2     /* one context per frame, frames registered one after the other */
#define MIA_HOST_DEVICE_BATCHED 3   /* on-device L-BFGS, all frames in one graph per iteration block */
#define MIA_HOST_DEVICE_PER_FRAME 4 /* on-device L-BFGS, one context (and graph) per frame, one after the other */
#define MIA_HOST_DEVICE_SPLIT 5     /* on-device L-BFGS, the frames in three batches of about equal pixel counts
                                       whose iterations overlap (mia_hip_optimize_many); fewer frames than
                                       three: one batch per frame */

typedef struct mia_host_series mia_host_series;

const char *mia_host_last_error(void);

/* n_frames independent registrations of dimension dim (2 or 3), frame sizes
 * sizes[3*i..] (x,y,z), transformation `spline:rate=<rate>` (cubic kernel),
 * image interpolation bspline:d=<degree>, cost image:cost=ssd (norm=false).
 * Coefficient grids follow the reference's formula
 * (mia/2d/transform/spline.cc:103-116, mia/3d/transform/spline.cc:130-143). */
int mia_host_series_create(int device, int dim, int n_frames, const uint32_t *sizes, float rate, int degree,
                           mia_host_series **out);

/* Frame i's moving (src) and fixed (ref) images.  Both are mapped together to
 * mean 0 and standard deviation 1 before registration, as the reference's
 * FScaleFilterCreator does (mia/template/nonrigidregister.cxx:147-184,
 * 204-210). */
int mia_host_series_set_frame(mia_host_series *s, int frame, const float *src, const float *ref);

/* Number of parameters of frame i and its offset in the concatenated vector. */
int mia_host_series_n_params(mia_host_series *s, int frame, size_t *n, size_t *offset);

/* Register every frame from the identity with the minimizer `minimizer`
 * ("lbfgs:maxiter=100[,memory=10,gtol=0,ftolr=0,step=1]" or MIA's
 * "gdsq:maxiter=100[,step=0.1,scale=2,xtola=0,gtola=0,ftolr=0]") in run mode
 * `mode`.  The device modes run lbfgs only (gtol = ftolr = 0).  Outputs: params_out (concatenated best parameters), costs_out
 * (best cost per frame), evals_out (evaluations per frame), seconds (wall time
 * of the optimisation loop, set-up excluded).  Any output may be NULL. */
int mia_host_series_run(mia_host_series *s, const char *minimizer, int mode, double *params_out, double *costs_out,
                        int *evals_out, double *seconds);

int mia_host_series_destroy(mia_host_series *s);

/* ---- multiresolution nonrigid registration ------------------------------ */

/* Per-level record of mia_host_nonrigid_register. */
typedef struct {
    uint32_t size[3];       /* image size of the level */
    uint32_t coeff_size[3]; /* coefficient grid after the level's refinement */
    int evaluations;        /* cost + gradient evaluations at the level */
    double cost;            /* best cost (costs + penalty) at the level */
    double seconds;         /* wall time of the level */
    double minimize_seconds; /* of seconds: in the minimisations (device optimizer set-up and graph capture
                                included) */
    double loop_seconds;    /* of those: the device optimizer's iteration loop (0 with a host minimizer) */
} mia_host_level_stats;

/* Register src to ref (dim 2 or 3, size x,y[,z], float images) as the
 * reference's nonrigid driver does (mia/template/nonrigidregister.cxx:192-296):
 * images mapped jointly to mean 0 / sigma 1, then for each of `levels`
 * levels (image size / 2^k): create or upscale the transformation, resample
 * the images, minimise, refine the coefficient grid, minimise again if it
 * grew.  transform: "spline:rate=..,penalty=[divcurl:weight=..]"; cost: one
 * or more full costs "image:cost=[ssd],weight=1" joined by '+'; minimizer:
 * "lbfgs:maxiter=.." or "gdsq:..".  Outputs (any may be NULL): the final
 * parameters (params_cap doubles available, *n_params set to the count),
 * the final coefficient grid, per-level stats (up to max_levels), the final
 * cost of the registered images at full size.  Runs on `device`. */
int mia_host_nonrigid_register(int device, int dim, const uint32_t *size, const float *src, const float *ref,
                               const char *transform, const char *cost, const char *minimizer, int levels,
                               double *params_out, size_t params_cap, size_t *n_params, uint32_t *coeff_size,
                               mia_host_level_stats *stats, int max_levels, double *final_cost);

/* ---- transformation files ----------------------------------------------- */

/* Save / load a spline transformation in the layout of MIA's "xml"
 * transformation IO (mia/3d/transio/xml.cc, mia/3d/transio/serialization.hh:36-79):
 * creator string, size (x, y[, z]), attributes, parameters.  Save creates
 * the transformation from `creator` and `size` and requires n to match its
 * degrees of freedom.  Load re-creates it from the file's creator string and
 * size and restores its attributes; outputs (any may be NULL): size,
 * coefficient grid, parameters (params_cap doubles available), their count,
 * the creator string.
 * Attributes (the transformation's CAttributedData map in string form,
 * mia/core/attributes.hh:259-349) travel as key/value pairs packed into one
 * buffer: "key\0value\0key\0value\0\0" (each string NUL-terminated, an empty
 * key ends the list; keys non-empty).  Save: `attributes` may be NULL (none).
 * Load: `attributes` (attributes_len bytes, may be NULL) receives the pairs
 * in key order; MIA_HIP_ERR_SIZE if they do not fit. */
int mia_host_transform_save(const char *path, int dim, const uint32_t *size, const char *creator,
                            const double *params, size_t n, const char *attributes);
int mia_host_transform_load(const char *path, int dim, uint32_t *size, uint32_t *coeff_size, double *params,
                            size_t params_cap, size_t *n_params, char *creator, int creator_len, char *attributes,
                            int attributes_len);

#ifdef __cplusplus
}
#endif

#endif
