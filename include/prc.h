/* prc.h -- C ABI of the MI355X positron-range-correction library (libprc_amd.so).
 *
 * The reference (jlherraiz/PRC_MLEM) ships one CUDA executable, prc.cu, whose only
 * programmatic interface is the internal `Projector` class (prc.cu:403-450) and the
 * solver loop in main() (prc.cu:659-710). This header is the thin extern "C" layer a
 * host program binds instead: opaque handles, plain pointers and sizes, int status
 * codes (0 = ok, negative = error, message in prc_last_error()), no C++ exceptions or
 * exit() across the boundary. Each entry point cites the reference code it replaces.
 *
 * Conventions: volumes are float32, X fastest, index (k*NY + j)*NX + i (prc.cu:188).
 * Pointers named dev_* are device pointers on the library's current HIP device; all
 * other pointers are host memory. Handles are not thread-safe; calls on one handle must
 * not overlap. Nothing in this library touches the GPU before the first create call.
 */
#ifndef PRC_H
#define PRC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRC_OK 0
#define PRC_ERR_INVALID (-1) /* bad argument (geometry, pointer, size) */
#define PRC_ERR_DEVICE (-2)  /* HIP runtime error */
#define PRC_ERR_MEMORY (-3)  /* device allocation failed */
#define PRC_ERR_NO_ACTIVITY (-4)

#define PRC_RANGE_BINS 4096 /* survival table bins on [0, cutoff] (prc.cu:61) */
#define PRC_NEAR_RAYS 2048  /* near-block rays per source (prc.cu:104) */

/* Weight storage precision of the projector. */
enum {
    PRC_PRECISION_FP32 = 0,  /* plain: FP32 weights, FP32 march (the correctness anchor) */
    PRC_PRECISION_MIXED = 1  /* FP16 weights (half the HBM stream), FP32 march and sums */
};

/* Projector flags. */
enum {
    PRC_FLAG_EXACT_RANGE = 1,  /* --exact-range: analytic survival per step (prc.cu:75, 625) */
    PRC_FLAG_NO_CACHE_NEAR = 2, /* accepted and ignored (the CLI maps --no-cache-near, the reference's
                                   memory-saving switch, to PRC_FLAG_MARCH_PASSES) */
    PRC_FLAG_NO_NEAR_MEMO = 4,  /* march every source's near block, even where 5x5x5 uniform
                                   neighbourhoods would share one (same result, for testing) */
    PRC_FLAG_MARCH_PASSES = 8   /* store the near block only and march the far rays in every forward
                                   and adjoint, as the build does by itself when both stored copies
                                   would not fit in device memory (the reference retraces when its
                                   near cache does not fit, prc.cu:411-418). Same operator; the
                                   forward's far part is summed by atomic additions, so its last
                                   bits may vary between runs */
};

typedef struct prc_projector prc_projector;
typedef struct prc_solver prc_solver;
typedef struct prc_batch prc_batch;

/* Geometry and model of one volume: the `Grid` of prc.cu:94-100 plus Options
 * (prc.cu:452-459) that shape the system matrix. */
typedef struct {
    int nx, ny, nz;    /* voxels; nx*ny*nz <= INT_MAX (prc.cu:547-551) */
    float sx, sy, sz;  /* voxel size [mm] */
    int isotope;       /* index from prc_isotope_find() */
    int n_rays;        /* far-field rays per voxel: even, >= 2 (prc.cu:540) */
    float support_rho; /* voxels with density >= support_rho hold activity (default 0.05) */
    float air_cap_mm;  /* a ray is dropped after this many mm of consecutive air (default 50) */
    int precision;     /* PRC_PRECISION_* */
    int flags;         /* PRC_FLAG_* */
} prc_geometry;

/* Solver settings (prc.cu:456-458). */
typedef struct {
    float beta;     /* RDP strength, >= 0; 0 = Richardson-Lucy */
    float gamma;    /* RDP edge parameter, >= 0 */
    int accelerate; /* Biggs-Andrews extrapolation on/off */
} prc_solve_params;

/* Sizes and counters of a built projector. */
typedef struct {
    int64_t n_voxels;
    int64_t n_active;        /* voxels with rho >= support_rho */
    int64_t n_rows;          /* X rows holding at least one active voxel */
    int64_t weight_entries;  /* source-major stored entries (the adjoint's planes, incl. padding) */
    int64_t target_entries;  /* target-major stored entries (the forward's planes, incl. padding) */
    int64_t weight_bytes;    /* device bytes of both stored copies */
    int64_t vrs_near;        /* voxel-ray steps of the near-block trace (normalisation) */
    int64_t vrs_far;         /* voxel-ray steps of one far-field pass over all sources */
    int64_t deposits;        /* nonzero stored far-field weights */
    int64_t near_uniform;    /* sources whose near block was copied from a uniform-density template */
    double norm_seconds;     /* wall time of the normalisation build */
    float cutoff_mm;         /* water-equivalent range cutoff (prc.cu:83-91, 587) */
    int64_t weight_planes;   /* source-major planes: entries of the adjoint's shift list */
    int64_t target_planes;   /* target-major planes: entries of the forward's shift list */
    int64_t far_marched;     /* 1 when the far rays are marched in every pass (near block stored only) */
    int64_t nonzero_entries; /* positive weights among the source-major entries (the rest are padding
                                of lanes whose ray ended before the segment's last stored step) */
} prc_stats;

/* Library identification. */
const char* prc_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* prc_last_error(void);

/* Isotope model (prc.cu:38-57, 570-584); host only, no GPU. */
int prc_isotope_count(void);
const char* prc_isotope_name(int index);
int prc_isotope_find(const char* name); /* index or -1 */
/* Water-equivalent distance that `fraction` of positrons does not exceed, FP32 bisection
 * on the renormalised model (prc.cu:83-91). */
float prc_range_cutoff(int isotope, float fraction);

/* The far ray set of prc.cu:118-129 as the device computes it (the march tables are built from
 * it): start offsets in voxels (3 floats per ray) and unit directions (3 floats per ray) of rays
 * 0 .. n - 1 of an n-ray set; n even, 2 <= n <= 2^26. For validation. */
int prc_ray_set(int n, float* host_start, float* host_dir);

/* Build the projector of one volume: density upload (clamped to >= 1e-3, prc.cu:603),
 * device range table, ray tables, and the normalisation pass that stores every
 * system-matrix weight once (replaces Projector::Projector, prc.cu:405-422).
 * host_rho: nx*ny*nz finite floats. host_table: optional 4097-entry survival table that
 * replaces the device-built one (validation); NULL = build on the device. */
int prc_projector_create(const prc_geometry* geometry, const float* host_rho, const float* host_table,
                         prc_projector** out);
void prc_projector_destroy(prc_projector* projector);
/* Run the normalisation pass again on the resident density (same result; reuses the weight
 * buffer). Waits for work queued on this projector, runs on the projector's own stream and
 * returns when the build is complete. For callers that time or repeat the build without
 * re-uploading. */
int prc_projector_normalise(prc_projector* projector);

/* Normalisation of several projectors at once, each on its own stream so that their kernels
 * overlap; same result as prc_projector_normalise on each. Returns when all are done. No
 * counterpart in the reference (one projector per run); for batches of volumes. */
int prc_projectors_normalise(int n, prc_projector* const* projectors);

/* dev_out = A dev_in, overwriting every voxel (Projector::forward, prc.cu:430-434). */
int prc_projector_forward(prc_projector* projector, const float* dev_in, float* dev_out);
/* The same as prc_projector_forward, allocating nothing. No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation). The staging vector f / W lives in a static device array of
 * 32 Mi floats (128 MiB, one per GPU, held until the process ends) instead of a buffer allocated on
 * first use; a volume that needs more (voxels plus two gather margins) is rejected with
 * PRC_ERR_INVALID. Calls that overlap on one GPU share that array and race. */
int prc_projector_forward_noalloc(prc_projector* projector, const float* dev_in, float* dev_out);
/* dev_out = A^T dev_in on active voxels; other voxels are left untouched
 * (Projector::adjoint, prc.cu:436). */
int prc_projector_adjoint(prc_projector* projector, const float* dev_in, float* dev_out);
/* Host-buffer forms of the two operators (copy in, apply, copy out). */
int prc_projector_forward_host(prc_projector* projector, const float* in, float* out);
int prc_projector_adjoint_host(prc_projector* projector, const float* in, float* out);
/* Per-voxel normalisation W (0 off support) (NormOp, prc.cu:232-237). */
int prc_projector_norm(prc_projector* projector, float* host_w);
/* The device survival table actually used (4097 floats) and its bins per mm. */
int prc_projector_table(prc_projector* projector, float* host_table, float* scale);
int prc_projector_stats(prc_projector* projector, prc_stats* stats);
/* Device pointer to the per-voxel W, for callers that chain device kernels. */
const float* prc_projector_norm_device(prc_projector* projector);

/* MAP-EM / Biggs-Andrews solver on one projector (prc.cu:659-710). host_g: measured
 * (blurred) activity; it is clamped to >= 0 and zeroed off support (prc.cu:604-612).
 * Returns PRC_ERR_NO_ACTIVITY when the supported counts sum to 0 (prc.cu:619-622).
 * The estimate starts as the preprocessed input; the initial floor and count
 * normalisation (prc.cu:669-672) run on the first call of prc_solver_iterate. */
int prc_solver_create(prc_projector* projector, const float* host_g, const prc_solve_params* params,
                      prc_solver** out);
void prc_solver_destroy(prc_solver* solver);
/* Restart from the preprocessed input already on the device (as after create). */
int prc_solver_reset(prc_solver* solver);
/* Run n more iterations on the device (no host synchronisation inside). */
int prc_solver_iterate(prc_solver* solver, int n);
/* Wait for queued work. */
int prc_solver_sync(prc_solver* solver);
/* Copy the current estimate to the host (prc.cu:713). */
int prc_solver_estimate(prc_solver* solver, float* host_out);
/* Device pointer to the current estimate (valid until the next iterate call). */
const float* prc_solver_estimate_device(prc_solver* solver);
/* host_out = A f of the current estimate (--reproject, prc.cu:714-717). */
int prc_solver_reproject(prc_solver* solver, float* host_out);
/* Sum of the current estimate in FP64 and the supported input counts. */
int prc_solver_sums(prc_solver* solver, double* estimate_sum, double* support_sum);

/* Batched correction of n independent volumes (no counterpart in the reference, which runs
 * one volume per process). The volumes' data are concatenated on the device; every stage of
 * an iteration is one launch over all volumes' work, and the iteration loop is captured once
 * as a HIP graph (per iteration count) and replayed. Results equal separate solver runs.
 * Volumes must share the precision and the acceleration flag; beta and gamma may differ. */
int prc_batch_create(int n, prc_projector* const* projectors, const float* const* host_g,
                     const prc_solve_params* params, prc_batch** out);
/* Normalise every projector of the batch again (same result as prc_projector_normalise on
 * each), on the batch's own stream: it returns once the builds' buffer sizes are known and
 * leaves the rest queued, so a prc_batch_run that follows sees the new weights while the
 * iterations of other batches keep running beside the build. No counterpart in the reference
 * (Projector::Projector, prc.cu:405-422, builds one volume on the default stream). */
int prc_batch_normalise(prc_batch* batch);
/* Restart every volume from its data and run `iterations` (asynchronous). The batch reads
 * the projectors' current weights, so a projector may be normalised again between runs. */
int prc_batch_run(prc_batch* batch, int iterations);
int prc_batch_sync(prc_batch* batch);
/* Copy volume v's estimate after the last run (synchronises). */
int prc_batch_estimate(prc_batch* batch, int v, float* host_out);
void prc_batch_destroy(prc_batch* batch);
/* create + run + estimates + destroy in one call. */
int prc_batch_correct(int n, prc_projector* const* projectors, const float* const* host_g,
                      const prc_solve_params* params, int iterations, float* const* host_out);

/* Kernel timing: while enabled, every launch of the library is bracketed by HIP events on its
 * stream (graph replay is bypassed). The report is JSON {name: {count, total_ms, mean_ms,
 * min_ms, max_ms}} and clears the records. No counterpart in the reference. */
int prc_profile_enable(int on);
int prc_profile_report(char* json, size_t capacity);

/* JSON array of the projector-pass tuning decisions made so far: per solver or batch, the mean
 * launch time of each launch shape of the forward and the adjoint ("grid": one wave per row
 * segment in launch order; "chunked": the same with each XCD's blocks renumbered to one contiguous
 * chunk of the segments; "queue@B": a persistent grid of B = 1, 2 or 4 blocks per CU taking
 * segments from XCD-local work queues; "+sc0": gathers served from L2 without allocating in the
 * CU's L1) and the one kept. No counterpart in the reference (prc.cu has one fixed launch shape);
 * PRC_TUNE=0 disables tuning (queue@2), PRC_TUNE=<shape> forces one ("queue" alone is queue@2). */
int prc_tuning_report(char* json, size_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* PRC_H */
