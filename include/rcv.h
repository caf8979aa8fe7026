/*
 * rcv.h -- C ABI of librcv.so, the MI355X (gfx950) implementation of the
 * recursive-control-variate inverse-rendering hot path.
 *
 * Conventions
 *   - Every entry point returns an int status: RCV_OK (0) or a negative
 *     RCV_E_* code. rcv_last_error() returns a thread-local message for the
 *     last failing call on the calling thread.
 *   - The caller allocates every device buffer (float pointers below are HIP
 *     device pointers unless stated otherwise); the library never frees caller
 *     memory. Host arrays (geometry, transforms) are copied during the call.
 *   - `stream` is a hipStream_t passed as void* (NULL = legacy default stream).
 *     All device work is enqueued on it; no call synchronises the host unless
 *     stated.
 *   - Scene handles are thread-compatible (one thread at a time); a CV state is
 *     not re-entrant: two rcv_cv_step calls on one state must be ordered on one
 *     stream.
 *   - Images are row-major (height, width, 3) float32. Textures are
 *     (height, width, channels); grids are (depth, height, width, channels).
 */
#ifndef RCV_H
#define RCV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RCV_OK 0
#define RCV_E_INVALID (-1)   /* bad argument / shape / id */
#define RCV_E_STATE (-2)     /* call not valid in the object's current state */
#define RCV_E_HIP (-3)       /* a HIP runtime call failed */
#define RCV_E_SCENE (-4)     /* scene check failed (see plugins/twostatepath.py:152-166) */
#define RCV_E_NOMEM (-5)

#define RCV_STATE_OLD 0
#define RCV_STATE_NEW 1

/* parameter kinds */
#define RCV_PARAM_TEXTURE 0 /* (ny, nx, channels) bitmap, bilinear, repeat */
#define RCV_PARAM_GRID 1    /* (nz, ny, nx, channels) grid, trilinear, clamp */
#define RCV_PARAM_VALUE 2   /* a constant of `channels` (1 or 3) values, nx = ny = nz = 1
                               (a homogeneous medium's sigma_t / albedo) */

/* bsdf types */
#define RCV_BSDF_DIFFUSE 0
#define RCV_BSDF_NULL 1

/* emitter types */
#define RCV_EMITTER_AREA 0     /* attached to a mesh shape */
#define RCV_EMITTER_CONSTANT 1 /* constant environment */

/* integrator types */
#define RCV_INTEGRATOR_PATH 0   /* twostateprb / prb  (surfaces)  */
#define RCV_INTEGRATOR_VOLPATH 1 /* twostateprbvolpath / prbvolpath */
#define RCV_INTEGRATOR_VOLPATH_SIMPLE 2 /* volpathsimple (plugins/volpathsimple.py):
                                           one state, DRT adjoint */

/* rcv_integrator_desc.flags for RCV_INTEGRATOR_VOLPATH_SIMPLE (bits set
 * disable; 0 = the reference defaults, plugins/volpathsimple.py:33-48) */
#define RCV_SIMPLE_NO_NEE 1
#define RCV_SIMPLE_NO_DRT 2
#define RCV_SIMPLE_NO_DRT_SUBSAMPLING 4
#define RCV_SIMPLE_NO_DRT_MIS 8

/* loss kinds */
#define RCV_LOSS_L1 0
#define RCV_LOSS_L2 1

typedef struct rcv_scene rcv_scene;

/* Integrator properties (plugins/twostatepath.py:106-118,147-150;
 * plugins/twostatevolpath.py:20-32,62-65). */
typedef struct rcv_integrator_desc {
  int32_t type;          /* RCV_INTEGRATOR_* */
  int32_t max_depth;     /* -1 = unbounded */
  int32_t rr_depth;
  int32_t hide_emitters; /* 0/1 */
  int32_t samples_per_thread; /* launch shape: camera samples traced serially
                                 per thread (1..spp, divides spp); 0 = auto.
                                 Does not change results beyond float
                                 summation order of the pixel average. */
  int32_t block_size;    /* threads per block, 64..256 multiple of 64; 0 = 256 */
  int32_t flags;         /* integrator-specific switches (RCV_SIMPLE_*) */
} rcv_integrator_desc;

/* A medium: heterogeneous, with an axis-aligned grid box (mitsuba
 * 'heterogeneous' with a 'gridvolume' sigma_t, as used by
 * experiments.py:52-69), or homogeneous (mitsuba 'homogeneous', the medium of
 * tutorial.ipynb:77-88): constant sigma_t and albedo, unbounded. A scene holds
 * at most one homogeneous medium. */
typedef struct rcv_medium_desc {
  int32_t sigma_t_param;   /* RCV_PARAM_GRID id (1 channel); homogeneous:
                              RCV_PARAM_VALUE id (1 channel) */
  int32_t albedo_param;    /* RCV_PARAM_GRID id (3 channels), homogeneous:
                              RCV_PARAM_VALUE id (3 channels); or -1 */
  float albedo_rgb[3];     /* used when albedo_param < 0 */
  float scale;             /* sigma_t = scale * grid (or value) */
  float bbox_min[3];       /* world-space box the grid spans (heterogeneous) */
  float bbox_max[3];
  int32_t sample_emitters; /* NEE inside the medium (default 1) */
  int32_t majorant_factor; /* 0 = one global majorant; k>0 = supergrid with
                              cells of k^3 voxels traversed by DDA
                              (plugins/twostatemedium.py:25-70); 0 when
                              homogeneous */
  int32_t homogeneous;     /* 1: homogeneous medium (bbox ignored) */
} rcv_medium_desc;

/* ---------------------------------------------------------------- library */
int rcv_version(void);
const char *rcv_last_error(void);

/* ------------------------------------------------------------------ scene
 * Replaces the scene assembled by mi.load_dict for the hot path: shapes,
 * BSDFs (incl. the 'twostate' adapter, plugins/twostatebsdf.py:4-56), media
 * (incl. 'twostatemedium', plugins/twostatemedium.py:5-200), emitters and the
 * (batch) sensor. Build with rcv_scene_add_*, then rcv_scene_commit. */
int rcv_scene_create(rcv_scene **out);
/* RCV_E_STATE (and nothing freed) while a batch holds the scene. */
int rcv_scene_destroy(rcv_scene *scene);

/* Declares a differentiable parameter (texture or grid). */
int rcv_scene_add_param(rcv_scene *scene, int32_t kind, int32_t nx, int32_t ny,
                        int32_t nz, int32_t channels, int32_t *out_id);
/* Binds the device data of a parameter for one state (RCV_STATE_OLD/NEW).
 * Replaces TwoState*.parameters_changed (plugins/twostatebsdf.py:46-56,
 * plugins/twostatemedium.py:189-200): instead of copying old<-new<-incoming
 * the host rebinds pointers. For grids bound to a medium, the state's maximum
 * (majorant, plugins/twostatemedium.py:73-75) is reduced on `stream`. */
int rcv_scene_bind_param(rcv_scene *scene, int32_t param, int32_t state,
                         const float *device_data, void *stream);
/* Changes the resolution of grid/texture parameter `param` after commit (the
 * resolution change of the upsampling step, utils.py:177-192, where the
 * Mitsuba volume or bitmap behind the key is replaced by a larger one). Both
 * states become unbound and must be bound again at the new size; majorant
 * supergrids of media over the grid are re-planned. RCV_E_STATE while a
 * batch holds the scene. */
int rcv_scene_resize_param(rcv_scene *scene, int32_t param, int32_t nx, int32_t ny,
                           int32_t nz, void *stream);
/* Medium.set_majorant_resolution_factor (plugins/twostatemedium.py:174-178,
 * called by adjust_majorant_res_factor, utils.py:27-43): 0 = one global
 * majorant, k > 0 = supergrid of ceil(n / k) cells per axis, rebuilt for the
 * bound states. RCV_E_STATE while a batch holds the scene. */
int rcv_scene_set_majorant_factor(rcv_scene *scene, int32_t medium, int32_t factor,
                                  void *stream);

int rcv_scene_add_bsdf(rcv_scene *scene, int32_t type, const float rgb[3],
                       int32_t texture_param, int32_t *out_id);
int rcv_scene_add_medium(rcv_scene *scene, const rcv_medium_desc *desc,
                         int32_t *out_id);
int rcv_scene_add_emitter(rcv_scene *scene, int32_t type, const float rgb[3],
                          int32_t *out_id);
/* Triangle mesh; `uv` may be NULL. interior/exterior are medium ids or -1;
 * `emitter` is an RCV_EMITTER_AREA id or -1. */
int rcv_scene_add_mesh(rcv_scene *scene, const float *vertices, const float *uv,
                       int32_t n_vertices, const int32_t *triangles,
                       int32_t n_triangles, int32_t bsdf, int32_t emitter,
                       int32_t interior, int32_t exterior, int32_t *out_id);
int rcv_scene_add_sphere(rcv_scene *scene, const float center[3], float radius,
                         int32_t bsdf, int32_t interior, int32_t exterior,
                         int32_t *out_id);
/* Oriented box: the image of [-1, 1]^3 under the affine map `to_world`
 * (row-major 3x4), i.e. Mitsuba's 'cube' shape, which the reference's volume
 * scenes use as the null-BSDF medium boundary (experiments.py:74-79). Traced
 * as one analytic primitive; hit points, outward face normals and the
 * per-face uv layout are those of the cube's 12-triangle mesh. Not an area
 * emitter (an emitting cube is added as a mesh). */
int rcv_scene_add_box(rcv_scene *scene, const float to_world[12], int32_t bsdf,
                      int32_t interior, int32_t exterior, int32_t *out_id);
/* A (batch) sensor: n_views perspective cameras side by side on one film of
 * film_width x film_height; view k covers columns [k*w/n, (k+1)*w/n).
 * to_world: n_views row-major 4x4 camera-to-world matrices (columns: left,
 * up, forward, origin); fov_x: horizontal field of view in degrees per view.
 * (mitsuba 'batch' + 'perspective', experiments.py:28-50.) */
int rcv_scene_add_sensor(rcv_scene *scene, int32_t film_width,
                         int32_t film_height, int32_t n_views,
                         const float *to_world, const float *fov_x,
                         int32_t *out_id);
/* Builds the BVH and uploads geometry. No add_* call is valid afterwards. */
int rcv_scene_commit(rcv_scene *scene, void *stream);
/* check_scene of the two-state integrators (plugins/twostatepath.py:152-166,
 * plugins/twostatevolpath.py:67-78): returns RCV_E_SCENE with a message if
 * the scene lacks the required two-state adapter. `twostate_flags` bit0 =
 * bsdf i is a two-state adapter (array of n_bsdf), same for media. */
int rcv_scene_check_twostate(const rcv_scene *scene, int32_t integrator_type,
                             const int32_t *bsdf_twostate, int32_t n_bsdf,
                             const int32_t *medium_twostate, int32_t n_media);

/* ----------------------------------------------------------------- render
 * Plain render of the NEW state: integrator.render(scene, sensor, seed, spp)
 * of the upstream prb / prbvolpath base class (reached from
 * plugins/metaintegrator.py:87-88). `img`: (film_h, film_w, 3). */
int rcv_render(const rcv_scene *scene, const rcv_integrator_desc *integrator,
               int32_t sensor, uint32_t seed, int32_t spp, float *img,
               void *stream);
/* Two-state render: render_twostates (plugins/twostatepath.py:179-223,
 * plugins/twostatevolpath.py:91-138). One path per sample carries both
 * throughputs; img_old = H, img_new = F. */
int rcv_render_twostates(const rcv_scene *scene,
                         const rcv_integrator_desc *integrator, int32_t sensor,
                         uint32_t seed, int32_t spp, float *img_old,
                         float *img_new, void *stream);
/* Adjoint: render_backward (plugins/cv_integrator.py:87-94, inherited PRB /
 * PRBVolpath adjoint on the NEW state). grad_img: (film_h, film_w, 3) =
 * dLoss/dimage. Gradients are ACCUMULATED (+=) into grad_ptrs[i], which has
 * the shape of parameter param_ids[i]. Path replay: nothing is taped. The
 * volume path tracer's adjoint (samples_per_thread 0) logs density
 * scatter-adds in a buffer that the library allocates on first use per
 * (device, stream), about 3.3 GB, and keeps until the process ends; calls on
 * one stream share it in stream order. */
int rcv_render_backward(const rcv_scene *scene,
                        const rcv_integrator_desc *integrator, int32_t sensor,
                        uint32_t seed, int32_t spp, const float *grad_img,
                        int32_t n_params, const int32_t *param_ids,
                        float *const *grad_ptrs, void *stream);
/* Forward mode: render_forward (plugins/cv_integrator.py:79-85,
 * plugins/metaintegrator.py:132-138, delegating to the upstream PRB /
 * PRBVolpath forward mode) of the NEW state. dimg (film_h, film_w, 3) =
 * d(image)/d(theta) . tangent, where tangents[i] (shaped like parameter
 * param_ids[i]) is the tangent of that parameter and the others are held
 * fixed. Paths are replayed with the same random numbers as the primal pass,
 * so for one seed <g, dimg> equals <adjoint(g), tangent> of
 * rcv_render_backward up to float rounding. RCV_E_INVALID for
 * volpathsimple, whose adjoint exists in backward mode only. */
int rcv_render_forward(const rcv_scene *scene,
                       const rcv_integrator_desc *integrator, int32_t sensor,
                       uint32_t seed, int32_t spp, int32_t n_params,
                       const int32_t *param_ids, const float *const *tangents,
                       float *dimg, void *stream);
/* Device event counters of the last render on this scene (tentative
 * collisions, NEE ratio-tracking steps, surface hits, paths): copies
 * `n` int64 counters to host memory `out` (synchronises `stream`). Counting
 * is enabled by rcv_scene_set_counting(scene, 1). */
int rcv_scene_set_counting(rcv_scene *scene, int32_t enabled);
int rcv_scene_read_counters(const rcv_scene *scene, int64_t *out, int32_t n,
                            void *stream);
/* Majorant supergrid of medium `medium` for one state (replaces
 * Medium.majorant_grid(), read at plugins/twostatemedium.py:25-29,52-53,
 * 82-84): writes the resolution (x, y, z) to `res` ((0, 0, 0) when the medium
 * uses one global majorant) and, when `out` is not null, copies the unscaled
 * cell maxima, x fastest, to host memory `out` of `capacity` floats
 * (synchronises `stream`). A medium created with majorant_factor k > 0 has
 * ceil(n / k) cells per axis of an n-voxel grid; each cell holds the maximum
 * of every voxel a trilinear lookup inside the cell reads. */
int rcv_scene_majorant_grid(const rcv_scene *scene, int32_t medium, int32_t state,
                            float *out, int64_t capacity, int32_t *res, void *stream);

/* --------------------------------------------------------------- CV state
 * Per-pixel control-variate state of CVIntegrator / MetaIntegrator
 * (plugins/cv_integrator.py:32-45, plugins/welford.py:44-57): eight float32
 * arrays of n_elem elements and a device int32 counter block
 * counters[0] = it, counters[1] = n (StatisticsEstimator.n), counters[2..3]
 * library scratch; all zero at init_buffers. */
typedef struct rcv_cv_state {
  float *img;
  float *v_n;
  float *w_s;
  float *mean_x;
  float *mean_y;
  float *var_x;
  float *var_y;
  float *cov;
  int32_t *counters;
} rcv_cv_state;

/* One fused combine: weights, NaN/Inf guard, clamp, blend, statistics update
 * and it += 1 (plugins/cv_integrator.py:60-76; pre_update ordering of
 * plugins/metaintegrator.py:95-96,110-111). H = img_old, F = img_new. The
 * counters are read and advanced on the device, so the call can be captured
 * in a graph and replayed. */
int rcv_cv_step(const rcv_cv_state *state, const float *H, const float *F,
                int64_t n_elem, int32_t warmup, float beta1, float beta2,
                int32_t pre_update, void *stream);
/* WelfordVarianceEstimator.update (plugins/welford.py:27-33): mean/var arrays
 * of n_elem. `counter` points to two device int32: counter[0] = n (advanced
 * on the device), counter[1] = scratch of the last-block election; both
 * start at zero. */
int rcv_welford_update(float *mean, float *var, int32_t *counter,
                       const float *x, int64_t n_elem, float beta1,
                       float beta2, void *stream);

/* ------------------------------------------------------- loss / optimizer */
/* loss_out[0] = mean(|img-ref|) or mean((img-ref)^2) (utils.py:21-25) and
 * grad_out = dloss/dimg (may be NULL). loss_out is a device float. */
int rcv_loss(const float *img, const float *ref, int64_t n_elem, int32_t kind,
             float *loss_out, float *grad_out, void *stream);
/* rcv_loss without any allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory at call time (code under HIP
 * graph capture, tight serving loops). Same arguments and results (bit for
 * bit) as rcv_loss; its partial sums live in a library-owned static device
 * array of 256 floats + a counter per device instead of a workspace that
 * rcv_loss allocates on its first call. Calls that overlap on one device
 * (concurrent streams) race on that array and must be ordered. */
int rcv_loss_noalloc(const float *img, const float *ref, int64_t n_elem,
                     int32_t kind, float *loss_out, float *grad_out,
                     void *stream);
/* Adam (mi.ad.Adam as used at optimize.py:90,127) fused with clamp
 * (utils.py:172-175): param_out = clamp(param_in - lr_t*m/(sqrt(v)+eps)),
 * with m, v updated in place; t is the 1-based step count. param_out may
 * equal param_in. */
int rcv_adam_step(float *param_out, const float *param_in, const float *grad,
                  float *m, float *v, int64_t n_elem, int32_t t, float lr,
                  float beta1, float beta2, float eps, float clamp_lo,
                  float clamp_hi, void *stream);

/* ------------------------------------------------------ batched jobs
 * Many independent optimisations (each: one committed scene of its own, one
 * optimised parameter, a method, its CV state and Adam state) advanced by one
 * device-resident iteration: two-state/plain render in persistent work-queue
 * kernels over all jobs, CV combine, loss gradient, path-replay adjoint, Adam
 * + clamp, two-state pointer rotation. Seeds (job seed + iteration), Adam step
 * counts and majorants advance on the device, so rcv_batch_step can capture
 * the iteration into a hipGraph once and replay it. Multi-job form of
 * run_opt's loop body (optimize.py:104-129); no reference counterpart exists
 * for the batching itself. All buffers are caller-owned device memory. */
typedef struct rcv_batch rcv_batch;

typedef struct rcv_batch_job {
  const rcv_scene *scene;          /* committed, used by this job only */
  rcv_integrator_desc integrator;
  int32_t sensor;
  uint32_t seed;                   /* iteration i renders with seed + i */
  int32_t spp, spp_grad;
  int32_t method;                  /* 0 = baseline (plain), 1 = cv_ps */
  int32_t loss;                    /* RCV_LOSS_* */
  int32_t param;                   /* optimised parameter id of `scene` */
  float *param_ring[3];            /* [0] old, [1] new (both = initial value), [2] scratch */
  float *adam_m, *adam_v, *grad;   /* parameter-sized, m = v = 0 initially */
  float lr, beta1, beta2, eps, clamp_lo, clamp_hi;
  const float *img_ref;            /* film-sized reference image */
  float *H, *F, *img, *grad_img;   /* film-sized work images */
  float *cv_state[8];              /* film-sized, zero: img v_n w_s mean_x mean_y var_x var_y cov */
  int32_t warmup;
  float cv_beta1, cv_beta2;
  int32_t pre_update;
  /* resume != 0: continue a job taken from another batch (rcv_batch_read_job)
   * with this iteration count, ring slots and pending majorant key;
   * otherwise start at iteration 0 with slots (0, 1, 2). */
  int32_t resume;
  int32_t iteration;
  int32_t ring_old, ring_new, ring_next;
  uint32_t next_key;
} rcv_batch_job;

#define RCV_BATCH_MAX_JOBS 4096

int rcv_batch_create(rcv_batch **out);
int rcv_batch_destroy(rcv_batch *batch);
int rcv_batch_add_job(rcv_batch *batch, const rcv_batch_job *job, int32_t *out_job);
/* Lays the jobs out in device memory; the scenes' parameter states must be
 * bound to param_ring[0] (OLD) and param_ring[1] (NEW) beforehand. */
int rcv_batch_commit(rcv_batch *batch, void *stream);
/* Launch shape of the persistent kernels: threads per block (64..512, a
 * multiple of 64, and at most what the device's LDS per block holds at 160 B
 * per thread: RCV_E_INVALID otherwise) and work units taken per queue grab
 * (64..8192, a multiple of 64); 0 = defaults. Invalidates a captured graph. */
int rcv_batch_set_launch(rcv_batch *batch, int32_t block_size, int32_t queue_grab);
/* One iteration of every job on `stream` (must not be the legacy default
 * stream when use_graph != 0: the first call captures a hipGraph, later calls
 * replay it). */
int rcv_batch_step(rcv_batch *batch, int32_t use_graph, void *stream);
/* Iteration count of a job and the ring slot holding its latest parameter
 * value (synchronises `stream`). */
int rcv_batch_job_state(const rcv_batch *batch, int32_t job, int32_t *iteration, int32_t *ring_slot,
                        void *stream);
/* Continuous batching: bit j of `active_bits` (uint32 words) selects whether
 * job j takes part in the next iterations; inactive jobs keep their state
 * untouched. Enqueued on `stream` (outside any captured graph). */
int rcv_batch_set_active(rcv_batch *batch, const uint32_t *active_bits, int32_t n_jobs, void *stream);
/* The resumable state of a job (see rcv_batch_job.resume); synchronises. */
int rcv_batch_read_job(const rcv_batch *batch, int32_t job, int32_t *iteration, int32_t *ring_old,
                       int32_t *ring_new, int32_t *ring_next, uint32_t *next_key, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* RCV_H */
