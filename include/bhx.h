/*
 * bhx.h - C ABI of the MI355X (gfx950) Kerr geodesic tracer.
 *
 * This is the drop-in boundary for the reference's frame API
 *     img::Image bh::render(const RenderConfig&, RenderStats&)      (src/render.h:72)
 * whose body (src/render.cpp:396-497) spawns a std::thread row pool and calls
 * the internal per-ray functions trace() (src/render.cpp:69-260), shade_hit()
 * (src/render.cpp:265-375) and shade() (src/render.cpp:382-393).  Those three
 * have internal linkage, so the only seam a maintainer can replace is the
 * thread pool inside render(): the entry points below are what such a
 * replacement calls.  The C++ wrapper that keeps the reference's own
 * signature is include/bhx_render.hpp.
 *
 * Conventions
 *   - Plain C: no C++ or torch types, fixed-size integer types, FP64 data.
 *   - Every function returns 0 on success or a negative BHX_E_* code; the
 *     message of the last failure on a context is bhx_last_error(ctx).
 *   - Functions whose name ends in _dev take DEVICE pointers and a hipStream_t
 *     passed as void* (NULL = the null stream); they are stream-ordered and do
 *     not synchronise unless stated.  All other pointers are host pointers.
 *   - A context's scratch (queue, initial states, counters) is reused by its
 *     next call: issue one context's calls on one stream, or synchronise
 *     between streams; use one context per host thread.
 *   - Ray index of sample (sx, sy) of pixel (x, y):
 *         ray = (y * width + x) * S*S + sy * S + sx,     S = sqrt_spp
 *     which is the order the reference visits samples (src/render.cpp:437-459).
 *   - SoA ray arrays are component-major: component c of ray i is at
 *     [c * n_rays + i].
 */
#ifndef BHX_H
#define BHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BHX_ABI_VERSION 3

/* Error codes. */
#define BHX_OK 0
#define BHX_E_INVALID (-1)    /* bad argument (NULL pointer, size, range)   */
#define BHX_E_HIP (-2)        /* a HIP runtime call failed                  */
#define BHX_E_NOMEM (-3)      /* device allocation failed                   */
#define BHX_E_TOO_LARGE (-4)  /* request exceeds a documented maximum       */

/* Termination class of a ray: the order of Hit::What (src/render.cpp:37). */
enum bhx_class {
    BHX_HORIZON = 0,
    BHX_DISC = 1,
    BHX_SUN = 2,
    BHX_STAR = 3,       /* neutron-star surface */
    BHX_SKY = 4,
    BHX_EXHAUSTED = 5
};

/* Tone curves of img::ToneMap (src/image.h:16). */
enum bhx_tonemap {
    BHX_TM_LINEAR = 0,
    BHX_TM_REINHARD = 1,
    BHX_TM_ACES = 2,
    BHX_TM_LOG = 3,
    BHX_TM_HDR = 4
};

/*
 * One scene: RenderConfig (src/render.h:15-53) flattened into POD, including
 * the members of Camera (src/scene.h:343-351), SunBody (src/scene.h:232-235),
 * HotSpot (src/scene.h:306-313), NeutronStar (src/neutronstar.h:43-67) and
 * Jet (src/jet.h:96-145, the values after Jet::configure, src/jet.h:261-311).
 * Units are the reference's: lengths in GM/c^2, angles in radians.
 */
typedef struct bhx_scene {
    double M_kg;              /* RenderConfig::M_kg                         */
    double spin;              /* RenderConfig::spin                         */
    double cam_r, cam_theta, cam_phi, cam_fov, cam_yaw, cam_pitch, cam_roll;
    double disc_r_out;        /* RenderConfig::disc_r_out                   */
    double eddington;         /* RenderConfig::eddington                    */
    double observer_time;     /* RenderConfig::observer_time                */
    double rtol, atol;        /* RenderConfig::rtol / atol                  */
    double r_escape;          /* RenderConfig::r_escape                     */
    /* SunBody */
    double sun_centre[3], sun_radius, sun_T_eff;
    /* HotSpot */
    double hs_r, hs_phi0, hs_sigma, hs_contrast, hs_omega;
    /* NeutronStar */
    double ns_M_kg, ns_R_m, ns_R, ns_T_eff, ns_spin_hz, ns_Omega;
    double ns_cap_tilt, ns_cap_radius, ns_cap_T, ns_cap_phase0;
    /* Jet (configured) */
    double jet_z_base, jet_z_top, jet_R_base, jet_k, jet_gamma_max, jet_z_sat;
    double jet_alpha, jet_u_peak, jet_u_wid, jet_spine, jet_spine_wid, jet_zeta;
    double jet_a, jet_Omega_F, jet_j_nu0, jet_r_bound;
    uint64_t seed;            /* RenderConfig::seed (jitter + star catalogue) */
    int32_t width, height, sqrt_spp, max_steps;
    int32_t disc_enabled, disc_sense, stars_enabled, star_count;
    int32_t sun_enabled, hs_enabled, hs_sense, ns_enabled;
    int32_t ns_caps, ns_two_caps, jet_enabled, threads;
    /* Jet model: 0 = the reference's Jet (src/jet.h), 1 = a conical jet that
     * the reference does not have: two cones about the spin axis of half
     * opening jet_cone_angle [rad] from jet_cone_r0 to jet_cone_r1 [GM/c^2],
     * radial outflow at Lorentz factor jet_gamma_max in the ZAMO frame,
     * comoving emissivity jet_j_nu0 (r0 / r)^2 nu^-jet_alpha with smooth
     * edges; integrated along the ray like the reference's jets. */
    int32_t jet_model, jet_reserved;
    double jet_cone_angle, jet_cone_r0, jet_cone_r1;
} bhx_scene;

/* RenderStats (src/render.h:55-69) plus the counters the reference lacks. */
typedef struct bhx_stats {
    int64_t rays, steps, captured, disc_hits, sun_hits, star_hits, jet_rays;
    int64_t escaped, exhausted;
    double max_norm_error, max_carter_drift, seconds, total_flux;
    int64_t attempts;         /* DP5 attempts, accepted + rejected           */
    int64_t rhs_evals;        /* geodesic_rhs evaluations on the GPU          */
    double trace_ms;          /* device time of the trace kernel (HIP events) */
    double shade_ms;          /* device time of the shading kernel            */
    double order_ms;          /* device time of the cost-ordered queue's pilot,
                                 sort and scan (0 with the centre-first order) */
} bhx_stats;

/* Per-ray outputs of the tracer (device pointers; any may be NULL). */
typedef struct bhx_rays {
    uint8_t *cls;             /* n_rays, enum bhx_class                       */
    double *y;                /* 8 x n_rays SoA: state Hit::y at termination  */
    double *jet;              /* n_rays: Hit::jet                             */
    int32_t *steps;           /* n_rays: accepted DP5 steps                   */
    int32_t *attempts;        /* n_rays: DP5 attempts incl. rejected          */
    double *norm_err;         /* n_rays: max |g^ab p_a p_b| / |terms| on ray  */
    double *carter;           /* n_rays: max |Q - Q0| / (|Q0| + 1e-12)        */
} bhx_rays;

/* Per-ray shading outputs (device pointers; any may be NULL). */
typedef struct bhx_shaded {
    double *xyz;              /* n_pixels x 3, interleaved like img::Image    */
    double *ray_xyz;          /* 3 x n_rays SoA: colour of each sample        */
    double *g;                /* n_rays: redshift factor (disc / NS), else 0  */
    double *dir;              /* 3 x n_rays SoA: escape direction (sky), else 0 */
} bhx_shaded;

typedef struct bhx_context bhx_context;

/* ---- context ---------------------------------------------------------- */

/* Opens device `device` and allocates the context's tables and counters. */
int bhx_create(int device, bhx_context **out);
int bhx_destroy(bhx_context *ctx);
const char *bhx_last_error(const bhx_context *ctx);
int bhx_abi_version(void);

/*
 * Star catalogue.  By default a scene with stars_enabled uses the reference's
 * generator, StarField(star_count, seed) (src/scene.h:62-107), built on the
 * host and binned on 360 x 720 cells (src/scene.h:173-181).  This call loads
 * an arbitrary catalogue instead (host arrays, n stars, unit directions,
 * colour temperature in K, irradiance in W m^-2); n = 0 returns to the
 * generator.  The context keeps its own copy.
 */
int bhx_set_catalogue(bhx_context *ctx, int32_t n, const double *x, const double *y,
                      const double *z, const double *temperature, const double *irradiance);

/* ---- frame API --------------------------------------------------------- */

/*
 * Drop-in for bh::render (src/render.h:72; body src/render.cpp:396-497):
 * traces and shades every sample of the scene on the GPU, averages the
 * samples of each pixel, and writes width*height*3 doubles of absolute CIE
 * XYZ radiance (row-major, like img::Image) to the HOST buffer `xyz`.
 * `stats` is accumulated like RenderStats (counters +=, errors max,
 * seconds assigned).  Synchronous.
 */
int bhx_render(bhx_context *ctx, const bhx_scene *scene, double *xyz, bhx_stats *stats);

/*
 * Tracer stage alone: trace() (src/render.cpp:69-260) for every sample of the
 * scene, one persistent launch over an atomic work queue.  Output arrays hold
 * n_rays = width*height*sqrt_spp^2 entries indexed by ray.  `stats` may be
 * NULL; when given, the call synchronises the stream and accumulates rays,
 * steps, attempts, RHS evaluations and the invariant maxima (the per-class
 * counters are filled by bhx_shade_dev).  The cost-ordered queue of a large
 * frame is built on a stream the context owns, joined to `stream` by events
 * (on `stream` itself while it is being captured): frames queued back to back
 * build their order while the previous frame traces and shades.
 */
int bhx_trace_dev(bhx_context *ctx, const bhx_scene *scene, const bhx_rays *rays,
                  bhx_stats *stats, void *stream);

/*
 * The tracer stage without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same per-ray outputs as bhx_trace_dev, bit for bit, and the same stats
 * counters (trace_ms and order_ms are left 0), but no context: every scratch
 * buffer (initial ray states, queue map, tile order and sort workspace,
 * counters, event parameters) lives in one static device array of about
 * 140 MiB that each GPU the library runs on holds until the process ends.
 * The queue is the centre-first tile order, built on the device.
 *   Limit: width * height * sqrt_spp^2 <= 1048576 samples and at most 32768
 *   tiles of 8 x 8 pixels; larger frames return BHX_E_TOO_LARGE.
 * Calls on one GPU share the array: calls that overlap on one GPU race, so
 * order them on one stream (or synchronise).  `device` is made current.
 * Failures return a BHX_E_* code; bhx_last_error(NULL) gives the message of
 * the calling thread's last failure.
 */
int bhx_trace_noalloc_dev(int32_t device, const bhx_scene *scene, const bhx_rays *rays, bhx_stats *stats,
                          void *stream);

/*
 * Shading stage: shade() (src/render.cpp:382-393) of every ray of the frame
 * from the tracer's outputs (cls, y, jet over ALL rays of the scene, ray_begin
 * 0), then the per-pixel average of src/render.cpp:460-462.
 */
int bhx_shade_dev(bhx_context *ctx, const bhx_scene *scene, const bhx_rays *rays,
                  const bhx_shaded *out, bhx_stats *stats, void *stream);

/* img::auto_exposure (src/image.cpp:22-45) on a device XYZ frame; the
 * result is written to the host double *exposure.  Synchronous. */
int bhx_auto_exposure_dev(bhx_context *ctx, const double *xyz, int64_t n_pixels, double key,
                          double *exposure, void *stream);

/* img::add_glare (src/image.cpp:47-106) on a device XYZ image (row-major,
 * width x height x 3 doubles), in place: the image plus a sigma = radius_px
 * and a sigma = 3 radius_px Gaussian of itself, weights 1 - strength,
 * 0.7 strength, 0.3 strength.  Bit-identical to the reference's CPU result.
 * No-op when strength <= 0 or radius_px <= 0; BHX_E_TOO_LARGE above a
 * radius of 113 px.  Stream-ordered (one host synchronisation of `stream`
 * before its taps are uploaded). */
int bhx_add_glare_dev(bhx_context *ctx, double *xyz, int32_t width, int32_t height, double strength,
                      double radius_px, void *stream);
/* img::develop (src/image.cpp:119-191): device XYZ frame -> device sRGB8. */
int bhx_develop_dev(bhx_context *ctx, const double *xyz, int64_t n_pixels, double exposure,
                    int32_t tonemap, double log_decades, uint8_t *rgb, void *stream);

/* img::auto_exposure then img::develop (src/image.cpp:22-45, 119-191) without
 * the host round trip: no counterpart in the reference, whose two calls
 * return the exposure to the caller in between.  The exposure stays on the
 * device - the same bits as bhx_auto_exposure_dev gives, the same sRGB8 as
 * bhx_develop_dev with it - so a frame's trace, shading and display pipeline
 * are queued on `stream` with no synchronisation.  exposure_dev, when not
 * NULL, is a device double that receives the exposure (n_pixels > 0).
 * Stream-ordered, asynchronous; calls that overlap on one context share its
 * selection buffers and race. */
int bhx_develop_auto_dev(bhx_context *ctx, const double *xyz, int64_t n_pixels, double key, int32_t tonemap,
                         double log_decades, uint8_t *rgb, double *exposure_dev, void *stream);

/* spec::blackbody_xyz (src/spectrum.cpp:123-134) of n temperatures, device
 * arrays, through the same 4096-entry table the shader uses. */
int bhx_blackbody_xyz_dev(bhx_context *ctx, const double *T, int64_t n, double *xyz,
                          void *stream);

/* ---- tuning ------------------------------------------------------------ */

/* Launch shape of the persistent tracer.  0 in a field keeps the default. */
typedef struct bhx_launch {
    int32_t block;            /* threads per workgroup (multiple of 64)       */
    int32_t waves_per_simd;   /* resident waves per SIMD the grid is sized for */
    int32_t chunk;            /* rays a wave takes from the queue at once     */
    int32_t refill;           /* refill when at least this many lanes idle    */
    int32_t queue_order;      /* single frames: BHX_ORDER_CENTRE or
                                 BHX_ORDER_COST (the default; frames of at
                                 least 4 rays per resident lane without a
                                 jet); batches ignore it                      */
} bhx_launch;

/* Work-queue order of a single frame.  Both trace the same rays to the same
 * bits; they differ only in when each 8 x 8 tile starts.
 *   CENTRE  tiles by distance from the image centre;
 *   COST    tiles longest-first, ranked by a pilot trace of one coarse ray
 *           per tile (rtol 1e-4, 16x the step-size cap, at most 32 steps)
 *           and the largest pilot cost in the tile's 3 x 3 neighbourhood, so
 *           the frame ends on its shortest rays. */
enum { BHX_ORDER_CENTRE = 1, BHX_ORDER_COST = 2 };

int bhx_set_launch(bhx_context *ctx, const bhx_launch *launch);
int bhx_get_launch(const bhx_context *ctx, bhx_launch *launch);

/* ---- batches ------------------------------------------------------------
 *
 * Many scenes traced as one job: what a caller of the reference does with a
 * loop of bh::render calls (src/render.h:72; the CLI's animation loop
 * src/main.cpp:931-997 is one), done as ray generation per scene, ONE
 * persistent tracer launch over every scene's rays (each lane reads its own
 * scene's constants), and shading per scene, captured once into a hipGraph.
 * Per-ray and per-pixel results are those of bhx_trace_dev / bhx_shade_dev
 * on each scene alone.  Limits: at most 128 scenes and 2e9 samples per batch;
 * scenes with stars must share star_count and seed (one catalogue).
 */
typedef struct bhx_batch bhx_batch;

/* Copies the scenes and allocates every device buffer of the batch. */
int bhx_batch_create(bhx_context *ctx, const bhx_scene *scenes, int32_t n, bhx_batch **out);
int bhx_batch_destroy(bhx_batch *batch);

/* Enqueue the whole batch on `stream`.  use_graph != 0: the first call
 * captures the sequence into a hipGraph (on a private stream) and every call
 * launches that graph; use_graph == 0: plain stream launches. */
int bhx_batch_run(bhx_batch *batch, int32_t use_graph, void *stream);

/* Device views of scene i after a run: the batch's per-ray arrays (ray j of
 * the scene is index ray_base + j; component c of the hit state is at
 * y[c * stride + ray_base + j]) and the scene's W*H*3 XYZ pixels. */
int bhx_batch_scene(const bhx_batch *batch, int32_t i, bhx_rays *rays, int64_t *ray_base,
                    int64_t *stride, double **xyz);

/* Copy scene i's results to HOST buffers (any may be NULL; synchronises
 * `stream`): class [n], hit state [8][n] component-major, jet [n], accepted
 * steps [n], attempts [n] (n = the scene's samples) and XYZ [W*H*3]. */
int bhx_batch_copy_scene(bhx_batch *batch, int32_t i, uint8_t *cls, double *y, double *jet,
                         int32_t *steps, int32_t *attempts, double *xyz, void *stream);

/* Per-scene RenderStats of the last run (n entries; synchronises `stream`);
 * batch_ms = device time of the whole batch. */
int bhx_batch_stats(bhx_batch *batch, bhx_stats *stats, double *batch_ms, void *stream);

/* Launch shape of the batch's tracer (0 in a field keeps the value). */
int bhx_batch_set_launch(bhx_batch *batch, const bhx_launch *launch);
int bhx_batch_get_launch(const bhx_batch *batch, bhx_launch *launch);

/* Autotuner: coordinate search over block size, occupancy, queue chunk and
 * refill threshold, each candidate timed over `reps` graph replays of this
 * batch; leaves the fastest shape set and returns it with its time. */
int bhx_batch_autotune(bhx_batch *batch, int32_t reps, bhx_launch *best, double *best_ms,
                       void *stream);

/* ---- differentiable render ----------------------------------------------
 *
 * The intensity image I per sample [W m^-2 sr^-1] - the bolometric disc
 * intensity sigma (g T)^4 / pi of a disc hit (the disc model of
 * src/disc.h:94-121, src/render.cpp:280-300) plus, with the conical jet
 * (jet_model 1), its band-integrated intensity kappa J, J the specific
 * intensity at nu0 integrated along the ray - and its discrete adjoint: given
 * dL/dI per sample, the gradient of L with respect to (spin, cam_theta [rad],
 * eddington, cam_r).  The forward records a checkpoint of the integrator
 * state every K accepted steps (64 bytes each) into a device pool of
 * chunks (16 records for K < 16, 8 for K < 64, else 4): one per ray to
 * start with and more as a ray needs them, so the pool holds about sum over
 * rays of (steps / K + 1) x 64 bytes rounded up to chunks, whatever
 * max_steps is; a
 * forward that finds the pool short still completes its image, and the next
 * backward grows the pool to the forward's count and records the same
 * forward again before it sweeps; the backward replays each K-step segment from its checkpoint with
 * the same arithmetic and sweeps it in reverse with the step sizes frozen,
 * and differentiates the disc crossing time by the implicit-function
 * theorem.  K = 1 records every accepted step instead (48 bytes: its start
 * (r, theta, p_r, p_theta), step size and jet flag; chunks of 128 records),
 * about steps x 48 bytes per ray, and the backward reads the records without
 * replaying; the gradient is the same, bit for bit, at every K.  No counterpart in the reference, which has no derivatives; the
 * forward image is the reference's (same tracer).  Scenes with a neutron
 * star, the reference's jets (jet_model 0), the Sun or a hot spot are
 * rejected.
 */
typedef struct bhx_adjoint bhx_adjoint;

int bhx_adjoint_create(bhx_context *ctx, const bhx_scene *scene, int32_t K, bhx_adjoint **out);
int bhx_adjoint_destroy(bhx_adjoint *adj);
/* New parameters for a frame of the same size (synchronises). */
int bhx_adjoint_set_scene(bhx_adjoint *adj, const bhx_scene *scene);
/* Recording forward: I (device, one double per sample) and the checkpoints,
 * then the backward's ray order (lit rays, I > 0, by step count, a device
 * radix sort).  Stream-ordered, no host synchronisation. */
int bhx_adjoint_forward(bhx_adjoint *adj, double *I, void *stream);
/* w (device, one double per sample) = dL/dI; grad (host, 4 doubles) =
 * dL/d(spin, cam_theta, eddington, cam_r) at the last forward's parameters.
 * Synchronous.  BHX_E_TOO_LARGE when the checkpoints would need more than
 * 128 GiB (raise K). */
int bhx_adjoint_backward(bhx_adjoint *adj, const double *w, double *grad, void *stream);
/* Device views of the last forward: class, hit state (8 x n SoA), steps. */
int bhx_adjoint_rays(const bhx_adjoint *adj, bhx_rays *rays);
/* The same copied to HOST buffers (any may be NULL; synchronises `stream`). */
int bhx_adjoint_copy_rays(bhx_adjoint *adj, uint8_t *cls, double *y, int32_t *steps, void *stream);
/* Autotuner of the checkpoint interval: for each of the n_K candidates Ks
 * (at most 16), one untimed and `reps` timed gradient passes (recording
 * forward + backward with dL/dI = w, device) at the adjoint's current scene;
 * leaves the fastest K set, then times the recording forward at each number
 * of waves per SIMD that fits and keeps the fastest too, and returns K and
 * the time per pass [ms] of the pair.  The gradient does not depend on
 * either (the replay reproduces every checkpoint).
 * A candidate whose records would exceed the pool's 128 GiB cap is skipped
 * (BHX_E_TOO_LARGE when none fits).  Synchronous; the next backward needs a
 * new forward. */
int bhx_adjoint_autotune(bhx_adjoint *adj, const double *w, const int32_t *Ks, int32_t n_K, int32_t reps,
                         int32_t *best_K, double *best_ms, void *stream);
/* The checkpoint interval in use. */
int bhx_adjoint_get_interval(const bhx_adjoint *adj, int32_t *K);
/* The recording forward's queue: a wave takes q_chunk consecutive slots at a
 * time and refills its idle lanes once `refill` of them wait (a ray's set-up
 * then runs on many lanes at once); 0 keeps a value.  Defaults: 32 and 1
 * (a single frame), 32 and 4 (a batch).
 * Results do not depend on them. */
int bhx_adjoint_set_launch(bhx_adjoint *adj, int32_t q_chunk, int32_t refill);
/* Waves per SIMD the recording forward runs (1 up to what fits, 0: what
 * fits, the default).  A frame of few rays per lane lasts as long as its
 * longest ray, which fewer waves per SIMD finish sooner; bhx_adjoint_autotune
 * picks it after K.  Results do not depend on it. */
int bhx_adjoint_set_forward_waves(bhx_adjoint *adj, int32_t waves);
int bhx_adjoint_get_forward_waves(const bhx_adjoint *adj, int32_t *waves);
/* A batch of n frames of one size and one physics (spin, disc, jet,
 * integrator, cam_r) that differ in their camera only (cam_theta, cam_phi,
 * fov, roll, jitter seed): one persistent forward over all n x samples rays,
 * so the GPU stays full across frames, and one backward that gives each
 * frame its own gradient.  I and w hold the frames back to back (frame f's
 * samples at f x samples_per_frame, each in bhx.h's index order); grad (host,
 * 4 x n doubles) gets frame f's dL_f/d(spin, cam_theta, eddington, cam_r) at
 * 4 f, L_f the part of L that frame f's samples carry: sum over f of
 * frame f's own gradient, not a shared parameter's total.  Each frame's
 * gradient is its rays' contributions summed in a fixed order
 * (deterministic); it equals the single-frame backward's to rounding (the
 * sums run in another order).  Hit states are not kept (bhx_adjoint_rays
 * gives y = NULL); bhx_adjoint_set_scene and bhx_adjoint_backward refuse a
 * batch.  At most 65535 frames and 2e9 samples. */
int bhx_adjoint_create_batch(bhx_context *ctx, const bhx_scene *scenes, int32_t n, int32_t K,
                             bhx_adjoint **out);
int bhx_adjoint_backward_batch(bhx_adjoint *adj, const double *w, double *grad, void *stream);
/* Checkpoint pool size in bytes, lit rays (I > 0) of the last forward, replayed segments
 * that did not reproduce their checkpoint (0 expected), kernel times. */
int bhx_adjoint_info(bhx_adjoint *adj, int64_t *ck_bytes, int64_t *lit_rays, int64_t *mismatches,
                     double *fwd_ms, double *bwd_ms);

/* ---- scene preparation (host only, no context, no GPU) ------------------ */

/* Normalise the jets of a scene from its spin and accretion rate: the CLI's
 * setup_jet (src/main.cpp:399-423) with Jet::configure (src/jet.h:261-311) -
 * the radiated fraction of the Blandford-Znajek power spread over the jets'
 * emissivity-weighted volume, here integrated by Gauss-Legendre quadrature
 * (the reference's midpoint sums give a j_nu0 0.7-1.7e-8 relative apart).
 * phi_flux is the dimensionless horizon flux (Jet::phi_flux, default 50) and
 * eps_rad the radiated fraction of the jet power (Jet::eps_rad, default 0.02),
 * src/jet.h:140-141, the CLI's --jet-flux and --jet-efficiency.
 * Reads M_kg, spin, eddington, disc_sense, jet_z_top, jet_z_sat and the jet
 * shape fields; writes jet_z_sat (capped at 0.7 z_top), jet_a, jet_Omega_F,
 * jet_j_nu0, jet_r_bound.  Clears jet_enabled when the scene has no disc, is a
 * neutron-star scene, or the Blandford-Znajek power is zero (a = 0). */
int bhx_setup_jet(bhx_scene *scene, double phi_flux, double eps_rad);

/* The same normalisation for the conical jet (jet_model = 1): the radiated
 * fraction eps_rad of the Blandford-Znajek power spread over both cones'
 * emissivity profile; writes jet_a, jet_Omega_F, jet_j_nu0 and jet_r_bound
 * (1.05 jet_cone_r1).  No counterpart in the reference. */
int bhx_setup_conical_jet(bhx_scene *scene, double phi_flux, double eps_rad);

/* Stable-branch star of gravitational mass M_kg for the polytrope
 * P = K_si rho^Gamma (SI units, rho the mass-energy density): what
 * ns::star_of_mass (src/tov.h:205-218) over ns::solve_tov (src/tov.h:65-177)
 * gives, central densities searched in the same range [2e17, 5e18] kg m^-3
 * and the surface at the same P = 1e-10 P_c.  The structure is integrated in
 * the enthalpy form (converged to ~2e-10 in R; the reference's radius is
 * 4-9e-8 smaller, its RK4 truncation error) and the stable branch is ended at
 * the true mass maximum rather than the best of 40 sampled densities.
 * Returns BHX_E_INVALID when no star of the range has that mass (heavier
 * than the maximum, or lighter than the star at 2e17 kg m^-3). */
int bhx_polytrope_star(double K_si, double Gamma, double M_kg, double *R_m, double *M_out);

/* A synthetic star catalogue (no counterpart in the reference, whose
 * StarField puts 45% of its stars near a galactic plane and goes to m = 12):
 * `count` stars with directions uniform on the sphere, apparent magnitude m
 * by inverse CDF of N(<m) ~ 10^(0.45 m) on [-1.5, 8] (the reference's count
 * law, src/scene.h:86-89), B-V ~ N(0.6, 0.4) mapped to colour temperature by
 * spec::bv_to_temperature (src/spectrum.cpp:158-162), irradiance
 * F0 10^(-0.4 m) with the reference's F0 (src/scene.h:100-101); one splitmix64
 * stream seeded with `seed`.  Fills count entries of each host array, ready
 * for bhx_set_catalogue. */
int bhx_uniform_sky(int32_t count, uint64_t seed, double *x, double *y, double *z, double *temperature,
                    double *irradiance);

/* Make a scene a neutron-star scene for a solved star (R_m metres, M_kg kg):
 * the mass and spin rules of the CLI's setup_neutron_star (src/main.cpp:
 * 503-512) and NeutronStar::set_from / set_spin (src/neutronstar.h:69-79);
 * ns_spin_hz must already be set.  Camera framing stays with the caller. */
int bhx_setup_neutron_star(bhx_scene *scene, double R_m, double M_kg);

/* ---- PNG / APNG ----------------------------------------------------------
 *
 * img::write_png(path, rgb, w, h) and img::write_apng(path, frames, w, h, fps,
 * loops) (src/image.h:45-50, src/image.cpp:193-778) as byte streams: the same
 * bytes as the reference's writer - per-scanline filter of minimum residual,
 * one dynamic-Huffman DEFLATE block from the greedy hash-chain LZ77 parse
 * (stored block when that is smaller), zlib, PNG / APNG 1.0 chunks with the
 * changed rectangle of each later frame.  The scanline filters and the match
 * search run on the device; the parse, the Huffman codes and the checksums
 * on the host.  Images are w x h x 3 bytes, row-major RGB.  The calls return
 * when the stream is complete (they synchronise `stream`).  The stream is
 * allocated by the library: release it with bhx_bytes_free.  Unlike the
 * reference (which writes a file and returns false on a size mismatch) the
 * sizes are arguments; invalid ones give BHX_E_INVALID.
 */
typedef struct bhx_bytes {
    uint8_t *data;
    size_t size;
} bhx_bytes;

/* rgb in host memory */
int bhx_png_encode(bhx_context *ctx, const uint8_t *rgb, int32_t width, int32_t height, bhx_bytes *out);
/* rgb in device memory (e.g. bhx_develop_dev's output), ordered on stream */
int bhx_png_encode_dev(bhx_context *ctx, const uint8_t *rgb, int32_t width, int32_t height, void *stream,
                       bhx_bytes *out);
/* n_frames images in host memory; loops = 0 plays forever */
int bhx_apng_encode(bhx_context *ctx, const uint8_t *const *frames, int32_t n_frames, int32_t width,
                    int32_t height, int32_t fps, int32_t loops, bhx_bytes *out);
/* n_frames device pointers (the array itself in host memory), ordered on stream */
int bhx_apng_encode_dev(bhx_context *ctx, const uint8_t *const *frames, int32_t n_frames, int32_t width,
                        int32_t height, int32_t fps, int32_t loops, void *stream, bhx_bytes *out);
void bhx_bytes_free(bhx_bytes *bytes);

#ifdef __cplusplus
}
#endif

#endif /* BHX_H */
