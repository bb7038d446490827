// params.h - launch arguments shared by the host API and the HIP kernels.
#pragma once

#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define BHX_HD __host__ __device__
#else
#define BHX_HD
#endif

namespace bhx {

// Jet (after Jet::configure, src/jet.h:96-145) and Sun (src/scene.h:232-235)
// parameters of the tracer's event branches, read from device memory.
struct EventParams {
    double z_base, z_top, R_base, k, gamma_max, z_sat, alpha, u_peak, u_wid, spine, spine_wid;
    double zeta, Omega_F, j_nu0, r_bound, gamma_zb;   // gamma_zb = Jet::gamma_at(z_base)
    double sun_c[3], sun_R;
    // conical jet (model 1): half opening, radial extent, outflow speed and
    // Lorentz factor; cone_cos = cos(cone_angle), cone_itaper =
    // 1 / (cos(0.8 cone_angle) - cos(cone_angle))
    double cone_angle, cone_r0, cone_r1, cone_beta, cone_gamma, cone_cos, cone_itaper;
    int32_t model, pad;
};

// Sampling of the conical jet along a step: kConeSub sub-intervals of
// 4-point Gauss-Legendre each, on the steps whose dense output can reach the
// cone.  With theta(s) = c1 + s c2 + s(1-s) c3 + s^2(1-s) c4 + s^2(1-s)^2 c5,
// |theta(s) - theta(0)| <= cone_reach = |c2| + |c3|/4 + 4|c4|/27 + |c5|/16 on
// [0, 1]; a step whose start lies further than cone_angle + cone_reach from
// the axis stays outside the cone, where the emission is zero.
constexpr int kConeSub = 8;
BHX_HD inline bool cone_reach(double th0, double c2, double c3, double c4, double c5,
                                           double cone_angle) {
    const double reach = fabs(c2) + 0.25 * fabs(c3) + (4.0 / 27.0) * fabs(c4) + 0.0625 * fabs(c5);
    const double psi0 = th0 < 3.14159265358979323846 - th0 ? th0 : 3.14159265358979323846 - th0;
    return psi0 - reach < cone_angle;
}

// One scene of a batched launch: the tracer's per-scene constants (the
// uniform fields of TraceArgs) plus its event parameters.  The batched tracer
// copies the table into LDS and each lane reads its own scene's row.
struct SceneK {
    double a, a2, r_h, r_stop, rtol, atol, r_escape, h0, r_in, r_out, ns_R;
    int32_t has_disc, has_ns, has_jet, has_sun, max_steps, pad;
    EventParams ev;
};
static_assert(sizeof(SceneK) % sizeof(double) == 0, "SceneK is copied as doubles");
constexpr int kMaxBatchScenes = 128;

// Pilot of the cost-ordered queue (one coarse ray per 8 x 8 tile): step-size
// cap kPilotHmax times the reference's hmax (render.cpp:130-131), at least
// these tolerances, at most kPilotSteps accepted steps.  Its attempt counts
// only rank the tiles.
constexpr double kPilotHmax = 24.0;
constexpr double kPilotRtol = 1e-4, kPilotAtol = 1e-6;
constexpr int kPilotSteps = 24;

// Ray generation: camera, jitter and the initial photon of every sample
// (src/render.cpp:437-452, src/scene.h:356-384, src/geodesic.h:215-231),
// written in work-queue order so the tracer's refills are coalesced loads.
struct RaygenArgs {
    double a, a2;
    double cam_r, cam_th, cam_ph;
    double e00, e03, e11, e22, e33;           // ZAMO tetrad, non-zero entries
    double gl_tt, gl_tp, gl_rr, gl_hh, gl_pp; // g_{mu nu} at the camera
    double cam_cs2, cam_sn2;                  // cos^2, clamped sin^2 of theta_cam (Carter Q0)
    double tanf, aspect, inv_S;
    double c_roll, s_roll, c_pitch, s_pitch, c_yaw, s_yaw;
    int32_t has_roll;
    int32_t W, H, S;
    uint64_t seed;
    int32_t tiles_x, n_tiles;
    const int32_t* tile_order;                // queue tile -> image tile (8 x 8 pixels)
    const int32_t* tile_start;                // queue tile -> first queue slot
    int32_t n_rays;                           // row stride of init (all rays of the launch)
    double* init;                             // kInitRows x n_rays, queue order
    int32_t* q2ray;                           // queue slot -> ray index
    // batched launches: this scene's first ray index, its id, and the
    // queue-slot -> scene map (NULL for a single scene)
    int32_t ray_base, scene_id;
    uint8_t* q2scene;
};

// Rows of the initial-state array.
enum InitRow { kI_T = 0, kI_R, kI_TH, kI_PH, kI_PR, kI_PTH, kI_PT, kI_PPH,
               kI_K1 = 8, kI_C0 = 14, kI_Q0 = 15, kInitRows = 16 };

// The tracer's loop: only what the integration and the events need, so
// every uniform value stays in scalar registers.
struct TraceArgs {
    // metric and integrator (src/render.cpp:71-86, src/render.h:45-48)
    double a, a2, r_h, r_stop, rtol, atol, r_escape, h0;
    // events (src/render.cpp:158-199)
    double r_in, r_out, ns_R;
    int32_t has_disc, has_ns, has_jet, has_sun, max_steps, refill, chunk;
    const EventParams* ev;          // jet and Sun parameters (device memory)
    int32_t n_rays;
    const double* init;             // kInitRows x n_rays, queue order
    const int32_t* q2ray;
    uint32_t* counter;              // next queue slot
    // batched launch: per-scene constants (n_scenes rows) and the queue-slot
    // -> scene map; the uniform fields above are then unused
    const SceneK* scenes;
    const uint8_t* q2scene;
    int32_t n_scenes;
    // outputs (SoA over ray index, see bhx.h)
    uint8_t* cls;
    double* y;
    double* jet;
    int32_t* steps;
    int32_t* attempts;
    double* norm_err;
    double* carter;
    // frame totals: [0] accepted steps, [1] attempts; [2..3] max norm / Carter (as bits)
    unsigned long long* totals;
};

// Device state of the exposure's radix selection between its passes: the
// count of positive finite luminances, and per order statistic its remaining
// rank and the key bits selected so far.
struct ExposureState {
    unsigned long long n_pos;
    unsigned long long k[2];
    unsigned long long prefix[2];
};

// Disc / NS / sky shading constants (src/render.cpp:265-393, src/disc.h,
// src/scene.h:111-223, src/neutronstar.h:90-125).
struct ShadeArgs {
    double a, a2, psf_sigma, observer_time;
    int32_t W, H, S, n_pix;
    int64_t n_rays;                    // row stride of the SoA ray arrays
    int64_t ray_base;                  // first ray of this scene in those arrays
    // Novikov-Thorne disc
    int32_t has_disc, disc_sg;
    double disc_a, r_in, r_out, flux_scale;   // flux_scale = mdot c^2 / r_g^2
    double pt_x0, pt_c[3], pt_x[3];           // Page-Thorne constants (disc.h:46-64)
    // hot spot
    int32_t hs_enabled;
    double hs_r, hs_phi0, hs_sigma, hs_contrast, hs_omega;
    // neutron star
    int32_t ns_enabled, ns_caps, ns_two_caps;
    double ns_R, ns_T_eff, ns_Omega, ns_cap_T, ns_cap_radius, ns_cap_phase0;
    double ns_m[3];                            // unused when caps follow the phase
    double ns_sin_tilt, ns_cos_tilt;
    // sky
    int32_t stars_enabled, n_cells;
    const uint32_t* cell_start;               // n_cells + 1 offsets into star rows
    const double* star;                       // 6 doubles per star: x y z irr cx cz
    double mw_L, mw_cx, mw_cz;                // Milky Way amplitude and colour ratios
    // Sun (render.cpp:301-325, scene.h:239-276)
    int32_t has_sun;
    double sun_c[3], sun_R, sun_T_eff, sun_self_z;
    // jet colour (spec::powerlaw_xyz of the jet's alpha)
    double jet_rgb[3];
    // spectrum table: 3 x 4096 natural logs (spectrum.cpp:97-114)
    const double* bb;
    // inputs
    const uint8_t* cls;
    const double* y;
    const double* jet;
    // outputs
    double* xyz;
    double* ray_xyz;
    double* g;
    double* dir;
    // per-block partials: class counts [6] as u64 and flux sums
    unsigned long long* class_counts;   // 7 counters: 6 classes + jet rays
    double* flux_partial;               // one per block
    // optional per-ray trace counters to reduce (batched launches, whose
    // tracer cannot total per scene): accepted steps, attempts, norm and
    // Carter maxima -> trace_totals[4] (sums, and maxima as bits)
    const int32_t* steps_in;
    const int32_t* attempts_in;
    const double* norm_in;
    const double* carter_in;
    unsigned long long* trace_totals;
};

}  // namespace bhx
