// params.h - launch arguments shared by the host API and the HIP kernels.
#pragma once

#include <stdint.h>

namespace bhx {

// Everything the tracer needs about one scene, precomputed on the host
// (camera tetrad, metric at the camera, horizon radii, disc edges).  Passed by
// value as a kernel argument, so every field lands in scalar registers.
struct TraceArgs {
    // metric and integrator (src/render.cpp:71-86, src/render.h:45-48)
    double a, a2, r_h, r_stop, rtol, atol, r_escape, h0;
    double pw_init;                 // pow(1e-4, 0.08): controller memory of a fresh ray
    // camera (src/scene.h:343-385, src/kerr.h:283-293, src/geodesic.h:215-231)
    double cam_r, cam_th, cam_ph;
    double e00, e03, e11, e22, e33;           // ZAMO tetrad, non-zero entries
    double gl_tt, gl_tp, gl_rr, gl_hh, gl_pp; // g_{mu nu} at the camera
    double cam_cs2, cam_sn2;                  // cos^2, clamped sin^2 of theta_cam (Carter Q0)
    double tanf, aspect, inv_S;
    double c_roll, s_roll, c_pitch, s_pitch, c_yaw, s_yaw;
    int32_t has_roll;
    int32_t W, H, S, max_steps;
    uint64_t seed;
    // events (src/render.cpp:158-199)
    int32_t has_disc, has_ns;
    double r_in, r_out, ns_R;
    // work queue: tiles of 8 x 8 pixels in the order given by tile_order
    int32_t tiles_x, n_tiles, refill;
    const int32_t* tile_order;
    uint32_t* tile_counter;
    // outputs (SoA over ray index, see bhx.h)
    int64_t n_rays;
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

// Disc / NS / sky shading constants (src/render.cpp:265-393, src/disc.h,
// src/scene.h:111-223, src/neutronstar.h:90-125).
struct ShadeArgs {
    double a, a2, psf_sigma, observer_time;
    int32_t W, H, S, n_pix;
    int64_t n_rays;
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
};

}  // namespace bhx
