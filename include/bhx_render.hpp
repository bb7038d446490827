/* bhx_render.hpp - the reference's C++ render entry point over the C ABI.
 *
 * Header-only C++17.  It maps a RenderConfig-shaped object (src/render.h:15-53
 * with Camera src/scene.h:343-351, SunBody src/scene.h:232-236, HotSpot
 * src/scene.h:306-313, NeutronStar src/neutronstar.h:43-67, Jet
 * src/jet.h:97-145) field by field onto bhx_scene, runs bhx_render on the GPU,
 * and merges the result into a RenderStats-shaped object the way render()
 * does (src/render.cpp:481-495: counters added, error maxima kept, seconds
 * assigned).  It is templated on those types, so the reference's own headers
 * drive it unchanged; a drop-in replacement of src/render.cpp is then
 *
 *     #include "render.h"
 *     #include "bhx_render.hpp"
 *     namespace bh {
 *     img::Image render(const RenderConfig& cfg, RenderStats& stats) {
 *         return bhx::render<img::Image>(cfg, stats);
 *     }
 *     }
 *
 * The same header gives img::write_png / img::write_apng (src/image.h:45-50)
 * on the GPU-assisted writer as bhx::write_png / bhx::write_apng
 * (blackhole_amd/integration/png_hip.cpp routes the reference's to them).
 *
 * Errors: the reference's render() has no failure mode; a HIP failure here
 * (no device, out of memory, a scene outside the documented limits) throws
 * std::runtime_error with the library's message.
 *
 * Threading: one bhx_context per host thread, created on first use on the
 * device named by the BHX_DEVICE environment variable (default 0) and kept
 * until the thread ends.  cfg.threads and cfg.quiet are accepted and have no
 * effect (there are no host workers and no progress output).
 */
#ifndef BHX_RENDER_HPP
#define BHX_RENDER_HPP

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "bhx.h"

namespace bhx {

template <class Cfg>
bhx_scene scene_of(const Cfg& c) {
    bhx_scene s;
    std::memset(&s, 0, sizeof s);
    s.M_kg = c.M_kg;
    s.spin = c.spin;
    s.cam_r = c.cam.r;
    s.cam_theta = c.cam.theta;
    s.cam_phi = c.cam.phi;
    s.cam_fov = c.cam.fov;
    s.cam_yaw = c.cam.yaw;
    s.cam_pitch = c.cam.pitch;
    s.cam_roll = c.cam.roll;
    s.width = c.width;
    s.height = c.height;
    s.sqrt_spp = c.sqrt_spp;
    s.disc_enabled = c.disc_enabled ? 1 : 0;
    s.disc_r_out = c.disc_r_out;
    s.eddington = c.eddington;
    s.disc_sense = c.disc_sense;
    s.stars_enabled = c.stars_enabled ? 1 : 0;
    s.star_count = c.star_count;
    s.seed = static_cast<uint64_t>(c.seed);
    s.sun_enabled = c.sun_enabled ? 1 : 0;
    for (int i = 0; i < 3; ++i) s.sun_centre[i] = c.sun.centre[i];
    s.sun_radius = c.sun.radius;
    s.sun_T_eff = c.sun.T_eff;
    s.hs_enabled = c.hotspot.enabled ? 1 : 0;
    s.hs_r = c.hotspot.r;
    s.hs_phi0 = c.hotspot.phi0;
    s.hs_sigma = c.hotspot.sigma;
    s.hs_contrast = c.hotspot.contrast;
    s.hs_omega = c.hotspot.omega;
    s.hs_sense = c.hotspot.sense;
    s.ns_enabled = c.star.enabled ? 1 : 0;
    s.ns_M_kg = c.star.M_kg;
    s.ns_R_m = c.star.R_m;
    s.ns_R = c.star.R;
    s.ns_T_eff = c.star.T_eff;
    s.ns_spin_hz = c.star.spin_hz;
    s.ns_Omega = c.star.Omega;
    s.ns_caps = c.star.caps ? 1 : 0;
    s.ns_cap_tilt = c.star.cap_tilt;
    s.ns_cap_radius = c.star.cap_radius;
    s.ns_cap_T = c.star.cap_T;
    s.ns_cap_phase0 = c.star.cap_phase0;
    s.ns_two_caps = c.star.two_caps ? 1 : 0;
    s.jet_enabled = c.jet.enabled ? 1 : 0;
    s.jet_z_base = c.jet.z_base;
    s.jet_z_top = c.jet.z_top;
    s.jet_R_base = c.jet.R_base;
    s.jet_k = c.jet.k;
    s.jet_gamma_max = c.jet.gamma_max;
    s.jet_z_sat = c.jet.z_sat;
    s.jet_alpha = c.jet.alpha;
    s.jet_u_peak = c.jet.u_peak;
    s.jet_u_wid = c.jet.u_wid;
    s.jet_spine = c.jet.spine;
    s.jet_spine_wid = c.jet.spine_wid;
    s.jet_zeta = c.jet.zeta;
    s.jet_a = c.jet.a;
    s.jet_Omega_F = c.jet.Omega_F;
    s.jet_j_nu0 = c.jet.j_nu0;
    s.jet_r_bound = c.jet.r_bound;
    s.observer_time = c.observer_time;
    s.rtol = c.rtol;
    s.atol = c.atol;
    s.r_escape = c.r_escape;
    s.max_steps = c.max_steps;
    s.threads = c.threads;
    return s;
}

/* Merge one frame's bhx_stats into a RenderStats-shaped object. */
template <class Stats>
void merge_stats(Stats& st, const bhx_stats& s) {
    st.rays += s.rays;
    st.steps += s.steps;
    st.captured += s.captured;
    st.disc_hits += s.disc_hits;
    st.sun_hits += s.sun_hits;
    st.star_hits += s.star_hits;
    st.jet_rays += s.jet_rays;
    st.total_flux += s.total_flux;
    st.escaped += s.escaped;
    st.exhausted += s.exhausted;
    if (s.max_norm_error > st.max_norm_error) st.max_norm_error = s.max_norm_error;
    if (s.max_carter_drift > st.max_carter_drift) st.max_carter_drift = s.max_carter_drift;
    st.seconds = s.seconds;
}

/* The calling thread's context (created on first use). */
inline bhx_context* thread_context() {
    struct Holder {
        bhx_context* ctx = nullptr;
        ~Holder() {
            if (ctx) bhx_destroy(ctx);
        }
    };
    thread_local Holder h;
    if (!h.ctx) {
        const char* env = std::getenv("BHX_DEVICE");
        const int dev = env ? std::atoi(env) : 0;
        bhx_context* c = nullptr;
        const int rc = bhx_create(dev, &c);
        if (rc != BHX_OK || !c)
            throw std::runtime_error("bhx_create(" + std::to_string(dev) + ") failed: " +
                                     std::to_string(rc));
        h.ctx = c;
    }
    return h.ctx;
}

/* render(cfg, stats) with the reference's semantics; Image needs a
 * (width, height) constructor and a contiguous `pixels` vector of three
 * doubles per pixel, row-major (src/image.h:18-27). */
template <class Image, class Cfg, class Stats>
Image render(const Cfg& cfg, Stats& stats) {
    const auto t0 = std::chrono::steady_clock::now();
    Image im(cfg.width, cfg.height);
    if (cfg.width == 0 || cfg.height == 0) {
        // no samples: the reference's loops do nothing and it returns the empty
        // image with only the time assigned (src/render.cpp:396-497)
        stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return im;
    }
    using Px = typename std::decay<decltype(im.pixels[0])>::type;
    static_assert(sizeof(Px) == 3 * sizeof(double), "pixel must be three doubles");
    const bhx_scene s = scene_of(cfg);
    bhx_stats st;
    std::memset(&st, 0, sizeof st);
    bhx_context* ctx = thread_context();
    const int rc = bhx_render(ctx, &s, reinterpret_cast<double*>(im.pixels.data()), &st);
    if (rc != BHX_OK)
        throw std::runtime_error(std::string("bhx_render failed: ") + bhx_last_error(ctx));
    merge_stats(stats, st);
    return im;
}

namespace detail {
inline bool write_bytes(const std::string& path, bhx_bytes& b) {
    FILE* f = std::fopen(path.c_str(), "wb");
    bool ok = f != nullptr;
    if (f) {
        ok = std::fwrite(b.data, 1, b.size, f) == b.size;
        std::fclose(f);
    }
    bhx_bytes_free(&b);
    return ok;
}
}  // namespace detail

/* img::write_png / img::write_apng (src/image.h:45-50) on the GPU-assisted
 * writer: the same file bytes as the reference's (src/image.cpp:193-778).
 * false on a size mismatch or when the file cannot be written, as the
 * reference; a HIP failure throws std::runtime_error. */
inline bool write_png(const std::string& path, const std::vector<uint8_t>& rgb, int w, int h) {
    if (w <= 0 || h <= 0 || static_cast<size_t>(w) * h * 3 != rgb.size()) return false;
    bhx_context* ctx = thread_context();
    bhx_bytes b{nullptr, 0};
    if (bhx_png_encode(ctx, rgb.data(), w, h, &b) != BHX_OK)
        throw std::runtime_error(std::string("bhx_png_encode failed: ") + bhx_last_error(ctx));
    return detail::write_bytes(path, b);
}

inline bool write_apng(const std::string& path, const std::vector<std::vector<uint8_t>>& frames, int w, int h,
                       int fps, int loops) {
    if (frames.empty() || w <= 0 || h <= 0) return false;
    std::vector<const uint8_t*> ptr;
    for (const auto& f : frames) {
        if (static_cast<size_t>(w) * h * 3 != f.size()) return false;
        ptr.push_back(f.data());
    }
    bhx_context* ctx = thread_context();
    bhx_bytes b{nullptr, 0};
    if (bhx_apng_encode(ctx, ptr.data(), static_cast<int32_t>(ptr.size()), w, h, fps, loops, &b) != BHX_OK)
        throw std::runtime_error(std::string("bhx_apng_encode failed: ") + bhx_last_error(ctx));
    return detail::write_bytes(path, b);
}

}  // namespace bhx

#endif /* BHX_RENDER_HPP */
