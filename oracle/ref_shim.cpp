// ref_shim.cpp - C ABI over the reference tracer, compiled from its sources.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's own src/spectrum.cpp and src/image.cpp where they lie under
// /root/reference, into oracle/_ref/libbhref.so.  It includes the reference's
// src/render.cpp so that the per-ray functions trace() and shade() (internal
// linkage there) can be called one ray at a time; nothing of the reference is
// copied into this repository.  Used to pin the oracle (tests/test_oracle_ref.py)
// and to write the golden fixtures (tests/golden/make_golden.py).

#include "render.cpp"  // the reference's src/render.cpp, from the include path

#include "../include/bhx.h"
#include "tov.h"

#include <cstring>

namespace {

bh::RenderConfig to_config(const bhx_scene* s) {
    bh::RenderConfig c;
    c.M_kg = s->M_kg;
    c.spin = s->spin;
    c.cam.r = s->cam_r; c.cam.theta = s->cam_theta; c.cam.phi = s->cam_phi;
    c.cam.fov = s->cam_fov; c.cam.yaw = s->cam_yaw; c.cam.pitch = s->cam_pitch;
    c.cam.roll = s->cam_roll;
    c.width = s->width; c.height = s->height; c.sqrt_spp = s->sqrt_spp;
    c.disc_enabled = s->disc_enabled != 0;
    c.disc_r_out = s->disc_r_out; c.eddington = s->eddington; c.disc_sense = s->disc_sense;
    c.stars_enabled = s->stars_enabled != 0; c.star_count = s->star_count; c.seed = s->seed;
    c.sun_enabled = s->sun_enabled != 0;
    for (int i = 0; i < 3; ++i) c.sun.centre[i] = s->sun_centre[i];
    c.sun.radius = s->sun_radius; c.sun.T_eff = s->sun_T_eff;
    c.hotspot.enabled = s->hs_enabled != 0; c.hotspot.r = s->hs_r; c.hotspot.phi0 = s->hs_phi0;
    c.hotspot.sigma = s->hs_sigma; c.hotspot.contrast = s->hs_contrast;
    c.hotspot.omega = s->hs_omega; c.hotspot.sense = s->hs_sense;
    c.star.enabled = s->ns_enabled != 0; c.star.M_kg = s->ns_M_kg; c.star.R_m = s->ns_R_m;
    c.star.R = s->ns_R; c.star.T_eff = s->ns_T_eff; c.star.spin_hz = s->ns_spin_hz;
    c.star.Omega = s->ns_Omega; c.star.caps = s->ns_caps != 0; c.star.cap_tilt = s->ns_cap_tilt;
    c.star.cap_radius = s->ns_cap_radius; c.star.cap_T = s->ns_cap_T;
    c.star.cap_phase0 = s->ns_cap_phase0; c.star.two_caps = s->ns_two_caps != 0;
    c.jet.enabled = s->jet_enabled != 0;
    c.jet.z_base = s->jet_z_base; c.jet.z_top = s->jet_z_top; c.jet.R_base = s->jet_R_base;
    c.jet.k = s->jet_k; c.jet.gamma_max = s->jet_gamma_max; c.jet.z_sat = s->jet_z_sat;
    c.jet.alpha = s->jet_alpha; c.jet.u_peak = s->jet_u_peak; c.jet.u_wid = s->jet_u_wid;
    c.jet.spine = s->jet_spine; c.jet.spine_wid = s->jet_spine_wid; c.jet.zeta = s->jet_zeta;
    c.jet.a = s->jet_a; c.jet.Omega_F = s->jet_Omega_F; c.jet.j_nu0 = s->jet_j_nu0;
    c.jet.r_bound = s->jet_r_bound;
    c.observer_time = s->observer_time;
    c.rtol = s->rtol; c.atol = s->atol; c.r_escape = s->r_escape; c.max_steps = s->max_steps;
    c.threads = s->threads; c.quiet = true;
    return c;
}

void from_stats(const bh::RenderStats& r, bhx_stats* o) {
    o->rays = r.rays; o->steps = r.steps; o->captured = r.captured; o->disc_hits = r.disc_hits;
    o->sun_hits = r.sun_hits; o->star_hits = r.star_hits; o->jet_rays = r.jet_rays;
    o->escaped = r.escaped; o->exhausted = r.exhausted; o->max_norm_error = r.max_norm_error;
    o->max_carter_drift = r.max_carter_drift; o->seconds = r.seconds; o->total_flux = r.total_flux;
}

}  // namespace

extern "C" {

int bhref_render(const bhx_scene* s, double* xyz, bhx_stats* st) {
    const bh::RenderConfig cfg = to_config(s);
    bh::RenderStats rs;
    const img::Image im = bh::render(cfg, rs);
    for (size_t i = 0; i < im.pixels.size(); ++i) {
        xyz[3 * i] = im.pixels[i].x;
        xyz[3 * i + 1] = im.pixels[i].y;
        xyz[3 * i + 2] = im.pixels[i].z;
    }
    if (st) from_stats(rs, st);
    return 0;
}

// Per-ray trace() + shade() with the jitter and camera of the reference's
// render() worker.  y is 8 x n SoA, ray_xyz 3 x n SoA.
int bhref_trace_list(const bhx_scene* s, const int64_t* ids, int64_t n, uint8_t* cls, double* y,
                     double* jet, int32_t* steps, double* g, double* ray_xyz) {
    const bh::RenderConfig cfg = to_config(s);
    std::unique_ptr<bh::NovikovThorneDisc> disc;
    if (cfg.disc_enabled)
        disc = std::make_unique<bh::NovikovThorneDisc>(cfg.M_kg, cfg.spin, cfg.disc_r_out,
                                                       cfg.eddington, cfg.disc_sense);
    std::unique_ptr<bh::StarField> stars;
    if (cfg.stars_enabled) stars = std::make_unique<bh::StarField>(cfg.star_count, cfg.seed);
    const bh::Tetrad tet = bh::zamo_tetrad(bh::geom_at(cfg.spin, cfg.cam.r, cfg.cam.theta));
    const double aspect = static_cast<double>(cfg.width) / cfg.height;
    const double psf_sigma = 0.55 * (cfg.cam.fov / cfg.width);
    const spec::XYZ jet_colour = cfg.jet.enabled ? spec::powerlaw_xyz(cfg.jet.alpha) : spec::XYZ{};
    const int S = std::max(1, cfg.sqrt_spp);
    const double inv_S = 1.0 / S;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t pix = ids[i] / (S * S);
        const int k = static_cast<int>(ids[i] % (S * S));
        const int px = static_cast<int>(pix % cfg.width), py = static_cast<int>(pix / cfg.width);
        bh::Rng rng(cfg.seed ^ (static_cast<uint64_t>(py) << 32) ^
                    (static_cast<uint64_t>(px) * 0x9E3779B97F4A7C15ull));
        double jx = 0.0, jy = 0.0;
        for (int q = 0; q <= k; ++q) {
            jx = ((q % S) + rng.uniform()) * inv_S;
            jy = ((q / S) + rng.uniform()) * inv_S;
        }
        const double u = (2.0 * (px + jx) / cfg.width - 1.0);
        const double v = (1.0 - 2.0 * (py + jy) / cfg.height);
        double dir[3];
        cfg.cam.ray_direction(u, v, aspect, dir);
        bh::RenderStats local;
        const bh::Hit hit = bh::trace(cfg, disc.get(), dir, tet, local);
        const spec::XYZ c = bh::shade(cfg, disc.get(), stars.get(), hit, psf_sigma, jet_colour, local);
        if (cls) cls[i] = static_cast<uint8_t>(hit.what);
        if (y) for (int q = 0; q < 8; ++q) y[q * n + i] = hit.y[q];
        if (jet) jet[i] = hit.jet;
        if (steps) steps[i] = static_cast<int32_t>(local.steps);
        if (g) {
            g[i] = 0.0;
            if (hit.what == bh::Hit::What::Disc && disc) {
                const double r = hit.y[bh::Y_R];
                if (disc->temperature(r) > 0.0)
                    g[i] = bh::redshift_factor(hit.y, disc->four_velocity(r));
            }
        }
        if (ray_xyz) { ray_xyz[i] = c.x; ray_xyz[n + i] = c.y; ray_xyz[2 * n + i] = c.z; }
    }
    return 0;
}

double bhref_auto_exposure(const double* xyz, int64_t npix, double key) {
    img::Image im(static_cast<int>(npix), 1);
    for (int64_t i = 0; i < npix; ++i) im.pixels[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    return img::auto_exposure(im, key);
}

void bhref_develop(const double* xyz, int64_t npix, double exposure, int tm, double log_decades,
                   uint8_t* rgb) {
    img::Image im(static_cast<int>(npix), 1);
    for (int64_t i = 0; i < npix; ++i) im.pixels[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    const std::vector<uint8_t> out = img::develop(im, exposure, static_cast<img::ToneMap>(tm), log_decades);
    std::memcpy(rgb, out.data(), out.size());
}

// img::add_glare (image.cpp:47-106) on a row-major w x h XYZ image, in place.
void bhref_add_glare(double* xyz, int32_t w, int32_t h, double strength, double radius_px) {
    img::Image im(w, h);
    const int64_t n = static_cast<int64_t>(w) * h;
    for (int64_t i = 0; i < n; ++i) im.pixels[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    img::add_glare(im, strength, radius_px);
    for (int64_t i = 0; i < n; ++i) {
        xyz[3 * i] = im.pixels[i].x; xyz[3 * i + 1] = im.pixels[i].y; xyz[3 * i + 2] = im.pixels[i].z;
    }
}

void bhref_blackbody_xyz(double T, double* o) {
    const spec::XYZ c = spec::blackbody_xyz(T);
    o[0] = c.x; o[1] = c.y; o[2] = c.z;
}

double bhref_blackbody_bolometric(double T) { return spec::blackbody_bolometric(T); }

// NovikovThorneDisc::temperature (disc.h:100-104) of the scene's disc at r.
double bhref_disc_temperature(const bhx_scene* s, double r) {
    const bh::NovikovThorneDisc d(s->M_kg, s->spin, s->disc_r_out, s->eddington, s->disc_sense);
    return d.temperature(r);
}

void bhref_star_radiance(int32_t count, uint64_t seed, const double* d, double sigma, double* o) {
    const bh::StarField sf(count, seed);
    const spec::XYZ c = sf.radiance(d, sigma);
    o[0] = c.x; o[1] = c.y; o[2] = c.z;
}

// main.cpp:399-423 setup_jet, on a scene whose jet_* fields hold the Jet
// defaults (or user values).
void bhref_setup_jet(bhx_scene* s) {
    bh::RenderConfig cfg = to_config(s);
    bh::Jet& j = cfg.jet;
    const bh::NovikovThorneDisc d(cfg.M_kg, cfg.spin, cfg.disc_r_out, cfg.eddington, cfg.disc_sense);
    j.z_sat = std::min(j.z_sat, 0.7 * j.z_top);
    j.configure(cfg.spin, d.mdot_si() * phys::c * phys::c, phys::r_g_metres(cfg.M_kg));
    s->jet_z_sat = j.z_sat; s->jet_a = j.a; s->jet_Omega_F = j.Omega_F; s->jet_j_nu0 = j.j_nu0;
    s->jet_r_bound = j.r_bound;
}

// ns::star_of_mass for the polytrope P = K rho^Gamma (tov.h:205-218, eos.h:153-167).
int bhref_polytrope_star(double K_si, double Gamma, double M_kg, double* R_m, double* M_out) {
    const ns::Polytrope eos(K_si, Gamma);
    const ns::Star st = ns::star_of_mass(eos, M_kg);
    *R_m = st.R;
    *M_out = st.M;
    return st.ok ? 0 : 1;
}

int bhref_polytrope_solve(double K_si, double Gamma, double rho_c, double* R_m, double* M_out) {
    const ns::Polytrope eos(K_si, Gamma);
    const ns::Star st = ns::solve_tov(eos, rho_c);
    *R_m = st.R;
    *M_out = st.M;
    return st.ok ? 0 : 1;
}

// img::write_png / img::write_apng (image.cpp:652-762) into a file at path.
int bhref_write_png(const char* path, const uint8_t* rgb, int32_t w, int32_t h) {
    const std::vector<uint8_t> v(rgb, rgb + static_cast<size_t>(w) * h * 3);
    return img::write_png(path, v, w, h) ? 0 : 1;
}

int bhref_write_apng(const char* path, const uint8_t* frames, int32_t n, int32_t w, int32_t h, int32_t fps,
                     int32_t loops) {
    const size_t sz = static_cast<size_t>(w) * h * 3;
    std::vector<std::vector<uint8_t>> f;
    for (int32_t i = 0; i < n; ++i) f.emplace_back(frames + i * sz, frames + (i + 1) * sz);
    return img::write_apng(path, f, w, h, fps, loops) ? 0 : 1;
}

}  // extern "C"
