/*
 * oracle.h - C ABI of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a CPU restatement of the
 * reference's per-pixel tracer (MattPar-Cell/Blackhole src/render.cpp,
 * geodesic.h, kerr.h, disc.h, scene.h, spectrum.cpp, image.cpp, jet.h,
 * neutronstar.h) used as the checker for the HIP path and as the CPU
 * baseline in bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (blackhole_amd/) never does.
 *
 * It follows the reference's operation order exactly and is built with
 * -O3 -ffp-contract=off (no FMA contraction, no fast-math), so that on the
 * same glibc it is bitwise equal to the reference compiled from source
 * (oracle/_ref, pinned in tests/test_oracle_ref.py and by the fixtures in
 * tests/golden/).
 */
#ifndef BH_ORACLE_H
#define BH_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../../include/bhx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A star catalogue in SoA form (host arrays). */
typedef struct oracle_catalogue {
    int32_t n;
    const double *x, *y, *z, *temperature, *irradiance;
} oracle_catalogue;

/* Per-ray outputs (host arrays, n entries; SoA component-major; any NULL). */
typedef struct oracle_rays {
    uint8_t *cls;
    double *y;          /* 8 x n */
    double *jet;
    int32_t *steps;
    int32_t *attempts;
    double *norm_err;
    double *carter;
    double *g;
    double *dir;        /* 3 x n */
    double *ray_xyz;    /* 3 x n */
} oracle_rays;

/* Whole frame (src/render.cpp:396-497).  cat == NULL: reference generator. */
int oracle_render(const bhx_scene *sc, const oracle_catalogue *cat, double *xyz,
                  bhx_stats *stats, int threads);

/* Trace + shade an explicit list of ray indices (index convention of bhx.h). */
int oracle_trace_list(const bhx_scene *sc, const oracle_catalogue *cat, const int64_t *ray_ids,
                      int64_t n, oracle_rays *out, int threads);

/* Trace + shade rays from explicit image-plane coordinates (u, v) in
 * [-1, 1] (render.cpp:447-451), e.g. to probe a pixel's neighbourhood. */
int oracle_trace_uv(const bhx_scene *sc, const oracle_catalogue *cat, const double *u,
                    const double *v, int64_t n, oracle_rays *out, int threads);
/* Image-plane coordinates of ray indices, with the reference's jitter. */
int oracle_ray_uv(const bhx_scene *sc, const int64_t *ids, int64_t n, double *u, double *v);

/* Trace one ray from an explicit initial photon state (no camera): used by
 * the physics checks.  Returns the class; y_out gets Hit::y. */
int oracle_trace_state(const bhx_scene *sc, const double *y0, double *y_out, double *jet,
                       int32_t *steps, double *norm_err, double *carter);

/* Pieces, for the property checks. */
void oracle_geodesic_rhs(double a, const double *y, double *dy);
void oracle_make_photon(double a, double r, double th, double phi, const double *dir,
                        double *y_out);
void oracle_invariants(double a, const double *y, double *E, double *Lz, double *Q,
                       double *norm, double *norm_rel);
/* DP5 integration with the reference's Stepper and a stop radius band. */
int oracle_integrate(double a, double *y, double h0, double rtol, double atol, int max_steps,
                     double r_min, double r_max, int64_t *nsteps);
double oracle_isco(double a, int sense);
double oracle_page_thorne_W(double a, double r, double r_isco);
double oracle_disc_temperature(const bhx_scene *sc, double r);
void oracle_flop_counts(int64_t *out);
void oracle_disc_intensity(const bhx_scene *sc, int64_t n, const uint8_t *cls, const double *r, const double *g,
                           double *out);
/* The conical jet (bhx_scene.jet_model 1, no reference counterpart). */
void oracle_cone_velocity(double a, double r, double th, double gamma, double *u);
void oracle_cone_shape(const bhx_scene *sc, int64_t n, const double *r, const double *cos_th, double *out);
void oracle_cone_gate(const bhx_scene *sc, int64_t n, const double *c, uint8_t *out);
double oracle_redshift(const double *y, const double *u);
double oracle_jet_boost(double alpha, double g);
void oracle_blackbody_xyz(double T, double *xyz);
void oracle_blackbody_xyz_exact(double T, double *xyz);
void oracle_powerlaw_xyz(double alpha, double *xyz);
void oracle_star_radiance(const bhx_scene *sc, const oracle_catalogue *cat, const double *d,
                          double sigma, double *xyz);
/* The reference's StarField(count, seed) catalogue, in generation order. */
int oracle_generate_catalogue(int32_t count, uint64_t seed, double *x, double *y, double *z,
                              double *temperature, double *irradiance);
/* The synthetic uniform sky of the 1024^2 frame (bhx_uniform_sky's table). */
int oracle_uniform_sky(int32_t count, uint64_t seed, double *x, double *y, double *z, double *temperature,
                       double *irradiance);
double oracle_auto_exposure(const double *xyz, int64_t n_pixels, double key);
/* img::add_glare (image.cpp:47-106) on a row-major w x h XYZ image, in place. */
void oracle_add_glare(double *xyz, int32_t w, int32_t h, double strength, double radius_px);
void oracle_develop(const double *xyz, int64_t n_pixels, double exposure, int tonemap,
                    double log_decades, uint8_t *rgb);

/* img::write_png / img::write_apng (image.cpp:193-778) as byte streams
 * (src/png.cpp).  rgb: w x h x 3 bytes, row-major; frames: n such images
 * back to back.  *len gets the stream's size; the bytes are copied to out
 * when cap >= *len (return 0), else 1 is returned.  -1: invalid arguments. */
int oracle_png_encode(const uint8_t *rgb, int32_t w, int32_t h, uint8_t *out, size_t cap, size_t *len);
int oracle_apng_encode(const uint8_t *frames, int32_t n, int32_t w, int32_t h, int32_t fps, int32_t loops,
                       uint8_t *out, size_t cap, size_t *len);
/* The writer's Huffman code-length builder (image.cpp:270-334). */
void oracle_huffman_lengths(const uint32_t *freq, int32_t n, int32_t max_bits, uint8_t *len);

#ifdef __cplusplus
}
#endif

#endif
