// shade.hip - shading, display and colour kernels for gfx950.
//
//   k_shade    shade_hit() + shade() of every ray (src/render.cpp:265-393) and
//              the per-pixel sample average (src/render.cpp:443-462); one
//              thread per pixel visiting its samples in the reference's order
//   k_develop  img::develop (src/image.cpp:119-191) to 8-bit sRGB
//   k_exp_hist / k_exp_select  the device radix selection behind
//              img::auto_exposure (image.cpp:22-45)
//   k_bb       spec::blackbody_xyz (src/spectrum.cpp:123-134) of a T array
#include "kerr_dev.h"
#include "params.h"
#include "../../include/bhx.h"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace bhx {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kLgMin = 1.6989700043360187;  // log10(50 K)
constexpr double kLgMax = 9.4771212547196626;  // log10(3e9 K)
constexpr int kTabN = 4096;
constexpr double kSigmaSB = 5.670374419e-8;

// spec::blackbody_xyz: log-log interpolation in the 4096-entry table.
__device__ __forceinline__ void bb_lookup(const double* __restrict__ tab, double T, double o[3]) {
    const double lg = log10(dclamp(T, 50.0, 3.0e9));
    const double f = (lg - kLgMin) / (kLgMax - kLgMin) * (kTabN - 1);
    // A NaN temperature (clamp keeps it) gives row 0 and u = NaN: a NaN colour,
    // as the reference's; the conversion is guarded so NaN never reaches (int).
    const int i = (f > 0.0) ? (int)dmin(f, (double)(kTabN - 2)) : 0;
    const double u = f - i;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        o[c] = exp(tab[c * kTabN + i] * (1.0 - u) + tab[c * kTabN + i + 1] * u);
}

// Page-Thorne W(r) with the spin-only roots and coefficients precomputed on
// the host (src/disc.h:46-64).
__device__ double pt_W(const ShadeArgs& P, double r) {
    const double x = sqrt(r);
    double W = x - P.pt_x0 - 1.5 * P.disc_a * log(x / P.pt_x0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (P.pt_c[i] != 0.0) W -= P.pt_c[i] * log((x - P.pt_x[i]) / (P.pt_x0 - P.pt_x[i]));
    return W;
}

// NovikovThorneDisc::temperature (src/disc.h:94-121) at r.
__device__ double disc_T(const ShadeArgs& P, double r) {
    if (r <= P.r_in || r > P.r_out) return 0.0;
    const double a = P.disc_a;
    const double sg = (double)P.disc_sg;
    const double sr = sqrt(r);
    const double r32 = r * sr;
    const double Om = sg / (r32 + sg * a);
    const double dd = r32 + sg * a;
    const double dOm = -sg * 1.5 * sr / (dd * dd);
    const double root = sqrt(r * r - 3.0 * r + 2.0 * sg * a * sr);
    const double E = (r * r - 2.0 * r + sg * a * sr) / (r * root);
    const double L = (sg * sr * (r * r - 2.0 * sg * a * sr + a * a)) / (r * root);
    // flux_dimensionless's rejections give 0, which flux() still multiplies by
    // the scale: 0 * NaN is NaN when the efficiency is NaN (a < 0 with a
    // prograde disc), and temperature() then rejects only F <= 0.
    const double den = E - Om * L;
    double F = 0.0;
    if (den > 0.0) {
        F = (-dOm) / (4.0 * kPi * r * den * den) * pt_W(P, r);
        if (!(F > 0.0 && isfinite(F))) F = 0.0;
    }
    F = F * P.flux_scale;
    if (F <= 0.0) return 0.0;
    return pow(F / kSigmaSB, 0.25);
}

// StarField::radiance + milky_way_radiance (src/scene.h:111-158, 202-223),
// cells as CSR over stars sorted by cell in index order.
__device__ void sky_radiance(const ShadeArgs& P, const double d[3], double o[3]) {
    o[0] = o[1] = o[2] = 0.0;
    const double sigma = P.psf_sigma;
    const double cutoff = 3.0 * sigma;
    const double inv2s2 = 1.0 / (2.0 * sigma * sigma);
    const double norm = 1.0 / (2.0 * kPi * sigma * sigma);
    const double theta = acos(dclamp(d[2], -1.0, 1.0));
    const double phi = atan2(d[1], d[0]);
    const int it = (int)(theta / kPi * 360);
    const double dlat = kPi / 360;
    const int span_t = max(1, (int)ceil(cutoff / dlat));
    const double sn = dmax(sin(theta), 1e-6);
    const double dlon = 2.0 * kPi / 720;
    const int span_p = max(1, (int)ceil(cutoff / (sn * dlon)));
    const int sp = (span_p * 2 + 1 >= 720) ? 360 : span_p;
    const int p0 = (int)((phi + kPi) / (2.0 * kPi) * 720);
    for (int da = -span_t; da <= span_t; ++da) {
        const int ti = it + da;
        if (ti < 0 || ti >= 360) continue;
        for (int db = -sp; db <= sp; ++db) {
            int pi = p0 + db;
            pi = ((pi % 720) + 720) % 720;
            const int cell = ti * 720 + pi;
            const uint32_t b = P.cell_start[cell], e = P.cell_start[cell + 1];
            for (uint32_t s = b; s < e; ++s) {
                const double* st = P.star + 6 * (size_t)s;
                const double dot = dclamp(d[0] * st[0] + d[1] * st[1] + d[2] * st[2], -1.0, 1.0);
                const double ang = (dot > 0.9999995) ? sqrt(dmax(0.0, 2.0 * (1.0 - dot))) : acos(dot);
                if (ang > cutoff) continue;
                const double w = norm * exp(-ang * ang * inv2s2) * st[3];
                o[0] += w * st[4];
                o[1] += w;
                o[2] += w * st[5];
            }
        }
    }
    // diffuse Galactic light, 21.5 mag/arcsec^2 (scene.h:202-223)
    const double ca = 0.8660254037844387, sa = 0.5;
    const double g0 = d[0], g1 = ca * d[1] + sa * d[2], g2 = -sa * d[1] + ca * d[2];
    const double sin_b = dclamp(g2, -1.0, 1.0);
    const double band = exp(-fabs(sin_b) / 0.055);
    const double l = atan2(g1, g0);
    const double bulge = 0.55 + 0.45 * exp(-(l * l) / (2.0 * 0.45 * 0.45));
    const double lane = 0.65 + 0.35 * sin(l * 7.3 + sin_b * 41.0) * sin(l * 3.1 - sin_b * 17.0);
    const double amp = P.mw_L * band * bulge * dmax(0.0, lane);
    o[0] += amp * P.mw_cx;
    o[1] += amp;
    o[2] += amp * P.mw_cz;
}

// Three waves per SIMD (146 VGPRs, no scratch).  The file is compiled without
// machine loop-invariant code motion: with it the backend hoists the FP64
// constants of every inlined acos / atan2 / exp / log / pow out of the
// per-sample loop into registers for the whole kernel, which took 256 VGPRs
// and 180 bytes of scratch per lane at two waves per SIMD.
// Block `blk` of a scene's shading (256 pixels).
__device__ __forceinline__ void shade_block(const ShadeArgs& P, int blk) {
    __shared__ unsigned long long counts[7];
    __shared__ double flux_w[4];
    __shared__ unsigned long long ttot[4];
    if (threadIdx.x < 7) counts[threadIdx.x] = 0;
    if (threadIdx.x < 4) ttot[threadIdx.x] = 0;
    __syncthreads();
    const int pix = blk * blockDim.x + threadIdx.x;
    const int S2 = P.S * P.S;
    const int64_t N = P.n_rays;
    unsigned int my_counts[7] = {0, 0, 0, 0, 0, 0, 0};
    double my_flux = 0.0;
    unsigned long long my_steps = 0, my_att = 0;
    double my_norm = 0.0, my_carter = 0.0;
    if (pix < P.n_pix) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < S2; ++k) {
            const int64_t ray = P.ray_base + (int64_t)pix * S2 + k;
            const int cls = P.cls[ray];
            if (P.steps_in) {
                my_steps += (unsigned long long)P.steps_in[ray];
                my_att += (unsigned long long)P.attempts_in[ray];
                my_norm = dmax(my_norm, P.norm_in[ray]);
                my_carter = dmax(my_carter, P.carter_in[ray]);
            }
            double y[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) y[c] = P.y[c * N + ray];
            const double jet = P.jet ? P.jet[ray] : 0.0;
            double col[3] = {0.0, 0.0, 0.0};
            double g = 0.0, dv[3] = {0.0, 0.0, 0.0};
            ++my_counts[cls < 6 ? cls : 5];
            if (cls == BHX_DISC && P.has_disc) {
                const double r = y[1];
                double T = disc_T(P, r);
                // render.cpp:278 rejects T <= 0 only: a NaN temperature (a < 0
                // with a prograde disc makes the efficiency NaN) goes on to g
                // and a NaN colour there, and does here.
                if (!(T <= 0.0)) {
                    if (P.hs_enabled) {
                        const double t_emit = P.observer_time + y[0];
                        const double phs = P.hs_phi0 + P.hs_omega * t_emit;
                        const double dx = r * cos(y[3]) - P.hs_r * cos(phs);
                        const double dy = r * sin(y[3]) - P.hs_r * sin(phs);
                        const double f = 1.0 + P.hs_contrast * exp(-0.5 * (dx * dx + dy * dy) /
                                                                     (P.hs_sigma * P.hs_sigma));
                        T *= pow(f, 0.25);
                    }
                    // keplerian_four_velocity (kerr.h:260-266), redshift_factor (render.cpp:31-34)
                    const double sg = (double)P.disc_sg, a = P.disc_a;
                    const double r32 = r * sqrt(r);
                    const double Om = sg / (r32 + sg * a);
                    const double dsc = r * r * r - 3.0 * r * r + 2.0 * sg * a * r32;
                    const double ut = (r32 + sg * a) / sqrt(dmax(1e-300, dsc));
                    const double ku = y[4] * ut + y[7] * (Om * ut);
                    g = (ku > 1e-12) ? 1.0 / ku : 0.0;
                    if (g > 0.0) bb_lookup(P.bb, g * T, col);
                }
            } else if (cls == BHX_STAR && P.ns_enabled) {
                // rotating neutron-star surface (render.cpp:327-356, neutronstar.h:90-125)
                const double r = y[1], th = y[2], ph = y[3];
                const double sn = sin(th);
                const double R = P.ns_R, Om = P.ns_Omega;
                const double norm2 = 1.0 - 2.0 / R - Om * Om * R * R * sn * sn;
                const double ut = (norm2 > 1e-12) ? 1.0 / sqrt(norm2) : 1.0;
                const double ku = y[4] * ut + y[7] * (Om * ut);
                if (ku > 1e-12) {
                    g = 1.0 / ku;
                    double T = P.ns_T_eff;
                    if (P.ns_caps) {
                        const double phase = P.ns_cap_phase0 + Om * (P.observer_time + y[0]);
                        const double m0 = P.ns_sin_tilt * cos(phase), m1 = P.ns_sin_tilt * sin(phase);
                        const double dot = sn * cos(ph) * m0 + sn * sin(ph) * m1 + cos(th) * P.ns_cos_tilt;
                        double ang = acos(dclamp(dot, -1.0, 1.0));
                        if (P.ns_two_caps) ang = dmin(ang, kPi - ang);
                        if (ang <= P.ns_cap_radius) {
                            const double edge = 0.9 * P.ns_cap_radius;
                            if (ang <= edge) {
                                T = P.ns_cap_T;
                            } else {
                                const double uu = (ang - edge) / (P.ns_cap_radius - edge);
                                const double w = 0.5 * (1.0 + cos(kPi * uu));
                                T = P.ns_T_eff + (P.ns_cap_T - P.ns_T_eff) * w;
                            }
                        }
                    }
                    const double f = 1.0 - 2.0 / r;
                    const double mu = (f > 0.0) ? dclamp(-y[5] * sqrt(f) / ku, 0.0, 1.0) : 1.0;
                    const double limb = (2.0 + 3.0 * mu) / 5.0;
                    bb_lookup(P.bb, g * T, col);
                    col[0] *= limb; col[1] *= limb; col[2] *= limb;
                }
            } else if (cls == BHX_SUN && P.has_sun) {
                // the Sun, at rest in Boyer-Lindquist coordinates (render.cpp:301-325)
                const double r = y[1];
                double sn, cs;
                sincos(y[2], &sn, &cs);
                const double S = r * r + P.a2 * cs * cs;
                const double gtt = -(1.0 - 2.0 * r / S);
                if (gtt < 0.0) {
                    const double ku = y[4] * (1.0 / sqrt(-gtt));
                    g = (ku > 1e-12) ? 1.0 / ku : 0.0;
                    if (g > 0.0) {
                        double sp, cp;
                        sincos(y[3], &sp, &cp);
                        const double R = sqrt(r * r + P.a2);
                        const double px = R * sn * cp, py = R * sn * sp, pz = r * cs;
                        Rhs f;
                        kerr_rhs(ray_k(P.a, P.a2, y[4], y[7]), y[1], y[2], y[5], y[6], f);
                        const double dr = f.d[1], dth = f.d[2], dph = f.d[3];
                        const double dR = (R > 0.0) ? r / R * dr : dr;
                        double v0 = dR * sn * cp + R * cs * dth * cp - R * sn * sp * dph;
                        double v1 = dR * sn * sp + R * cs * dth * sp + R * sn * cp * dph;
                        double v2 = dr * cs - r * sn * dth;
                        const double vn = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                        if (vn > 0.0) { v0 /= vn; v1 /= vn; v2 /= vn; }
                        // Eddington limb darkening (scene.h:266-276)
                        double limb = 1.0;
                        double n0 = px - P.sun_c[0], n1 = py - P.sun_c[1], n2 = pz - P.sun_c[2];
                        const double len = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
                        if (len > 0.0) {
                            n0 /= len; n1 /= len; n2 /= len;
                            const double mu = dclamp(-(n0 * v0 + n1 * v1 + n2 * v2), 0.0, 1.0);
                            limb = (2.0 + 3.0 * mu) / 5.0;
                        }
                        bb_lookup(P.bb, g * P.sun_T_eff / (1.0 + P.sun_self_z), col);
                        col[0] *= limb; col[1] *= limb; col[2] *= limb;
                    }
                }
            } else if (cls == BHX_SKY) {
                // escape direction (render.cpp:15-26, 358-368)
                Rhs f;
                kerr_rhs(ray_k(P.a, P.a2, y[4], y[7]), y[1], y[2], y[5], y[6], f);
                const double r = y[1];
                const double R = sqrt(r * r + P.a2);
                double sn, cs, sp, cp;
                sincos(y[2], &sn, &cs);
                sincos(y[3], &sp, &cp);
                const double dr = f.d[1], dth = f.d[2], dph = f.d[3];
                const double dR = (R > 0.0) ? r / R * dr : dr;
                double v0 = dR * sn * cp + R * cs * dth * cp - R * sn * sp * dph;
                double v1 = dR * sn * sp + R * cs * dth * sp + R * sn * cp * dph;
                double v2 = dr * cs - r * sn * dth;
                const double vn = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                if (vn > 0.0) {
                    v0 /= vn; v1 /= vn; v2 /= vn;
                    dv[0] = v0; dv[1] = v1; dv[2] = v2;
                    if (P.stars_enabled) sky_radiance(P, dv, col);
                }
            }
            if (jet > 0.0) {
                ++my_counts[6];
                col[0] += P.jet_rgb[0] * jet;
                col[1] += P.jet_rgb[1] * jet;
                col[2] += P.jet_rgb[2] * jet;
            }
            if (P.ray_xyz) { P.ray_xyz[ray] = col[0]; P.ray_xyz[N + ray] = col[1]; P.ray_xyz[2 * N + ray] = col[2]; }
            if (P.g) P.g[ray] = g;
            if (P.dir) { P.dir[ray] = dv[0]; P.dir[N + ray] = dv[1]; P.dir[2 * N + ray] = dv[2]; }
            acc[0] += col[0]; acc[1] += col[1]; acc[2] += col[2];
        }
        const double w = 1.0 / (S2);
        if (P.xyz) {
            P.xyz[3 * (size_t)pix] = acc[0] * w;
            P.xyz[3 * (size_t)pix + 1] = acc[1] * w;
            P.xyz[3 * (size_t)pix + 2] = acc[2] * w;
        }
        my_flux = acc[1] * w;
    }
    // block reduction: class counters and the flux partial (fixed order)
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        unsigned long long v = my_counts[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (__lane_id() == 0 && v) atomicAdd(&counts[c], v);
    }
    for (int off = 32; off > 0; off >>= 1) my_flux += __shfl_xor(my_flux, off);
    if (__lane_id() == 0) flux_w[threadIdx.x / 64] = my_flux;
    if (P.steps_in) {
        for (int off = 32; off > 0; off >>= 1) {
            my_steps += __shfl_xor(my_steps, off);
            my_att += __shfl_xor(my_att, off);
            my_norm = dmax(my_norm, __shfl_xor(my_norm, off));
            my_carter = dmax(my_carter, __shfl_xor(my_carter, off));
        }
        if (__lane_id() == 0) {
            atomicAdd(&ttot[0], my_steps);
            atomicAdd(&ttot[1], my_att);
            atomicMax(&ttot[2], (unsigned long long)__double_as_longlong(my_norm));  // >= 0: orders as bits
            atomicMax(&ttot[3], (unsigned long long)__double_as_longlong(my_carter));
        }
    }
    __syncthreads();
    if (threadIdx.x < 7 && counts[threadIdx.x]) atomicAdd(&P.class_counts[threadIdx.x], counts[threadIdx.x]);
    if (P.steps_in && threadIdx.x == 0) {
        atomicAdd(&P.trace_totals[0], ttot[0]);
        atomicAdd(&P.trace_totals[1], ttot[1]);
        atomicMax(&P.trace_totals[2], ttot[2]);
        atomicMax(&P.trace_totals[3], ttot[3]);
    }
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) s += flux_w[w];
        P.flux_partial[blk] = s;
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_shade(const ShadeArgs P) {
    shade_block(P, blockIdx.x);
}

// Every scene of a batch in one launch: block b shades block b - first[s] of
// scene s = scene[b] (scene arguments from the table).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_shade_batch(const ShadeArgs* __restrict__ table, const int32_t* __restrict__ scene,
              const int32_t* __restrict__ first) {
    const int s = scene[blockIdx.x];
    shade_block(table[s], (int)blockIdx.x - first[s]);
}

// ---- display ---------------------------------------------------------------
__device__ __forceinline__ double aces(double x) {
    const double A = 2.51, B = 0.03, C = 2.43, D = 0.59, E = 0.14;
    return dclamp((x * (A * x + B)) / (x * (C * x + D) + E), 0.0, 1.0);
}
__device__ __forceinline__ double srgb_oetf(double v) {
    v = dclamp(v, 0.0, 1.0);
    return (v <= 0.0031308) ? 12.92 * v : 1.055 * pow(v, 1.0 / 2.4) - 0.055;
}

// The exposure of img::auto_exposure (src/image.cpp:37-44) from the device
// selection's state: key / (99.5th percentile), at most 3 key / (99.9th), 1
// without a positive pixel.  IEEE divisions, as the host computes it.
__device__ __forceinline__ double exposure_of(const ExposureState& st, double key) {
    if (st.n_pos == 0ull) return 1.0;
    const double hi = __longlong_as_double((long long)st.prefix[0]);
    const double peak = __longlong_as_double((long long)st.prefix[1]);
    double e = 1.0;
    if (hi > 0.0) e = key / hi;
    if (peak > 0.0) e = dmin(e, 3.0 * key / peak);
    return e;
}

// `st` non-null: the exposure is exposure_of(*st, key), computed where the
// selection left it, with no host round trip (and stored to *e_out by the
// first thread when e_out is non-null); otherwise `exposure` is used.
__global__ void k_develop(const double* __restrict__ xyz, int64_t n, double exposure,
                          const ExposureState* __restrict__ st, double key, double* __restrict__ e_out,
                          int tm, double log_k, uint8_t* __restrict__ rgb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (st) {
        exposure = exposure_of(*st, key);
        if (e_out && i == 0) *e_out = exposure;
    }
    if (i >= n) return;
    double c0 = xyz[3 * i], c1 = xyz[3 * i + 1], c2 = xyz[3 * i + 2];
    if (!isfinite(c0) || !isfinite(c1) || !isfinite(c2)) c0 = c1 = c2 = 0.0;
    c0 *= exposure; c1 *= exposure; c2 *= exposure;
    // XYZ (D65) -> linear sRGB (IEC 61966-2-1), gamut clamp towards white
    double v[3];
    v[0] = 3.2404542 * c0 - 1.5371385 * c1 - 0.4985314 * c2;
    v[1] = -0.9692660 * c0 + 1.8760108 * c1 + 0.0415560 * c2;
    v[2] = 0.0556434 * c0 - 0.2040259 * c1 + 1.0572252 * c2;
    const double m = dmin(dmin(v[0], v[1]), v[2]);
    if (!(m >= 0.0)) { v[0] -= m; v[1] -= m; v[2] -= m; }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        double x = dmax(v[ch], 0.0);
        switch (tm) {
            case BHX_TM_LINEAR: x = dclamp(x, 0.0, 1.0); break;
            case BHX_TM_REINHARD: x = x / (1.0 + x); break;
            case BHX_TM_ACES: x = aces(x); break;
            case BHX_TM_HDR: {
                const double cap = 0.12, w = 0.5;
                const double denom = log10(1.0 + log_k);
                const double f = aces(x);
                const double l = dclamp(log10(1.0 + log_k * dmax(x, 0.0)) / denom, 0.0, 1.0);
                const double lift = cap * tanh(w * l / cap);
                x = dclamp(f + lift * (1.0 - f), 0.0, 1.0);
                break;
            }
            case BHX_TM_LOG:
            default: {
                const double denom = log10(1.0 + log_k);
                x = dclamp(log10(1.0 + log_k * dmax(x, 0.0)) / denom, 0.0, 1.0);
                break;
            }
        }
        rgb[3 * i + ch] = (uint8_t)llround(255.0 * srgb_oetf(x));
    }
}

// img::auto_exposure's two order statistics (image.cpp:22-45) by radix
// selection on the bits of the positive finite luminances (positive doubles
// order like their bit patterns), on the device from start to end: six digit
// passes (12, 12, 12, 12, 12 and 4 bits from the top), each a histogram pass
// over the frame and a one-block selection of each statistic's digit.  The
// state between passes (count, ranks, prefixes) stays in device memory.
constexpr int kExpBins = 4096;
__host__ __device__ constexpr int exp_shift(int pass) { return pass < 5 ? 52 - 12 * pass : 0; }
__host__ __device__ constexpr int exp_width(int pass) { return pass < 5 ? 12 : 4; }

// Histograms of this pass's digit, one per statistic, of the keys whose bits
// above the digit equal that statistic's prefix; per-block counts in LDS.
__global__ void __launch_bounds__(256) k_exp_hist(const double* __restrict__ xyz, int64_t n, int pass,
                                                  const ExposureState* __restrict__ state,
                                                  unsigned int* __restrict__ hist) {
    __shared__ unsigned int h[2 * kExpBins];
    for (int i = threadIdx.x; i < 2 * kExpBins; i += blockDim.x) h[i] = 0u;
    __syncthreads();
    const int shift = exp_shift(pass), width = exp_width(pass), top = shift + width;
    const unsigned long long mask = (1ull << width) - 1ull;
    const unsigned long long p0 = state->prefix[0], p1 = state->prefix[1];
    const bool live = pass == 0 || state->n_pos > 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; live && i < n; i += stride) {
        const double Y = xyz[3 * i + 1];
        if (!(Y > 0.0 && isfinite(Y))) continue;
        const unsigned long long key = (unsigned long long)__double_as_longlong(Y);
        const unsigned int digit = (unsigned int)((key >> shift) & mask);
        if (pass == 0) {
            atomicAdd(&h[digit], 1u);
            continue;
        }
        const unsigned long long above = key >> top;
        if (above == p0) atomicAdd(&h[digit], 1u);
        if (above == p1) atomicAdd(&h[kExpBins + digit], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * kExpBins; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// One block of 256 threads: the digit of each statistic's rank in this pass's
// histogram (16 bins per thread, a block scan of the per-thread sums), the
// new prefix and remaining rank; clears the histograms for the next pass.
// Pass 0 also counts the positive finite luminances n and sets the ranks
// k = floor(q (n - 1)), q = 0.999 and 0.9999, as the reference does.
__global__ void __launch_bounds__(256) k_exp_select(int pass, ExposureState* __restrict__ state,
                                                    unsigned int* __restrict__ hist) {
    __shared__ unsigned long long part[256];
    __shared__ unsigned long long total;
    const int t = threadIdx.x;
    constexpr int kPer = kExpBins / 256;
    for (int j = 0; j < 2; ++j) {
        const unsigned int* hj = hist + (pass == 0 ? 0 : j * kExpBins);
        unsigned long long mine = 0;
        for (int b = 0; b < kPer; ++b) mine += hj[t * kPer + b];
        part[t] = mine;
        __syncthreads();
        // inclusive scan of the 256 partial sums (Hillis-Steele)
        for (int off = 1; off < 256; off <<= 1) {
            const unsigned long long v = (t >= off) ? part[t - off] : 0ull;
            __syncthreads();
            part[t] += v;
            __syncthreads();
        }
        if (pass == 0 && j == 0 && t == 255) {
            total = part[255];
            state->n_pos = total;
            if (total > 0) {
                for (int q = 0; q < 2; ++q) {
                    const double frac = q == 0 ? 0.999 : 0.9999;
                    unsigned long long k = (unsigned long long)(frac * (double)(total - 1));
                    state->k[q] = k < total - 1 ? k : total - 1;
                }
            }
        }
        __syncthreads();
        if (state->n_pos > 0) {
            const unsigned long long k = state->k[j];
            const unsigned long long before = part[t] - mine;
            if (k >= before && k < part[t]) {   // the rank falls in this thread's bins
                unsigned long long cum = before;
                int b = 0;
                for (; b < kPer; ++b) {
                    const unsigned long long c = hj[t * kPer + b];
                    if (k < cum + c) break;
                    cum += c;
                }
                state->k[j] = k - cum;
                state->prefix[j] = (state->prefix[j] << exp_width(pass)) | (unsigned long long)(t * kPer + b);
            }
        }
        __syncthreads();
    }
    for (int i = t; i < 2 * kExpBins; i += 256) hist[i] = 0u;
}

__global__ void k_bb(const double* __restrict__ tab, const double* __restrict__ T, int64_t n,
                     double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double o[3];
    bb_lookup(tab, T[i], o);
    out[3 * i] = o[0]; out[3 * i + 1] = o[1]; out[3 * i + 2] = o[2];
}

__global__ void k_fill_u64(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}
__global__ void k_fill_u32(unsigned int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

hipError_t launch_shade(const ShadeArgs& args, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_shade, dim3(blocks), dim3(256), 0, s, args);
    return hipGetLastError();
}
hipError_t launch_shade_batch(const ShadeArgs* table, const int32_t* block_scene, const int32_t* block_first,
                              int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_shade_batch, dim3(blocks), dim3(256), 0, s, table, block_scene, block_first);
    return hipGetLastError();
}
hipError_t launch_develop(const double* xyz, int64_t n, double exposure, int tm, double log_k,
                          uint8_t* rgb, hipStream_t s) {
    const int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_develop, dim3(blocks), dim3(256), 0, s, xyz, n, exposure,
                       static_cast<const ExposureState*>(nullptr), 0.0, static_cast<double*>(nullptr), tm, log_k,
                       rgb);
    return hipGetLastError();
}
hipError_t launch_develop_auto(const double* xyz, int64_t n, const ExposureState* state, double key,
                               double* e_out, int tm, double log_k, uint8_t* rgb, hipStream_t s) {
    const int blocks = (int)std::max<int64_t>(1, (n + 255) / 256);
    hipLaunchKernelGGL(k_develop, dim3(blocks), dim3(256), 0, s, xyz, n, 1.0, state, key, e_out, tm, log_k, rgb);
    return hipGetLastError();
}
hipError_t launch_exposure(const double* xyz, int64_t n, ExposureState* state, unsigned int* hist,
                           hipStream_t s) {
    const int blocks = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + 255) / 256));
    for (int pass = 0; pass < 6; ++pass) {
        hipLaunchKernelGGL(k_exp_hist, dim3(blocks), dim3(256), 0, s, xyz, n, pass,
                           static_cast<const ExposureState*>(state), hist);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_exp_select, dim3(1), dim3(256), 0, s, pass, state, hist);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
hipError_t launch_bb(const double* tab, const double* T, int64_t n, double* out, hipStream_t s) {
    const int blocks = (int)((n + 255) / 256);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bb, dim3(blocks), dim3(256), 0, s, tab, T, n, out);
    return hipGetLastError();
}
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}
hipError_t launch_zero_u32(unsigned int* p, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace bhx
