// tracer.hip - persistent FP64 geodesic tracer for gfx950.
//
// Reproduces trace() of the reference (src/render.cpp:69-260): backwards
// photon from the camera, Dormand-Prince 5(4) with FSAL, RMS error norm and
// the PI controller of src/geodesic.h:130-194, the hmax clamp before each step
// (render.cpp:130-131), the 40-retry / |h| < 1e-14 horizon rule (:133-136),
// per-step invariants (:140-145), and the events in the reference's order:
// neutron-star surface (:158-169), horizon (:172-177), equatorial disc
// crossing with 60-step bisection on the dense output (:180-199), escape
// (:247-251), step budget (:257-259).
//
// Execution model (MI355X-first, not the reference's thread-per-row pool):
//   * one ray per lane; the whole ray state lives in VGPRs for its lifetime;
//   * persistent grid sized to the resident-wave capacity of the chip;
//   * each wave owns an 8 x 8-pixel tile taken from a global atomic counter
//     (tiles ordered centre-first so the long photon-ring rays start early);
//   * when at least `refill` lanes have finished, the wave ballots its idle
//     lanes and hands them the next rays of its tile by prefix rank (mbcnt);
//   * one loop iteration = one DP5 attempt on every live lane, so rejected
//     steps, events and refills never stall the other 63 lanes for long.
#include "kerr_dev.h"
#include "params.h"
#include "../../include/bhx.h"

#include <hip/hip_runtime.h>

namespace bhx {
namespace {

// Dormand-Prince 5(4) tableau, error weights and Hairer's dense-output
// weights (Hairer, Norsett & Wanner, Solving ODEs I, II.4).
constexpr double A21 = 1.0 / 5.0;
constexpr double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
constexpr double A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
constexpr double A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0, A53 = 64448.0 / 6561.0,
                 A54 = -212.0 / 729.0;
constexpr double A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0, A63 = 46732.0 / 5247.0,
                 A64 = 49.0 / 176.0, A65 = -5103.0 / 18656.0;
constexpr double B1 = 35.0 / 384.0, B3 = 500.0 / 1113.0, B4 = 125.0 / 192.0,
                 B5 = -2187.0 / 6784.0, B6 = 11.0 / 84.0;
constexpr double E1 = B1 - 5179.0 / 57600.0, E3 = B3 - 7571.0 / 16695.0,
                 E4 = B4 - 393.0 / 640.0, E5 = B5 - (-92097.0 / 339200.0),
                 E6 = B6 - 187.0 / 2100.0, E7 = 0.0 - 1.0 / 40.0;
constexpr double D1 = -12715105075.0 / 11282082432.0, D3 = 87487479700.0 / 32700410799.0,
                 D4 = -10690763975.0 / 1880347072.0, D5 = 701980252875.0 / 199316789632.0,
                 D6 = -1453857185.0 / 822651844.0, D7 = 69997945.0 / 29380423.0;
// PI controller exponents alpha = 0.7/5, beta = 0.4/5, safety 0.9 (geodesic.h:170).
constexpr double kAlpha = 0.7 / 5.0, kBeta = 0.4 / 5.0;

constexpr int kTile = 8;

__device__ __forceinline__ double dense_comp(double c1, double c2, double c3, double c4, double c5,
                                             double s) {
    const double s1 = 1.0 - s;
    return c1 + s * (c2 + s1 * (c3 + s * (c4 + s1 * c5)));
}

// cos(th) * c0 <= 0 for c0 = cos of a double (never 0): the sign of cos(th)
// is read off th near the equator, where cos(fl(pi/2)) = 6.1e-17 > 0 and the
// next double up has a negative cosine; elsewhere cos is evaluated.
__device__ __forceinline__ bool crossed(double th, double c0) {
    if (th > 0.5 && th < 2.6) {
        const bool cpos = th <= 1.5707963267948966;
        return (c0 > 0.0) ? !cpos : cpos;
    }
    return cos(th) * c0 <= 0.0;
}

__device__ __forceinline__ unsigned long long dbl_bits(double x) {
    return static_cast<unsigned long long>(__double_as_longlong(x));
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_trace(const TraceArgs P) {
    // ---- lane state ------------------------------------------------------
    bool live = false;
    int64_t ray = 0;
    double y[6], k1[6];                 // t, r, th, ph, p_r, p_th
    double pt = 0.0, pph = 0.0;         // conserved p_t, p_phi
    double h = 0.0, pw = 0.0, c_prev = 0.0, q0raw = 0.0, iq0 = 0.0;
    double mx_norm = 0.0, mx_carter = 0.0;
    int nsteps = 0, natt = 0, tries = 0;
    unsigned long long tot_steps = 0, tot_att = 0;
    double lane_norm = 0.0, lane_carter = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) { y[c] = 0.0; k1[c] = 0.0; }

    // ---- wave-uniform queue state ------------------------------------------
    int t_next = 0, t_count = 0, t_x0 = 0, t_y0 = 0, t_w = 1;
    bool q_done = false;
    const int S2 = P.S * P.S;
    const int lane = __lane_id();

    for (;;) {
        // ---- refill idle lanes from the wave's tile ------------------------
        unsigned long long idle = __ballot(!live);
        int n_idle = __popcll(idle);
        if (!q_done && (n_idle >= P.refill || n_idle == 64)) {
            while (idle != 0ull) {
                if (t_next >= t_count) {
                    uint32_t t = 0;
                    if (lane == 0) t = atomicAdd(P.tile_counter, 1u);
                    t = __builtin_amdgcn_readfirstlane(t);
                    if (t >= (uint32_t)P.n_tiles) { q_done = true; break; }
                    const int tile = P.tile_order[t];
                    const int tx = tile % P.tiles_x, ty = tile / P.tiles_x;
                    t_x0 = tx * kTile;
                    t_y0 = ty * kTile;
                    t_w = min(kTile, P.W - t_x0);
                    const int t_h = min(kTile, P.H - t_y0);
                    t_count = t_w * t_h * S2;
                    t_next = 0;
                }
                const int avail = t_count - t_next;
                const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                if (!live && rank < avail) {
                    // ---- start a ray: jitter, camera, photon, k1 --------------
                    const int q = t_next + rank;
                    const int pin = q / S2, k = q - pin * S2;
                    const int px = t_x0 + pin % t_w, py = t_y0 + pin / t_w;
                    ray = ((int64_t)py * P.W + px) * S2 + k;
                    const uint64_t sd = P.seed ^ ((uint64_t)py << 32) ^ ((uint64_t)px * 0x9E3779B97F4A7C15ull);
                    const int sx = k % P.S, sy = k / P.S;
                    const double jx = (sx + splitmix_u01(sd, 2 * (uint64_t)k + 1)) * P.inv_S;
                    const double jy = (sy + splitmix_u01(sd, 2 * (uint64_t)k + 2)) * P.inv_S;
                    const double u = (2.0 * (px + jx) / P.W - 1.0);
                    const double v = (1.0 - 2.0 * (py + jy) / P.H);
                    // Camera::ray_direction (scene.h:356-384)
                    double cx = u * P.tanf, cy = v * P.tanf / P.aspect;
                    const double cz = 1.0;
                    if (P.has_roll) {
                        const double xr = P.c_roll * cx - P.s_roll * cy;
                        const double yr = P.s_roll * cx + P.c_roll * cy;
                        cx = xr; cy = yr;
                    }
                    const double y2 = P.c_pitch * cy - P.s_pitch * cz;
                    const double z2 = P.s_pitch * cy + P.c_pitch * cz;
                    const double x2 = P.c_yaw * cx + P.s_yaw * z2;
                    const double z3 = -P.s_yaw * cx + P.c_yaw * z2;
                    const double v0 = -z3, v1 = -y2, v2 = x2;
                    const double n = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                    const double d0 = v0 / n, d1 = v1 / n, d2 = v2 / n;
                    // make_photon, past-directed (geodesic.h:215-231)
                    const double pu0 = -P.e00;
                    const double pu1 = d0 * P.e11;
                    const double pu2 = d1 * P.e22;
                    const double pu3 = -(P.e03 - d2 * P.e33);
                    pt = P.gl_tt * pu0 + P.gl_tp * pu3;
                    const double pr0 = P.gl_rr * pu1;
                    const double ph0 = P.gl_hh * pu2;
                    pph = P.gl_tp * pu0 + P.gl_pp * pu3;
                    y[0] = 0.0; y[1] = P.cam_r; y[2] = P.cam_th; y[3] = P.cam_ph;
                    y[4] = pr0; y[5] = ph0;
                    // Carter Q at the camera (kerr.h:342-348)
                    const double E = -pt;
                    q0raw = ph0 * ph0 + P.cam_cs2 * (P.a2 * (0.0 - E * E) + pph * pph / P.cam_sn2);
                    iq0 = 1.0 / (fabs(q0raw) + 1e-12);
                    Rhs f;
                    kerr_rhs(P.a, P.a2, y[1], y[2], pt, y[4], y[5], pph, f);
#pragma unroll
                    for (int c = 0; c < 6; ++c) k1[c] = f.d[c];
                    c_prev = f.cs;
                    h = P.h0;
                    pw = P.pw_init;
                    nsteps = natt = tries = 0;
                    mx_norm = mx_carter = 0.0;
                    // Hit::y defaults to the initial state (render.cpp:88-89)
                    if (P.y) {
                        const int64_t N = P.n_rays;
                        P.y[ray] = 0.0; P.y[N + ray] = y[1]; P.y[2 * N + ray] = y[2];
                        P.y[3 * N + ray] = y[3]; P.y[4 * N + ray] = pt; P.y[5 * N + ray] = pr0;
                        P.y[6 * N + ray] = ph0; P.y[7 * N + ray] = pph;
                    }
                    live = true;
                    if (P.max_steps <= 0) live = false;  // loop never runs: Exhausted at y0
                    if (!live) {
                        if (P.cls) P.cls[ray] = BHX_EXHAUSTED;
                        if (P.jet) P.jet[ray] = 0.0;
                        if (P.steps) P.steps[ray] = 0;
                        if (P.attempts) P.attempts[ray] = 0;
                        if (P.norm_err) P.norm_err[ray] = 0.0;
                        if (P.carter) P.carter[ray] = 0.0;
                    }
                }
                t_next += min(__popcll(idle), avail);
                idle = __ballot(!live);
            }
        }
        // A refill runs whenever all 64 lanes are idle and only stops with every
        // lane live or the queue drained, so no live lane means no work left.
        if (__ballot(live) == 0ull) break;
        if (!live) continue;

        // ---- one DP5 attempt (geodesic.h:130-194) ---------------------------
        if (tries == 0) {
            const double hmax = 0.05 * dmax(1.0, y[1] - P.r_h) + 0.01;
            if (h > hmax) h = hmax;
        }
        double tmp[6], k2[6], k3[6], k4[6], k5[6], k6[6], yn[6];
        Rhs f;
#pragma unroll
        for (int c = 0; c < 6; ++c) tmp[c] = y[c] + h * A21 * k1[c];
        kerr_rhs(P.a, P.a2, tmp[1], tmp[2], pt, tmp[4], tmp[5], pph, f);
#pragma unroll
        for (int c = 0; c < 6; ++c) { k2[c] = f.d[c]; tmp[c] = y[c] + h * (A31 * k1[c] + A32 * k2[c]); }
        kerr_rhs(P.a, P.a2, tmp[1], tmp[2], pt, tmp[4], tmp[5], pph, f);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            k3[c] = f.d[c];
            tmp[c] = y[c] + h * (A41 * k1[c] + A42 * k2[c] + A43 * k3[c]);
        }
        kerr_rhs(P.a, P.a2, tmp[1], tmp[2], pt, tmp[4], tmp[5], pph, f);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            k4[c] = f.d[c];
            tmp[c] = y[c] + h * (A51 * k1[c] + A52 * k2[c] + A53 * k3[c] + A54 * k4[c]);
        }
        kerr_rhs(P.a, P.a2, tmp[1], tmp[2], pt, tmp[4], tmp[5], pph, f);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            k5[c] = f.d[c];
            tmp[c] = y[c] + h * (A61 * k1[c] + A62 * k2[c] + A63 * k3[c] + A64 * k4[c] + A65 * k5[c]);
        }
        kerr_rhs(P.a, P.a2, tmp[1], tmp[2], pt, tmp[4], tmp[5], pph, f);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            k6[c] = f.d[c];
            yn[c] = y[c] + h * (B1 * k1[c] + B3 * k3[c] + B4 * k4[c] + B5 * k5[c] + B6 * k6[c]);
        }
        Rhs g7;
        kerr_rhs(P.a, P.a2, yn[1], yn[2], pt, yn[4], yn[5], pph, g7);
        double err2 = 0.0, c5[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double e = h * (E1 * k1[c] + E3 * k3[c] + E4 * k4[c] + E5 * k5[c] + E6 * k6[c] +
                                  E7 * g7.d[c]);
            const double sc = P.atol + P.rtol * dmax(fabs(y[c]), fabs(yn[c]));
            const double q = e * rcp(sc);
            err2 += q * q;   // the p_t, p_phi terms are exactly 0
            c5[c] = h * (D1 * k1[c] + D3 * k3[c] + D4 * k4[c] + D5 * k5[c] + D6 * k6[c] +
                         D7 * g7.d[c]);
        }
        double err = sqrt(err2 * 0.125);
        if (!isfinite(err)) err = 1e10;
        const double lg = log(dmax(err, 1e-16));
        double fac = 0.9 * exp(-kAlpha * lg) * pw;
        fac = dclamp(fac, 0.2, 5.0);
        ++natt;

        int term = -1;  // class once the ray ends
        if (err <= 1.0) {
            // ---- accepted: invariants (render.cpp:140-145) -------------------
            ++nsteps;
            tries = 0;
            {
                const double p1 = yn[4], p2 = yn[5];
                const double t0 = g7.gtt * pt * pt, t1 = 2.0 * g7.gtp * pt * pph;
                const double t2 = g7.grr * p1 * p1, t3 = g7.ghh * p2 * p2, t4 = g7.gpp * pph * pph;
                const double nrm = t0 + t1 + t2 + t3 + t4;
                const double sc = fabs(t0) + fabs(t1) + fabs(t2) + fabs(t3) + fabs(t4);
                const double rel = (sc > 0.0) ? nrm / sc : nrm;
                mx_norm = dmax(mx_norm, fabs(rel));
                const double E = -pt;
                const double Q = p2 * p2 + g7.cs * g7.cs * (P.a2 * (0.0 - E * E) + pph * pph * g7.isn2);
                mx_carter = dmax(mx_carter, fabs(Q - q0raw) * iq0);
            }
            double c1[6], c2[6], c3[6], c4[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                c1[c] = y[c];
                c2[c] = yn[c] - y[c];
                c3[c] = h * k1[c] - c2[c];
                c4[c] = c2[c] - h * g7.d[c] - c3[c];
            }
            double yo[6];
            if (P.has_ns && yn[1] <= P.ns_R) {
                // ---- neutron-star surface (render.cpp:158-169) -----------------
                double lo = 0.0, hi = 1.0;
                for (int it = 0; it < 60; ++it) {
                    const double mid = 0.5 * (lo + hi);
                    if (dense_comp(c1[1], c2[1], c3[1], c4[1], c5[1], mid) <= P.ns_R) hi = mid; else lo = mid;
                }
                const double se = 0.5 * (lo + hi);
#pragma unroll
                for (int c = 0; c < 6; ++c) yo[c] = dense_comp(c1[c], c2[c], c3[c], c4[c], c5[c], se);
                term = BHX_STAR;
            } else if (yn[1] <= P.r_stop) {
                // ---- horizon (render.cpp:172-177) -------------------------------
#pragma unroll
                for (int c = 0; c < 6; ++c) yo[c] = yn[c];
                term = BHX_HORIZON;
            } else {
                if (P.has_disc && c_prev * g7.cs < 0.0) {
                    // ---- equatorial crossing (render.cpp:180-199) ---------------
                    double lo = 0.0, hi = 1.0;
                    for (int it = 0; it < 60; ++it) {
                        const double mid = 0.5 * (lo + hi);
                        if (crossed(dense_comp(c1[2], c2[2], c3[2], c4[2], c5[2], mid), c_prev)) hi = mid;
                        else lo = mid;
                    }
                    const double se = 0.5 * (lo + hi);
                    const double rc = dense_comp(c1[1], c2[1], c3[1], c4[1], c5[1], se);
                    if (rc >= P.r_in && rc <= P.r_out) {
#pragma unroll
                        for (int c = 0; c < 6; ++c) yo[c] = dense_comp(c1[c], c2[c], c3[c], c4[c], c5[c], se);
                        term = BHX_DISC;
                    }
                }
                if (term < 0 && yn[1] > P.r_escape) {
                    // ---- escape (render.cpp:247-251) --------------------------------
#pragma unroll
                    for (int c = 0; c < 6; ++c) yo[c] = yn[c];
                    term = BHX_SKY;
                }
            }
            if (term < 0) {
#pragma unroll
                for (int c = 0; c < 6; ++c) { y[c] = yn[c]; k1[c] = g7.d[c]; }
                c_prev = g7.cs;
                pw = exp(kBeta * dmax(lg, -9.210340371976182));   // pow(max(err, 1e-4), beta)
                h *= fac;
                if (nsteps >= P.max_steps) {
#pragma unroll
                    for (int c = 0; c < 6; ++c) yo[c] = y[c];
                    term = BHX_EXHAUSTED;
                }
            }
            if (term >= 0 && P.y) {
                const int64_t N = P.n_rays;
                P.y[ray] = yo[0]; P.y[N + ray] = yo[1]; P.y[2 * N + ray] = yo[2];
                P.y[3 * N + ray] = yo[3]; P.y[4 * N + ray] = pt; P.y[5 * N + ray] = yo[4];
                P.y[6 * N + ray] = yo[5]; P.y[7 * N + ray] = pph;
            }
        } else {
            // ---- rejected (geodesic.h:192, render.cpp:135) ----------------------
            h *= dmin(fac, 1.0);
            if (++tries > 40 || fabs(h) < 1e-14) term = BHX_HORIZON;  // Hit::y stays y0
        }
        if (term >= 0) {
            if (P.cls) P.cls[ray] = (uint8_t)term;
            if (P.jet) P.jet[ray] = 0.0;
            if (P.steps) P.steps[ray] = nsteps;
            if (P.attempts) P.attempts[ray] = natt;
            if (P.norm_err) P.norm_err[ray] = mx_norm;
            if (P.carter) P.carter[ray] = mx_carter;
            tot_steps += nsteps;
            tot_att += natt;
            lane_norm = dmax(lane_norm, mx_norm);
            lane_carter = dmax(lane_carter, mx_carter);
            live = false;
        }
    }

    // ---- frame totals: one atomic per wave ---------------------------------
    for (int off = 32; off > 0; off >>= 1) {
        tot_steps += __shfl_xor(tot_steps, off);
        tot_att += __shfl_xor(tot_att, off);
        lane_norm = dmax(lane_norm, __shfl_xor(lane_norm, off));
        lane_carter = dmax(lane_carter, __shfl_xor(lane_carter, off));
    }
    if (lane == 0) {
        atomicAdd(&P.totals[0], tot_steps);
        atomicAdd(&P.totals[1], tot_att);
        atomicMax(&P.totals[2], dbl_bits(lane_norm));     // non-negative doubles order as bits
        atomicMax(&P.totals[3], dbl_bits(lane_carter));
    }
}

}  // namespace

// Host launcher: one persistent grid.
hipError_t launch_trace(const TraceArgs& args, int grid, int block, hipStream_t stream) {
    hipLaunchKernelGGL(k_trace, dim3(grid), dim3(block), 0, stream, args);
    return hipGetLastError();
}

hipError_t trace_occupancy(int block, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_trace, block, 0);
}

}  // namespace bhx
