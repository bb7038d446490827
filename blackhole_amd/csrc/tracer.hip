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
//   k_raygen  one thread per sample: jitter, camera, photon, first RHS; the
//             initial state is written as 16 SoA rows in work-queue order
//             (8 x 8-pixel tiles, centre-first so the long photon-ring rays
//             start early);
//   k_trace   persistent grid sized to the resident-wave capacity; one ray per
//             lane with its whole state in VGPRs; a wave takes `chunk` queue
//             slots from a global atomic counter and, whenever at least
//             `refill` lanes have finished, hands its idle lanes the next slots
//             by ballot + prefix rank (mbcnt) - consecutive slots, so the
//             refill loads are coalesced; one loop iteration = one DP5 attempt
//             on every live lane.
#include "dense_dev.h"
#include "dp5.h"
#include "raygen_dev.h"
#include "events_dev.h"
#include "kerr_dev.h"
#include "params.h"
#include "../../include/bhx.h"

#include <hip/hip_runtime.h>

namespace bhx {
namespace {

using namespace dp5;

constexpr int kTile = 8;
// per-lane LDS slot of a waiting conical-jet sample: dense-output coefficients
// of r, theta, p_r (5 each), s_jet, h
constexpr int kPend = 17;

// Four-point Gauss-Legendre nodes and weights on [0, 1] (render.cpp:47-50).
__constant__ const double kGLx[4] = {0.06943184420297371, 0.33000947820757187, 0.66999052179242813,
                                     0.93056815579702629};
__constant__ const double kGLw[4] = {0.17392742256872693, 0.32607257743127307, 0.32607257743127307,
                                     0.17392742256872693};

// A wave-uniform double moved into scalar registers.
__device__ __forceinline__ double uniform_d(double x) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(b));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(b >> 32));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// ---------------------------------------------------------------------------
// Ray generation: one block of 64 threads per queue tile.
// ---------------------------------------------------------------------------
// Sample k of pixel (px, py): camera, photon, Carter Q0 and the first RHS,
// written to queue slot q of the SoA initial-state rows (row stride N).
__device__ __forceinline__ void gen_sample(const RaygenArgs& P, int px, int py, int k, int q, int ray,
                                           int64_t N) {
    double d[3];
    camera_dir(P, px, py, k, d);
    double pt, pr, pth, pph;
    camera_photon(P, d, pt, pr, pth, pph);
    // Carter Q at the camera (kerr.h:342-348)
    const double E = -pt;
    const double q0 = pth * pth + P.cam_cs2 * (P.a2 * (0.0 - E * E) + pph * pph / P.cam_sn2);
    Rhs f;
    kerr_rhs(ray_k(P.a, P.a2, pt, pph), P.cam_r, P.cam_th, pr, pth, f);
    double* I = P.init;
    I[kI_T * N + q] = 0.0;
    I[kI_R * N + q] = P.cam_r;
    I[kI_TH * N + q] = P.cam_th;
    I[kI_PH * N + q] = P.cam_ph;
    I[kI_PR * N + q] = pr;
    I[kI_PTH * N + q] = pth;
    I[kI_PT * N + q] = pt;
    I[kI_PPH * N + q] = pph;
#pragma unroll
    for (int c = 0; c < 6; ++c) I[(kI_K1 + c) * N + q] = f.d[c];
    I[kI_C0 * N + q] = f.cs;
    I[kI_Q0 * N + q] = q0;
    P.q2ray[q] = ray;
}

// Queue tile t of a scene: its 8 x 8 pixels' samples at their queue slots.
__device__ __forceinline__ void raygen_tile(const RaygenArgs& P, int t) {
    if (t >= P.n_tiles) return;
    const int tile = P.tile_order[t];
    const int tx = tile % P.tiles_x, ty = tile / P.tiles_x;
    const int x0 = tx * kTile, y0 = ty * kTile;
    const int tw = min(kTile, P.W - x0), th = min(kTile, P.H - y0);
    const int S2 = P.S * P.S;
    const int count = tw * th * S2;
    const int base = P.tile_start[t];
    const int64_t N = P.n_rays;
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        const int pin = i / S2, k = i - pin * S2;
        const int px = x0 + pin % tw, py = y0 + pin / tw;
        const int q = base + i;
        gen_sample(P, px, py, k, q, P.ray_base + (py * P.W + px) * S2 + k, N);
        if (P.q2scene) P.q2scene[q] = (uint8_t)P.scene_id;
    }
}

__global__ void __launch_bounds__(64) k_raygen(const RaygenArgs P) { raygen_tile(P, blockIdx.x); }

// Every scene of a batch in one launch: block b generates queue tile
// b - first[s] of scene s = scene[b] (scene arguments from the table).
__global__ void __launch_bounds__(64) k_raygen_batch(const RaygenArgs* __restrict__ table,
                                                     const int32_t* __restrict__ scene,
                                                     const int32_t* __restrict__ first) {
    const int s = scene[blockIdx.x];
    raygen_tile(table[s], (int)blockIdx.x - first[s]);
}

// Pilot rays of the cost-ordered queue: one per 8 x 8 tile, at the tile's
// centre pixel (clamped into a partial edge tile), first sample; slot = ray =
// row-major tile index, row stride n_tiles.
__global__ void __launch_bounds__(256) k_pilot_gen(const RaygenArgs P) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P.n_tiles) return;
    const int tx = t % P.tiles_x, ty = t / P.tiles_x;
    const int px = min(tx * kTile + kTile / 2, P.W - 1), py = min(ty * kTile + kTile / 2, P.H - 1);
    gen_sample(P, px, py, 0, t, t, P.n_tiles);
}

// ---------------------------------------------------------------------------
// The persistent tracer.
// ---------------------------------------------------------------------------
// The tracer body, shared by the single-scene kernel (kBatch = false: every
// scene constant is a kernel argument, uniform, in scalar registers) and the
// batched kernel (kBatch = true: each lane's ray belongs to its own scene; the
// spin and tolerances live in the lane's registers, everything else in the
// LDS copy `sk` of the scene table).  kLean drops the jet and Sun branches at
// compile time (scenes of disc, neutron star, horizon and sky only): the
// kernel then holds its state in 168 VGPRs and runs three waves per SIMD
// instead of two; the arithmetic of every branch it keeps is unchanged.
#define SK(f) (kBatch ? sk[sid].f : P.f)
#define EV (kBatch ? sk[sid].ev : *P.ev)
#define HAS_JET (!kLean && SK(has_jet))
#define HAS_SUN (!kLean && SK(has_sun))
template <bool kBatch, bool kPilot = false, bool kLean = false>
__device__ __forceinline__ void trace_body(const TraceArgs& P, const SceneK* sk, double* pend,
                                           unsigned long long* blk) {
    // ---- per-scene constants -------------------------------------------------
    double la = P.a, la2 = P.a2, lrtol = P.rtol, latol = P.atol;
    int sid = 0;

    // ---- lane state ------------------------------------------------------
    bool live = false;
    int q = 0;                          // queue slot of the lane's ray
    double y[6], k1[6];                 // t, r, th, ph, p_r, p_th
    double pc[3] = {0.0, 0.0, 0.0};     // y's pseudo-Cartesian point (jet cull, Sun): a step's end is the next one's start
    double pt = 0.0, pph = 0.0;         // conserved p_t, p_phi
    double kapt = 0.0, kapph = 0.0, ka2pt2 = 0.0, kpph2 = 0.0;   // their products in RayK
    double h = 0.0, lpw = 0.0, c_prev = 0.0, q0raw = 0.0, iq0 = 0.0;   // lpw = beta log(err_prev)
    double mx_norm = 0.0, mx_carter = 0.0, jet_acc = 0.0;
    int nsteps = 0, natt = 0, tries = 0;
    // conical jet: a sample of the last step waits (jet_pend) in the lane's LDS
    // slot my_pend; a ray that ended writes its jet integral at the next loop
    // head (jet_out), after that sample
    bool jet_pend = false, jet_out = false;
    double* my_pend = kLean ? nullptr : pend + threadIdx.x * kPend;
#pragma unroll
    for (int c = 0; c < 6; ++c) { y[c] = 0.0; k1[c] = 0.0; }

    // ---- per-step event constants ------------------------------------------
    // The single-scene kernel reads the jet cull and the Sun test's constants
    // once, into scalar registers: read inside the step loop, each would
    // expose a full memory latency on every step.  The batched kernel reads
    // them from its LDS table.
    double ev_r_bound = 0.0, ev_cone_angle = 0.0, ev_sun_R = 0.0, ev_sun_c[3] = {0.0, 0.0, 0.0};
    int ev_model = 0;
    if constexpr (!kBatch && !kLean) {
        if (P.has_jet || P.has_sun) {
            ev_r_bound = uniform_d(P.ev->r_bound);
            ev_cone_angle = uniform_d(P.ev->cone_angle);
            ev_sun_R = uniform_d(P.ev->sun_R);
#pragma unroll
            for (int c = 0; c < 3; ++c) ev_sun_c[c] = uniform_d(P.ev->sun_c[c]);
            ev_model = __builtin_amdgcn_readfirstlane(P.ev->model);
        }
    }
#define EVS(f) (kBatch ? sk[sid].ev.f : ev_##f)

    // ---- wave-uniform queue state ------------------------------------------
    int c_next = 0, c_end = 0;
    bool q_done = false;
    const int lane = __lane_id();
    const int64_t N = P.n_rays;

    for (;;) {
        // ---- conical jet (model 1): kConeSub x 4-point Gauss-Legendre -------
        // on the dense output over [0, s_jet] of each lane's last step.  A
        // sampled step costs ~10 plain steps, so the wave shares the nodes:
        // here at the loop head every lane is active, each pass serves two
        // sampling lanes (owners), 32 lanes per owner evaluate one node each
        // from the owner's LDS slot, and a butterfly sums each half - the
        // balanced tree over the nodes in order.  (A wave that is not whole
        // here lets each owner sum its own nodes along the same tree.)
        {
            const unsigned long long owners = __ballot(jet_pend);
            if (owners != 0ull) {
                const bool coop = __ballot(true) == ~0ull;
                unsigned long long rem = owners;
                double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0, l4 = 0.0, jsum = 0.0;
                const int n_pass = coop ? (__popcll(owners) + 1) / 2 : kConeSub * 4;
#pragma unroll 1
                for (int pass = 0; pass < n_pass; ++pass) {
                    int owner = lane, j = pass, A = -1, B = -1;
                    if (coop) {
                        A = __builtin_ctzll(rem);
                        rem &= rem - 1;
                        if (rem) {
                            B = __builtin_ctzll(rem);
                            rem &= rem - 1;
                        }
                        owner = (lane >= 32 && B >= 0) ? B : A;
                        j = lane & 31;
                    }
                    const double* op = pend + (threadIdx.x - lane + owner) * kPend;
                    const double opt = __shfl(pt, owner), opph = __shfl(pph, owner);
                    double ola = la, ola2 = la2;
                    int osid = 0;
                    if constexpr (kBatch) {
                        osid = __shfl(sid, owner);
                        ola = __shfl(la, owner);
                        ola2 = __shfl(la2, owner);
                    }
                    const EventParams& JO = kBatch ? sk[osid].ev : *P.ev;
                    // node j: sub-interval j / 4, Gauss point j % 4
                    const int qd = j & 3;
                    const double sv = ((j >> 2) + kGLx[qd]) * (1.0 / kConeSub) * op[15];
                    const double rs = dense_comp(op[0], op[1], op[2], op[3], op[4], sv);
                    const double ths = dense_comp(op[5], op[6], op[7], op[8], op[9], sv);
                    double sn, cs;
                    sincos_near(ths, &sn, &cs);
                    double v = 0.0;
                    const double sh = cone_shape(JO, rs, cs);
                    if (sh > 0.0) {
                        const double prs = dense_comp(op[10], op[11], op[12], op[13], op[14], sv);
                        const double b = cone_boost(JO, ola, ola2, rs, sn, cs, opt, prs, opph);
                        if (b > 0.0) v = kGLw[qd] * sh * b;
                    }
                    if (coop) {
#pragma unroll
                        for (int off = 1; off < 32; off <<= 1) v = v + __shfl_xor(v, off);
                        const double sA = __shfl(v, 0), sB = __shfl(v, 32);
                        if (lane == A) jsum = sA;
                        if (lane == B) jsum = sB;
                    } else {
                        // node j closes the subtrees of its trailing one bits
                        const int t = __builtin_ctz(~j);
                        double c = v;
                        if (t > 0) c = l0 + c;
                        if (t > 1) c = l1 + c;
                        if (t > 2) c = l2 + c;
                        if (t > 3) c = l3 + c;
                        if (t > 4) c = l4 + c;
                        if (t == 0) l0 = c;
                        else if (t == 1) l1 = c;
                        else if (t == 2) l2 = c;
                        else if (t == 3) l3 = c;
                        else if (t == 4) l4 = c;
                        else jsum = c;
                    }
                }
                if (jet_pend) {
                    jet_acc += EV.j_nu0 * jsum * my_pend[16] * my_pend[15] * (1.0 / kConeSub);
                    jet_pend = false;
                }
            }
        }
        if (jet_out) {   // the ray ended on the last step
            if (P.jet) P.jet[P.q2ray[q]] = jet_acc;
            jet_out = false;
        }

        // ---- refill idle lanes from the wave's chunk -----------------------
        unsigned long long idle = __ballot(!live);
        const int n_idle = __popcll(idle);
        if (!q_done && (n_idle >= P.refill || n_idle == 64)) {
            while (idle != 0ull) {
                if (c_next >= c_end) {
                    uint32_t b = 0;
                    if (lane == 0) b = atomicAdd(P.counter, (uint32_t)P.chunk);
                    b = __builtin_amdgcn_readfirstlane(b);
                    if (b >= (uint64_t)N) { q_done = true; break; }
                    c_next = (int)b;
                    c_end = (int)min(N, (int64_t)b + P.chunk);
                }
                const int avail = c_end - c_next;
                const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                if (!live && rank < avail) {
                    q = c_next + rank;
                    if constexpr (kBatch) {
                        sid = P.q2scene[q];
                        la = sk[sid].a;
                        la2 = sk[sid].a2;
                        lrtol = sk[sid].rtol;
                        latol = sk[sid].atol;
                    }
                    const double* I = P.init;
#pragma unroll
                    for (int c = 0; c < 6; ++c) {
                        y[c] = I[c * N + q];
                        k1[c] = I[(kI_K1 + c) * N + q];
                    }
                    pt = I[kI_PT * N + q];
                    pph = I[kI_PPH * N + q];
                    {
                        const RayK K0 = ray_k(la, la2, pt, pph);
                        kapt = K0.apt;
                        kapph = K0.apph;
                        ka2pt2 = K0.a2pt2;
                        kpph2 = K0.pph2;
                    }
                    c_prev = I[kI_C0 * N + q];
                    if (HAS_JET || HAS_SUN) to_cart(la2, y[1], y[2], y[3], pc);
                    q0raw = I[kI_Q0 * N + q];
                    iq0 = rcp(fabs(q0raw) + 1e-12);
                    h = SK(h0);
                    lpw = kBeta * kLog1em4;   // err_prev = 1e-4 (geodesic.h:197)
                    nsteps = natt = tries = 0;
                    mx_norm = mx_carter = jet_acc = 0.0;
                    live = SK(max_steps) > 0;
                    if (!live) {  // the step loop never runs: Exhausted at y0
                        const int ray = P.q2ray[q];
                        if (P.cls) P.cls[ray] = BHX_EXHAUSTED;
                        if (P.y) {
                            const int64_t M = N;
                            P.y[ray] = y[0]; P.y[M + ray] = y[1]; P.y[2 * M + ray] = y[2];
                            P.y[3 * M + ray] = y[3]; P.y[4 * M + ray] = pt; P.y[5 * M + ray] = y[4];
                            P.y[6 * M + ray] = y[5]; P.y[7 * M + ray] = pph;
                        }
                        if (P.jet) P.jet[ray] = 0.0;
                        if (P.steps) P.steps[ray] = 0;
                        if (P.attempts) P.attempts[ray] = 0;
                        if (P.norm_err) P.norm_err[ray] = 0.0;
                        if (P.carter) P.carter[ray] = 0.0;
                    }
                }
                c_next += min(__popcll(idle), avail);
                idle = __ballot(!live);
            }
        }
        // A refill runs whenever all 64 lanes are idle and only stops with every
        // lane live or the queue drained, so no live lane means no work left.
        if (__ballot(live) == 0ull) break;
        if (!live) continue;

        // ---- one DP5 attempt (geodesic.h:130-194) ---------------------------
        {
            double hmax = 0.05 * dmax(1.0, y[1] - SK(r_h)) + 0.01;
            if constexpr (kPilot) hmax *= kPilotHmax;   // the pilot's coarse steps
            h = (tries == 0 && h > hmax) ? hmax : h;   // clamp once per step, not per retry
        }
        double tmp[6], k2[6], k3[6], k4[6], k5[6], k6[6], yn[6];
        RayK K;
        K.a = la; K.a2 = la2; K.pt = pt; K.pph = pph;
        K.apt = kapt; K.apph = kapph; K.a2pt2 = ka2pt2; K.pph2 = kpph2;
        Rhs f;
        // stage inputs y + sum_j (h a_ij) k_j: the step size is folded into
        // the tableau once per attempt (an FMA per term, no separate scaling);
        // they only feed the right-hand side, the new state keeps y + h sum b_j k_j
        const double h21 = h * A21;
#pragma unroll
        for (int c = 0; c < 6; ++c) tmp[c] = fma(h21, k1[c], y[c]);
        kerr_rhs(K, tmp[1], tmp[2], tmp[4], tmp[5], f);
        const double h31 = h * A31, h32 = h * A32;
#pragma unroll
        for (int c = 0; c < 6; ++c) { k2[c] = f.d[c]; tmp[c] = fma(h32, k2[c], fma(h31, k1[c], y[c])); }
        kerr_rhs(K, tmp[1], tmp[2], tmp[4], tmp[5], f);
        const double h41 = h * A41, h42 = h * A42, h43 = h * A43;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            k3[c] = f.d[c];
            tmp[c] = fma(h43, k3[c], fma(h42, k2[c], fma(h41, k1[c], y[c])));
        }
        kerr_rhs(K, tmp[1], tmp[2], tmp[4], tmp[5], f);
        const double h51 = h * A51, h52 = h * A52, h53 = h * A53, h54 = h * A54;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            k4[c] = f.d[c];
            tmp[c] = fma(h54, k4[c], fma(h53, k3[c], fma(h52, k2[c], fma(h51, k1[c], y[c]))));
        }
        kerr_rhs(K, tmp[1], tmp[2], tmp[4], tmp[5], f);
        const double h61 = h * A61, h62 = h * A62, h63 = h * A63, h64 = h * A64, h65 = h * A65;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            k5[c] = f.d[c];
            tmp[c] = fma(h65, k5[c], fma(h64, k4[c], fma(h63, k3[c], fma(h62, k2[c], fma(h61, k1[c], y[c])))));
        }
        kerr_rhs(K, tmp[1], tmp[2], tmp[4], tmp[5], f);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            k6[c] = f.d[c];
            yn[c] = y[c] + h * (B1 * k1[c] + B3 * k3[c] + B4 * k4[c] + B5 * k5[c] + B6 * k6[c]);
        }
        Rhs g7;
        kerr_rhs(K, yn[1], yn[2], yn[4], yn[5], g7);
        double err2 = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double e = h * (E1 * k1[c] + E3 * k3[c] + E4 * k4[c] + E5 * k5[c] + E6 * k6[c] +
                                  E7 * g7.d[c]);
            // max of two magnitudes: y is an accepted state, so never NaN
            const double sc = latol + lrtol * fmax(fabs(y[c]), fabs(yn[c]));
            const double qe = e * rcp_coarse(sc);
            err2 += qe * qe;   // the p_t, p_phi terms are exactly 0
        }
        // The RMS error is err = sqrt(u), u = err2 / 8 (geodesic.h:157-166),
        // used only through err <= 1 and log err, so the square root is never
        // taken: sqrt(u) <= 1 exactly when u <= 1 + 2^-52 (sqrt rounds 1 + 2^-52
        // to 1), and log err = log(u) / 2.  A non-finite u stands for err = 1e10
        // (geodesic.h:166); min(u, 1e20) maps inf and NaN there, and any
        // u >= 1e20 is rejected with fac clamped to 0.2 all the same.
        const double u = fmin(err2 * 0.125, 1e20);
        // fac = 0.9 err^-alpha err_prev^beta (geodesic.h:170-173) as one
        // exponential, exp(beta log err_prev - alpha log err); u, lg, fac are
        // never NaN here, so max/min are the reference's std::max/std::clamp
        const double lg = 0.5 * log_pos(fmax(u, 1e-16 * 1e-16));
        const double fac = fmin(fmax(0.9 * exp_near(lpw - kAlpha * lg), 0.2), 5.0);
        ++natt;
        // Hairer's fifth dense-output coefficient of component c, built only on
        // the steps that locate an event or sample the jet
        auto dense_c5 = [&](int c) {
            return h * (D1 * k1[c] + D3 * k3[c] + D4 * k4[c] + D5 * k5[c] + D6 * k6[c] + D7 * g7.d[c]);
        };

        int term = -1;  // class once the ray ends
        double yo[6];
        if (u <= 1.0 + 0x1p-52) {
            // ---- accepted: invariants (render.cpp:140-145) -------------------
            ++nsteps;
            tries = 0;
            {
                // the five terms of g^{ab} p_a p_b (kerr.h:301-321) from the FSAL geometry
                const double p1 = yn[4], p2 = yn[5], rn = yn[1];
                const double iQ = g7.iS * g7.iD;
                const double A = g7.rpa * g7.rpa - la2 * g7.D * g7.sn2;
                const double t0 = -A * iQ * K.pt * K.pt;
                const double t1 = -4.0 * la * rn * iQ * K.pt * K.pph;
                const double t2 = g7.D * g7.iS * p1 * p1, t3 = g7.iS * p2 * p2;
                const double t4 = (g7.D - la2 * g7.sn2) * iQ * g7.isn2 * K.pph2;
                const double nrm = t0 + t1 + t2 + t3 + t4;
                const double sc = fabs(t0) + fabs(t1) + fabs(t2) + fabs(t3) + fabs(t4);
                const double rel = (sc > 0.0) ? nrm * rcp(sc) : nrm;
                mx_norm = dmax(mx_norm, fabs(rel));
                const double E = -pt;
                const double Q = p2 * p2 + g7.cs * g7.cs * (la2 * (0.0 - E * E) + pph * pph * g7.isn2);
                mx_carter = dmax(mx_carter, fabs(Q - q0raw) * iq0);
            }
            const bool ns_hit = SK(has_ns) && yn[1] <= SK(ns_R);
            const bool horizon = !ns_hit && yn[1] <= SK(r_stop);
            bool disc_cross = !ns_hit && !horizon && SK(has_disc) && c_prev * g7.cs < 0.0;
            if (disc_cross) {
                // Where the step's r(s) stays outside [r_in, r_out] the crossing
                // cannot hit the disc, and its root is not located: on [0, 1],
                // r(s) - (c1 + s c2) = s(1-s)(c3 + s(c4 + (1-s) c5)) is at most
                // |c3|/4 + 4|c4|/27 + |c5|/16, and c1 + s c2 lies between the
                // end points.  The 1e-12 margins cover the rounding of the bound
                // and of the root's own evaluation, so no hit is ever skipped.
                const double c2r = yn[1] - y[1];
                const double c3r = h * k1[1] - c2r;
                const double c4r = c2r - h * g7.d[1] - c3r;
                const double dev = 0.25 * fabs(c3r) + (4.0 / 27.0) * fabs(c4r) + 0.0625 * fabs(dense_c5(1));
                const double lo = fmin(y[1], yn[1]) - dev, hi = fmax(y[1], yn[1]) + dev;
                disc_cross = lo <= SK(r_out) * (1.0 + 1e-12) && hi >= SK(r_in) * (1.0 - 1e-12);
            }
            // jet cull and Sun candidate need the pseudo-Cartesian end points of the step
            bool near_jet = false, sun_cand = false;
            double p0[3], p1[3];
            if (HAS_JET || HAS_SUN) {
#pragma unroll
                for (int c = 0; c < 3; ++c) p0[c] = pc[c];
                to_cart(la2, yn[1], yn[2], yn[3], p1);
                if (HAS_JET) near_jet = seg_dist(p0, p1) < EVS(r_bound);        // render.cpp:151-152
                // conical jet: only steps whose dense output can reach the cone
                if (near_jet && EVS(model) == 1) {
                    const double c2t = yn[2] - y[2];
                    const double c3t = h * k1[2] - c2t;
                    near_jet = cone_reach(y[2], c2t, c3t, c2t - h * g7.d[2] - c3t, dense_c5(2), EVS(cone_angle));
                }
                if (HAS_SUN && !ns_hit && !horizon) {                             // render.cpp:203-209
                    const double e0 = p1[0] - p0[0], e1 = p1[1] - p0[1], e2 = p1[2] - p0[2];
                    const double seg_len = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
                    const double f0 = p0[0] - EVS(sun_c)[0], f1 = p0[1] - EVS(sun_c)[1], f2 = p0[2] - EVS(sun_c)[2];
                    sun_cand = sqrt(f0 * f0 + f1 * f1 + f2 * f2) <= seg_len + EVS(sun_R) * 1.5;
                }
            }
            double s_jet = -1.0;   // fraction of the step the jet integral covers
            double c5[6];          // set when the event branch below runs
            if (horizon) {
                // ---- horizon (render.cpp:172-177): after the NS test, before the disc
#pragma unroll
                for (int c = 0; c < 6; ++c) yo[c] = yn[c];
                term = BHX_HORIZON;
                s_jet = 1.0;
            }
            if (ns_hit || disc_cross || sun_cand || near_jet) {
                // dense output of the step (geodesic.h:177-185)
                double c1[6], c2[6], c3[6], c4[6];
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    c1[c] = y[c];
                    c2[c] = yn[c] - y[c];
                    c3[c] = h * k1[c] - c2[c];
                    c4[c] = c2[c] - h * g7.d[c] - c3[c];
                    c5[c] = dense_c5(c);
                }
                if (ns_hit) {
                    // ---- neutron-star surface (render.cpp:158-169) -------------
                    const double se = root_r(c1[1], c2[1], c3[1], c4[1], c5[1], SK(ns_R), y[1], yn[1]);
#pragma unroll
                    for (int c = 0; c < 6; ++c) yo[c] = dense_comp(c1[c], c2[c], c3[c], c4[c], c5[c], se);
                    term = BHX_STAR;
                    s_jet = se;
                }
                if (disc_cross) {
                    // ---- equatorial crossing (render.cpp:180-199) ---------------
                    const double se = root_theta(c1[2], c2[2], c3[2], c4[2], c5[2], c_prev, y[2], yn[2]);
                    const double rc = dense_comp(c1[1], c2[1], c3[1], c4[1], c5[1], se);
                    if (rc >= SK(r_in) && rc <= SK(r_out)) {
#pragma unroll
                        for (int c = 0; c < 6; ++c) yo[c] = dense_comp(c1[c], c2[c], c3[c], c4[c], c5[c], se);
                        term = BHX_DISC;
                        s_jet = se;
                    }
                }
                if (term < 0 && sun_cand) {
                    // ---- the Sun: 12 probes then 50 bisections (render.cpp:210-240)
                    const double* C = EV.sun_c;
                    const double R2 = EV.sun_R * EV.sun_R;
                    auto fsun = [&](double sv) {
                        double q[3];
                        to_cart(la2, dense_comp(c1[1], c2[1], c3[1], c4[1], c5[1], sv),
                                dense_comp(c1[2], c2[2], c3[2], c4[2], c5[2], sv),
                                dense_comp(c1[3], c2[3], c3[3], c4[3], c5[3], sv), q);
                        const double dx = q[0] - C[0], dy = q[1] - C[1], dz = q[2] - C[2];
                        return dx * dx + dy * dy + dz * dz - R2;
                    };
                    if (fsun(0.0) <= 0.0) {  // camera already inside: a hit at the step start
#pragma unroll
                        for (int c = 0; c < 6; ++c) yo[c] = y[c];
                        term = BHX_SUN;
                    } else {
                        double prev = 0.0;
                        for (int kp = 1; kp <= 12 && term < 0; ++kp) {
                            const double sv = (double)kp / 12;
                            if (fsun(sv) <= 0.0) {
                                double lo = prev, hi = sv;
                                for (int it = 0; it < 50; ++it) {
                                    const double mid = 0.5 * (lo + hi);
                                    if (fsun(mid) <= 0.0) hi = mid; else lo = mid;
                                }
                                const double se = 0.5 * (lo + hi);
#pragma unroll
                                for (int c = 0; c < 6; ++c) yo[c] = dense_comp(c1[c], c2[c], c3[c], c4[c], c5[c], se);
                                term = BHX_SUN;
                                s_jet = se;
                            }
                            prev = sv;
                        }
                    }
                }
                if (term < 0) s_jet = 1.0;   // nothing stopped the ray: the whole step counts
                if (near_jet && s_jet > 0.0 && EVS(model) != 1) {
                    // ---- jet emission: 4-point Gauss-Legendre (render.cpp:107-122)
                    const EventParams& J = EV;
                    double sum = 0.0;
#pragma unroll 1
                    for (int qd = 0; qd < 4; ++qd) {
                        const double sv = kGLx[qd] * s_jet;
                        const double rs = dense_comp(c1[1], c2[1], c3[1], c4[1], c5[1], sv);
                        const double ths = dense_comp(c1[2], c2[2], c3[2], c4[2], c5[2], sv);
                        double sn, cs;
                        sincos_near(ths, &sn, &cs);
                        const double sh = jet_shape(J, rs * cs, sqrt(rs * rs + la2) * sn);
                        if (sh <= 0.0) continue;
                        const double prs = dense_comp(c1[4], c2[4], c3[4], c4[4], c5[4], sv);
                        const double b = jet_boost(J, la, la2, rs, ths, pt, prs,
                                                   dense_comp(c1[5], c2[5], c3[5], c4[5], c5[5], sv), pph);
                        if (b <= 0.0) continue;
                        sum += kGLw[qd] * sh * b;
                    }
                    jet_acc += J.j_nu0 * sum * h * s_jet;
                }
            }
            if (near_jet && s_jet > 0.0 && EVS(model) == 1) {
                // ---- conical jet (model 1): the sample is taken at the next loop
                // head, where the whole wave is converged (see there); its dense
                // output (r, theta, p_r), s_jet and h wait in the lane's LDS slot
                const int cc[3] = {1, 2, 4};
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const int c = cc[m];
                    const double d2 = yn[c] - y[c];
                    const double d3 = h * k1[c] - d2;
                    my_pend[5 * m] = y[c];
                    my_pend[5 * m + 1] = d2;
                    my_pend[5 * m + 2] = d3;
                    my_pend[5 * m + 3] = d2 - h * g7.d[c] - d3;
                    my_pend[5 * m + 4] = c5[c];
                }
                my_pend[15] = s_jet;
                my_pend[16] = h;
                jet_pend = true;
            }
            if (term < 0 && yn[1] > SK(r_escape)) {
                // ---- escape (render.cpp:247-251), after the jet sample ----------
#pragma unroll
                for (int c = 0; c < 6; ++c) yo[c] = yn[c];
                term = BHX_SKY;
            }
            if (term < 0) {
#pragma unroll
                for (int c = 0; c < 6; ++c) { y[c] = yn[c]; k1[c] = g7.d[c]; }
                if (HAS_JET || HAS_SUN) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) pc[c] = p1[c];
                }
                c_prev = g7.cs;
                lpw = kBeta * dmax(lg, kLog1em4);   // err_prev = max(err, 1e-4)
                h *= fac;
                if (nsteps >= SK(max_steps)) {
#pragma unroll
                    for (int c = 0; c < 6; ++c) yo[c] = y[c];
                    term = BHX_EXHAUSTED;
                }
            }
        } else {
            // ---- rejected (geodesic.h:192, render.cpp:135) ----------------------
            h *= fmin(fac, 1.0);
            if (++tries > 40 || fabs(h) < 1e-14) {
                term = BHX_HORIZON;  // Hit::y keeps the initial photon state
                const double* I = P.init;
#pragma unroll
                for (int c = 0; c < 6; ++c) yo[c] = I[c * N + q];
            }
        }
        if (term >= 0) {
            const int ray = P.q2ray[q];
            if (P.cls) P.cls[ray] = (uint8_t)term;
            if (P.y) {
                const int64_t M = N;
                P.y[ray] = yo[0]; P.y[M + ray] = yo[1]; P.y[2 * M + ray] = yo[2];
                P.y[3 * M + ray] = yo[3]; P.y[4 * M + ray] = pt; P.y[5 * M + ray] = yo[4];
                P.y[6 * M + ray] = yo[5]; P.y[7 * M + ray] = pph;
            }
            jet_out = true;   // the jet integral is written at the next loop head
            if (P.steps) P.steps[ray] = nsteps;
            if (P.attempts) P.attempts[ray] = natt;
            if (P.norm_err) P.norm_err[ray] = mx_norm;
            if (P.carter) P.carter[ray] = mx_carter;
            atomicAdd(&blk[0], (unsigned long long)nsteps);
            atomicAdd(&blk[1], (unsigned long long)natt);
            atomicMax(&blk[2], (unsigned long long)__double_as_longlong(mx_norm));  // >= 0: orders as bits
            atomicMax(&blk[3], (unsigned long long)__double_as_longlong(mx_carter));
            live = false;
        }
    }

}
#undef SK
#undef EV
#undef EVS
#undef HAS_JET
#undef HAS_SUN

__device__ __forceinline__ void flush_totals(const TraceArgs& P, unsigned long long* blk) {
    // frame totals: one global atomic per block
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&P.totals[0], blk[0]);
        atomicAdd(&P.totals[1], blk[1]);
        atomicMax(&P.totals[2], blk[2]);
        atomicMax(&P.totals[3], blk[3]);
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_trace(const TraceArgs P) {
    __shared__ unsigned long long blk[4];  // steps, attempts, max norm bits, max Carter bits
    extern __shared__ double pend_raw[];   // kPend doubles per lane: waiting jet samples
    if (threadIdx.x < 4) blk[threadIdx.x] = 0ull;
    __syncthreads();
    trace_body<false>(P, nullptr, pend_raw, blk);
    flush_totals(P, blk);
}

// Scenes without jets or the Sun: the same tracer without those branches.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_trace_lean(const TraceArgs P) {
    __shared__ unsigned long long blk[4];
    if (threadIdx.x < 4) blk[threadIdx.x] = 0ull;
    __syncthreads();
    trace_body<false, false, true>(P, nullptr, nullptr, blk);
    flush_totals(P, blk);
}

// Pilot of the cost-ordered queue: the same tracer with the coarse hmax, at
// the tolerances and step cap its arguments carry.
template <bool kLean>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_trace_pilot(const TraceArgs P) {
    __shared__ unsigned long long blk[4];
    extern __shared__ double pend_raw[];
    if (threadIdx.x < 4) blk[threadIdx.x] = 0ull;
    __syncthreads();
    trace_body<false, true, kLean>(P, nullptr, pend_raw, blk);
    flush_totals(P, blk);
}

// Batched: the scene table (n_scenes rows) is copied into dynamic LDS first.
// kLean: no scene of the batch has jets or the Sun.
template <bool kLean>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_trace_batch(const TraceArgs P) {
    __shared__ unsigned long long blk[4];
    extern __shared__ double sk_raw[];
    SceneK* sk = reinterpret_cast<SceneK*>(sk_raw);
    if (threadIdx.x < 4) blk[threadIdx.x] = 0ull;
    {
        const double* src = reinterpret_cast<const double*>(P.scenes);
        const int nd = P.n_scenes * (int)(sizeof(SceneK) / sizeof(double));
        for (int i = threadIdx.x; i < nd; i += blockDim.x) sk_raw[i] = src[i];
    }
    __syncthreads();
    trace_body<true, false, kLean>(P, sk, sk_raw + P.n_scenes * (sizeof(SceneK) / sizeof(double)), blk);
    flush_totals(P, blk);
}

}  // namespace

hipError_t launch_raygen(const RaygenArgs& args, hipStream_t stream) {
    hipLaunchKernelGGL(k_raygen, dim3(args.n_tiles), dim3(64), 0, stream, args);
    return hipGetLastError();
}

hipError_t launch_raygen_batch(const RaygenArgs* table, const int32_t* block_scene, const int32_t* block_first,
                               int blocks, hipStream_t stream) {
    hipLaunchKernelGGL(k_raygen_batch, dim3(blocks), dim3(64), 0, stream, table, block_scene, block_first);
    return hipGetLastError();
}

// Host launchers: one persistent grid.
// Dynamic LDS: [scene table (batched only)] [kPend doubles per lane].
static size_t pend_bytes(int block) { return sizeof(double) * kPend * (size_t)block; }

// The lean kernel serves every scene without jets or the Sun.
bool trace_is_lean(const TraceArgs& args) { return !args.has_jet && !args.has_sun; }

hipError_t launch_trace(const TraceArgs& args, int grid, int block, hipStream_t stream) {
    if (trace_is_lean(args))
        hipLaunchKernelGGL(k_trace_lean, dim3(grid), dim3(block), 0, stream, args);
    else
        hipLaunchKernelGGL(k_trace, dim3(grid), dim3(block), pend_bytes(block), stream, args);
    return hipGetLastError();
}

hipError_t launch_pilot(const RaygenArgs& gen, const TraceArgs& args, int block, hipStream_t stream) {
    hipLaunchKernelGGL(k_pilot_gen, dim3((gen.n_tiles + 255) / 256), dim3(256), 0, stream, gen);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // every wave takes one chunk of 64 pilot rays: n_tiles / 64 waves in all
    const int waves = (gen.n_tiles + 63) / 64;
    const int grid = (waves * 64 + block - 1) / block;
    if (trace_is_lean(args))
        hipLaunchKernelGGL(k_trace_pilot<true>, dim3(grid), dim3(block), 0, stream, args);
    else
        hipLaunchKernelGGL(k_trace_pilot<false>, dim3(grid), dim3(block), pend_bytes(block), stream, args);
    return hipGetLastError();
}

// The batch's has_jet / has_sun say whether any of its scenes has them.
hipError_t launch_trace_batch(const TraceArgs& args, int grid, int block, hipStream_t stream) {
    const size_t table = sizeof(SceneK) * (size_t)args.n_scenes;
    if (trace_is_lean(args))
        hipLaunchKernelGGL(k_trace_batch<true>, dim3(grid), dim3(block), table, stream, args);
    else
        hipLaunchKernelGGL(k_trace_batch<false>, dim3(grid), dim3(block), table + pend_bytes(block), stream, args);
    return hipGetLastError();
}

hipError_t trace_occupancy(int block, bool lean, int* blocks_per_cu) {
    if (lean) return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_trace_lean, block, 0);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_trace, block, pend_bytes(block));
}

hipError_t trace_batch_occupancy(int block, int n_scenes, bool lean, int* blocks_per_cu) {
    const size_t table = sizeof(SceneK) * (size_t)n_scenes;
    if (lean) return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_trace_batch<true>, block, table);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_trace_batch<false>, block,
                                                        table + pend_bytes(block));
}

}  // namespace bhx
