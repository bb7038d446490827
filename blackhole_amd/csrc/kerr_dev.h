// kerr_dev.h - FP64 device building blocks of the Kerr tracer (gfx950).
//
// Boyer-Lindquist Kerr metric, its closed-form r/theta partials and the
// Hamiltonian right-hand side, restructured for CDNA4's FP64 VALU:
//   * the 15 divisions of the textbook form (src/kerr.h:107-176) become one
//     Newton-refined reciprocal of Sigma Delta sin^2 theta shared by every
//     component (separable form of the contraction, see kerr_rhs);
//   * p_t and p_phi are constants of the motion (their RHS is exactly 0,
//     src/geodesic.h:64-65), so a ray carries them as scalars and integrates
//     only the six components (t, r, theta, phi, p_r, p_theta);
//   * one branch-free sincos per evaluation; the geometry of the last stage
//     (FSAL) is reused for the per-step invariants and the equatorial sign test.
// Results agree with the reference expressions to a few ulp; the rounding
// differences are far below the integrator's own error estimate noise.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bhx {

// 1/x to <= 1 ulp: hardware estimate + two Newton steps.  Non-finite or zero
// x gives a non-finite result, which the step controller rejects like the
// reference's inf/NaN from a division by zero.
__host__ __device__ __forceinline__ double rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / x;   // host builds of these functions (adjoint self-checks)
#endif
}

// 1/x to 2.2e-15 relative (hardware estimate, 2^-24.4, + one Newton step;
// measured over 4M arguments): for the error norm's scales, a step-size
// heuristic that needs no more.
__host__ __device__ __forceinline__ double rcp_coarse(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
#else
    return 1.0 / x;
#endif
}

// sin and cos of |x| <= 1e5 to <= 1 ulp of glibc away from their zeros, and
// to 1 ulp + |k| 4e-33 absolute next to x = k pi/2, where the two-term
// reduction's dropped term k (pi/2 - hi - lo) shows (tests/test_transcendentals.py):
// Cody-Waite reduction by pi/2 with a two-term split evaluated by FMA, then the minimax
// kernels of Sun's fdlibm (__kernel_sin / __kernel_cos, 1993) on
// [-pi/4, pi/4].  About 35 VALU ops, few registers and no branch, against ~70
// for the library routine whose Payne-Hanek branch would split every RHS into
// separate basic blocks.  theta cannot leave that range: each DP5 step moves it
// by less than pi and a ray takes at most 20000 steps.  NaN stays NaN.
// sincos_parts stops before the quadrant's common sign: x = r + q pi/2 gives
// (sin x, cos x) = (a, b), both signs flipped when bit 1 of q is set; it
// returns a, b and that flip as a sign-bit mask.  A caller that needs only
// sin^2, cos^2 and sin x cos x (the flips cancel in the product) skips it.
__host__ __device__ __forceinline__ void sincos_parts(double x, double* sa, double* cb, uint64_t* flip) {
    const double k = rint(x * 0.63661977236758134308);
    double r = fma(-k, 1.5707963267948966, x);
    r = fma(-k, 6.123233995736766e-17, r);
    const double z = r * r;
    const double ps = fma(z, fma(z, fma(z, fma(z, fma(z, 1.58969099521155010221e-10,
        -2.50507602534068634195e-08), 2.75573137070700676789e-06), -1.98412698298579493134e-04),
        8.33333333332248946124e-03), -1.66666666666666324348e-01);
    const double sr = fma(r * z, ps, r);
    const double pc = fma(z, fma(z, fma(z, fma(z, fma(z, -1.13596475577881948265e-11,
        2.08757232129817482790e-09), -2.75573143513906633035e-07), 2.48015872894767294178e-05),
        -1.38888888888741095749e-03), 4.16666666666666019037e-02);
    const double hz = 0.5 * z;
    const double w = fma(-0.5, z, 1.0);   // 1 - hz, hz exact
    const double cr = w + (((1.0 - w) - hz) + z * z * pc);
    // x = r + q pi/2: (sin x, cos x) = (sr, cr), (cr, -sr), (-sr, -cr), (-cr, sr)
    // for q = 0..3 (mod 4): pick by the low bit; bit 1 flips both signs
    const int q = static_cast<int>(k);
    const bool odd = (q & 1) != 0;
    *sa = odd ? cr : sr;
    *cb = odd ? -sr : cr;
    *flip = static_cast<uint64_t>(q & 2) << 62;
}

// Flip the sign of x where `flip` has the sign bit set (an exclusive or).
__host__ __device__ __forceinline__ double flip_sign(double x, uint64_t flip) {
    return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, x) ^ flip);
}

__host__ __device__ __forceinline__ void sincos_near(double x, double* s, double* c) {
    double a, b;
    uint64_t flip;
    sincos_parts(x, &a, &b, &flip);
    *s = flip_sign(a, flip);
    *c = flip_sign(b, flip);
}

// fma(a, b, c) for a constant c: on the device one v_fma_f64 taking c from a
// scalar register pair (the backend would otherwise copy c into vector
// registers to use the accumulating form, two extra vector moves per term of
// a Horner evaluation).
__host__ __device__ __forceinline__ double fma_k(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
#else
    return fma(a, b, c);
#endif
}

// exp(x) for |x| < 700 to <= 1 ulp (checked against glibc): Cody-Waite
// reduction by ln 2 (fdlibm's two-part constant), degree-13 Taylor-Horner on
// |r| <= ln(2)/2 (remainder < 4e-18), scale by 2^k.
__host__ __device__ __forceinline__ double exp_near(double x) {
    const double k = rint(x * 1.4426950408889634);
    double r = fma(-k, 6.93147180369123816490e-01, x);
    r = fma(-k, 1.90821492927058770002e-10, r);
    double p = fma_k(1.0 / 6227020800.0, r, 1.0 / 479001600.0);
    p = fma_k(p, r, 1.0 / 39916800.0);
    p = fma_k(p, r, 1.0 / 3628800.0);
    p = fma_k(p, r, 1.0 / 362880.0);
    p = fma_k(p, r, 1.0 / 40320.0);
    p = fma_k(p, r, 1.0 / 5040.0);
    p = fma_k(p, r, 1.0 / 720.0);
    p = fma_k(p, r, 1.0 / 120.0);
    p = fma_k(p, r, 1.0 / 24.0);
    p = fma_k(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, static_cast<int>(k));
}

// log(x) for finite normal x > 0 to <= 1 ulp (checked against glibc): the
// reduction and minimax kernel of fdlibm's __ieee754_log (Sun, 1993).
__host__ __device__ __forceinline__ double log_pos(double x) {
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) { m *= 2.0; e -= 1; }
    const double f = m - 1.0;
    const double s = f * rcp(2.0 + f);
    const double z = s * s, w = z * z;
    const double t2 = z * fma_k(w, fma_k(w, fma_k(w, 1.479819860511658591e-01, 1.818357216161805012e-01),
                                         2.857142874366239149e-01), 6.666666666666735130e-01);
    const double R = fma(w, fma_k(w, fma_k(w, 1.531383769920937332e-01, 2.222219843214978396e-01),
                                  3.999999999940941908e-01), t2);   // t2 + t1
    const double hfsq = 0.5 * f * f;
    const double dk = e;
    return dk * 6.93147180369123816490e-01 - ((hfsq - (s * (hfsq + R) + dk * 1.90821492927058770002e-10)) - f);
}

// true for finite x; the compiler builtin, so host and device builds agree
__host__ __device__ __forceinline__ bool finite_d(double x) { return __builtin_isfinite(x); }

__host__ __device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max
__host__ __device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min
__host__ __device__ __forceinline__ double dclamp(double v, double lo, double hi) {             // std::clamp
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}

// Per-ray constants of the momentum equations (p_t, p_phi are conserved).
struct RayK {
    double a, a2, pt, pph;
    double apt, apph, a2pt2, pph2;   // a p_t, a p_phi, a^2 p_t^2, p_phi^2
};
__host__ __device__ __forceinline__ RayK ray_k(double a, double a2, double pt, double pph) {
    RayK k;
    k.a = a; k.a2 = a2; k.pt = pt; k.pph = pph;
    k.apt = a * pt; k.apph = a * pph; k.a2pt2 = a2 * pt * pt; k.pph2 = pph * pph;
    return k;
}

// Right-hand side plus the geometry the per-step invariants reuse.
struct Rhs {
    double d[6];                    // dt, dr, dtheta, dphi, dp_r, dp_theta
    double cs, sn2, isn2;           // cos(theta), clamped sin^2, 1/sin^2
    double iS, iD, D, rpa;          // 1/Sigma, 1/Delta, Delta, r^2 + a^2
};

// Hamilton's equations dx/dl = g^{mu nu} p_nu, dp/dl = -1/2 d(g^{ab}) p_a p_b
// (src/geodesic.h:50-71) for the Kerr inverse metric of src/kerr.h:107-176,
// written through the separable form of the contraction
//     Sigma g^{ab} p_a p_b = Delta p_r^2 + p_th^2 + U^2 / sin^2 - P^2 / Delta,
//     P = (r^2 + a^2) p_t + a p_phi,   U = p_phi + a p_t sin^2,
// whose r and theta partials are
//     d_r  = [Delta' p_r^2 - 4 r p_t P / Delta + P^2 Delta' / Delta^2 - 2 r H] / Sigma,
//     d_th = 2 sin cos [a^2 p_t^2 - p_phi^2 / sin^4 + a^2 H] / Sigma,
// with H = g^{ab} p_a p_b.  Algebraically identical to the component-wise
// partials of the reference (checked against the oracle's RHS to rounding),
// about 40% fewer operations, and a single reciprocal of Sigma Delta sin^2.
// The same with sin and cos of theta given (sn, cs).
__host__ __device__ __forceinline__ void kerr_rhs_sc(const RayK& K, double r, double sn, double cs, double pr,
                                                     double pth, Rhs& o) {
    // |sin theta| >= 1e-9 with its sign kept (kerr.h:34,51), as one max and a
    // sign copy; a NaN angle gives +-1e-9 here but its cosine stays NaN
    sn = copysign(fmax(fabs(sn), 1e-9), sn);
    const double sn2 = sn * sn;
    const double r2 = r * r;
    const double rpa = fma(r, r, K.a2);
    const double S = r2 + K.a2 * (cs * cs);
    const double D = rpa - 2.0 * r;
    const double SD = S * D;
    const double iX = rcp(SD * sn2);          // 1 / (Sigma Delta sin^2)
    const double iSD = sn2 * iX;              // 1 / (Sigma Delta)
    const double iS = D * iSD;
    const double iD = S * iSD;
    const double isn2 = SD * iX;
    const double P = rpa * K.pt + K.apph;
    const double PiD = P * iD;
    const double U = K.pph + K.apt * sn2;
    const double pr2 = pr * pr;
    const double N = D * pr2 + pth * pth + U * U * isn2 - P * PiD;
    const double H = N * iS;
    const double Dr = 2.0 * r - 2.0;
    const double br = Dr * pr2 + PiD * (Dr * PiD - 4.0 * r * K.pt) - 2.0 * r * H;
    const double bt = K.a2pt2 + K.a2 * H - K.pph2 * isn2 * isn2;
    o.d[0] = (K.a * U - rpa * PiD) * iS;
    o.d[1] = D * pr * iS;
    o.d[2] = pth * iS;
    o.d[3] = (K.apt + K.pph * isn2 - K.a * PiD) * iS;
    o.d[4] = -0.5 * iS * br;
    o.d[5] = -(sn * cs) * iS * bt;
    o.cs = cs;
    o.sn2 = sn2;
    o.isn2 = isn2;
    o.iS = iS;
    o.iD = iD;
    o.D = D;
    o.rpa = rpa;
}

// The right-hand side depends on sin theta and cos theta only through sin^2,
// cos^2 and sin cos, in which the quadrant's common sign flip cancels: the
// unflipped pair goes in, and only the cosine kept for the event tests (o.cs)
// gets its sign - the same bits as with the flipped pair.
__host__ __device__ __forceinline__ void kerr_rhs(const RayK& K, double r, double th, double pr, double pth,
                                         Rhs& o) {
    double sa, cb;
    uint64_t flip;
    sincos_parts(th, &sa, &cb, &flip);
    kerr_rhs_sc(K, r, sa, cb, pr, pth, o);
    o.cs = flip_sign(cb, flip);
}

// splitmix64 (the reference's Rng, src/scene.h:27-44), counter form: draw n
// (1-based) of a stream seeded with s is mix(s + n * golden).
__host__ __device__ __forceinline__ uint64_t splitmix_draw(uint64_t seed, uint64_t n) {
    uint64_t z = seed + n * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ double splitmix_u01(uint64_t seed, uint64_t n) {
    return (double)(splitmix_draw(seed, n) >> 11) * 0x1.0p-53;
}

}  // namespace bhx
