// selfcheck.cpp - the adjoint's hand-written reverse mode against central
// finite differences of the same forward code, on the CPU (host builds of the
// device functions).  Built and run by tests/test_adjoint_selfcheck.py:
//   rhs    vector-Jacobian product of the Hamiltonian right-hand side
//   step   one Dormand-Prince step with the step size frozen
//   term   the terminal step: r at the equatorial crossing of the dense output
//   ray    d I / d(spin, theta_cam, eddington, r_cam) of whole disc-hit rays
//          against finite differences of the full adaptive forward
//   jet    the same for a scene with the conical jet: rays lit by the jet
//          (any class) and disc hits seen through it
// Prints one line per check: name, count, worst relative error.
#include "../../blackhole_amd/csrc/adjoint_frame.h"

#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

using namespace bhx;
using namespace bhx::adj;

static double rel(double a, double b, double scale) { return std::fabs(a - b) / std::max(std::fabs(b), scale); }

static void check_rhs(std::mt19937_64& rng) {
    std::uniform_real_distribution<double> U(0.0, 1.0);
    double worst = 0.0;
    int n = 0;
    for (int t = 0; t < 200; ++t) {
        const double a = 0.998 * U(rng), pt = -1.0, pph = -5.0 + 10.0 * U(rng);
        const double x[4] = {3.0 + 40.0 * U(rng), 0.3 + 2.5 * U(rng), -2.0 + 4.0 * U(rng), -6.0 + 12.0 * U(rng)};
        const double w[4] = {U(rng) - 0.5, U(rng) - 0.5, U(rng) - 0.5, U(rng) - 0.5};
        double xb[4] = {}, cb[3] = {};
        rhs_vjp(ray_k(a, a * a, pt, pph), x, w, xb, cb);
        auto J = [&](const double* xx, double aa, double ptt, double pphh) {
            double k[4];
            rhs_x(ray_k(aa, aa * aa, ptt, pphh), xx, k);
            return w[0] * k[0] + w[1] * k[1] + w[2] * k[2] + w[3] * k[3];
        };
        double fd[7];
        for (int i = 0; i < 4; ++i) {
            double xp[4] = {x[0], x[1], x[2], x[3]}, xm[4] = {x[0], x[1], x[2], x[3]};
            const double h = 1e-6 * std::max(1.0, std::fabs(x[i]));
            xp[i] += h; xm[i] -= h;
            fd[i] = (J(xp, a, pt, pph) - J(xm, a, pt, pph)) / (2 * h);
        }
        const double h = 1e-6;
        fd[4] = (J(x, a + h, pt, pph) - J(x, a - h, pt, pph)) / (2 * h);
        fd[5] = (J(x, a, pt + h, pph) - J(x, a, pt - h, pph)) / (2 * h);
        fd[6] = (J(x, a, pt, pph + h) - J(x, a, pt, pph - h)) / (2 * h);
        const double ad[7] = {xb[0], xb[1], xb[2], xb[3], cb[0], cb[1], cb[2]};
        double scale = 0.0;
        for (int i = 0; i < 7; ++i) scale = std::max(scale, std::fabs(fd[i]));
        for (int i = 0; i < 7; ++i) worst = std::max(worst, rel(ad[i], fd[i], 1e-3 * scale));
        ++n;
    }
    std::printf("rhs %d %.3e\n", n, worst);
}

static void step_fwd(const RayK& K, const double x[4], double h, double yn[4]) {
    double ynb[4] = {0, 0, 0, 0}, kb[7][4] = {}, xb[4] = {}, cb[3] = {};
    (void)ynb; (void)kb; (void)xb; (void)cb;
    const double A[6][5] = {{0, 0, 0, 0, 0}, {A21, 0, 0, 0, 0}, {A31, A32, 0, 0, 0}, {A41, A42, A43, 0, 0},
                            {A51, A52, A53, A54, 0}, {A61, A62, A63, A64, A65}};
    const double B[6] = {B1, 0.0, B3, B4, B5, B6};
    double k[6][4], X[4];
    for (int i = 0; i < 6; ++i) {
        for (int c = 0; c < 4; ++c) {
            double acc = 0.0;
            for (int j = 0; j < i; ++j) acc += A[i][j] * k[j][c];
            X[c] = x[c] + h * acc;
        }
        rhs_x(K, X, k[i]);
    }
    for (int c = 0; c < 4; ++c) {
        double acc = 0.0;
        for (int j = 0; j < 6; ++j) acc += B[j] * k[j][c];
        yn[c] = x[c] + h * acc;
    }
}

static void check_step(std::mt19937_64& rng) {
    std::uniform_real_distribution<double> U(0.0, 1.0);
    double worst = 0.0;
    int n = 0;
    for (int t = 0; t < 100; ++t) {
        const double a = 0.998 * U(rng), pt = -1.0, pph = -5.0 + 10.0 * U(rng);
        const double x[4] = {4.0 + 40.0 * U(rng), 0.5 + 2.0 * U(rng), -1.0 + 2.0 * U(rng), -3.0 + 6.0 * U(rng)};
        const double h = -(0.05 + 0.5 * U(rng));
        const double w[4] = {U(rng) - 0.5, U(rng) - 0.5, U(rng) - 0.5, U(rng) - 0.5};
        double ynb[4] = {w[0], w[1], w[2], w[3]}, kb[7][4] = {}, xb[4] = {}, cb[3] = {};
        step_vjp(ray_k(a, a * a, pt, pph), x, h, ynb, kb, false, xb, cb);
        auto J = [&](const double* xx, double aa, double ptt, double pphh) {
            double yn[4];
            step_fwd(ray_k(aa, aa * aa, ptt, pphh), xx, h, yn);
            return w[0] * yn[0] + w[1] * yn[1] + w[2] * yn[2] + w[3] * yn[3];
        };
        double fd[7];
        for (int i = 0; i < 4; ++i) {
            double xp[4] = {x[0], x[1], x[2], x[3]}, xm[4] = {x[0], x[1], x[2], x[3]};
            const double d = 1e-6 * std::max(1.0, std::fabs(x[i]));
            xp[i] += d; xm[i] -= d;
            fd[i] = (J(xp, a, pt, pph) - J(xm, a, pt, pph)) / (2 * d);
        }
        const double d = 1e-6;
        fd[4] = (J(x, a + d, pt, pph) - J(x, a - d, pt, pph)) / (2 * d);
        fd[5] = (J(x, a, pt + d, pph) - J(x, a, pt - d, pph)) / (2 * d);
        fd[6] = (J(x, a, pt, pph + d) - J(x, a, pt, pph - d)) / (2 * d);
        const double ad[7] = {xb[0], xb[1], xb[2], xb[3], cb[0], cb[1], cb[2]};
        double scale = 0.0;
        for (int i = 0; i < 7; ++i) scale = std::max(scale, std::fabs(fd[i]));
        for (int i = 0; i < 7; ++i) worst = std::max(worst, rel(ad[i], fd[i], 1e-3 * scale));
        ++n;
    }
    std::printf("step %d %.3e\n", n, worst);
}

// r at the equatorial crossing of the dense output of one frozen step.
static bool cross_r(const RayK& K, const double x[4], double h, double* r_hit, double* s_out) {
    Ray s{};
    s.y[1] = x[0]; s.y[2] = x[1]; s.y[4] = x[2]; s.y[5] = x[3];
    s.pt = K.pt; s.pph = K.pph; s.h = h;
    Rhs f;
    kerr_rhs(K, x[0], x[1], x[2], x[3], f);
    for (int c = 0; c < 6; ++c) s.k1[c] = f.d[c];
    Att A;
    attempt(K, s, 1e-9, 1e-11, A);
    if (!(f.cs * A.g7.cs < 0.0)) return false;
    double cr[5], ct[5];
    dense_coeffs(s, A, 1, cr);
    dense_coeffs(s, A, 2, ct);
    const double se = root_theta(ct[0], ct[1], ct[2], ct[3], ct[4], f.cs, s.y[2], A.yn[2]);
    *r_hit = dense_comp(cr[0], cr[1], cr[2], cr[3], cr[4], se);
    *s_out = se;
    return true;
}

static void check_terminal(std::mt19937_64& rng) {
    std::uniform_real_distribution<double> U(0.0, 1.0);
    double worst = 0.0;
    int n = 0;
    for (int t = 0; t < 400 && n < 100; ++t) {
        const double a = 0.998 * U(rng), pt = -1.0, pph = -4.0 + 8.0 * U(rng);
        // start just above the equator (theta < pi/2) moving towards larger theta
        const double x[4] = {6.0 + 6.0 * U(rng), 1.5707963267948966 - 0.001 - 0.004 * U(rng), -0.5 + U(rng),
                             1.0 + 3.0 * U(rng)};
        const double h = 0.5 + 1.5 * U(rng);
        const RayK K = ray_k(a, a * a, pt, pph);
        double r0, s0;
        if (!cross_r(K, x, h, &r0, &s0)) continue;
        double xb[4] = {}, cb[3] = {};
        const Frame F0{};   // no jet: the seed is r(s*) alone
        terminal_vjp(F0, K, x, h, s0, 1.0, false, 0.0, xb, cb);
        double fd[7];
        bool ok = true;
        auto R = [&](const double* xx, double aa, double ptt, double pphh, double* out) {
            double s_;
            ok = ok && cross_r(ray_k(aa, aa * aa, ptt, pphh), xx, h, out, &s_);
        };
        for (int i = 0; i < 4; ++i) {
            double xp[4] = {x[0], x[1], x[2], x[3]}, xm[4] = {x[0], x[1], x[2], x[3]};
            const double d = 1e-6 * std::max(1.0, std::fabs(x[i]));
            xp[i] += d; xm[i] -= d;
            double rp, rm;
            R(xp, a, pt, pph, &rp);
            R(xm, a, pt, pph, &rm);
            fd[i] = (rp - rm) / (2 * d);
        }
        const double d = 1e-6;
        double rp, rm;
        R(x, a + d, pt, pph, &rp); R(x, a - d, pt, pph, &rm); fd[4] = (rp - rm) / (2 * d);
        R(x, a, pt + d, pph, &rp); R(x, a, pt - d, pph, &rm); fd[5] = (rp - rm) / (2 * d);
        R(x, a, pt, pph + d, &rp); R(x, a, pt, pph - d, &rm); fd[6] = (rp - rm) / (2 * d);
        if (!ok) continue;
        const double ad[7] = {xb[0], xb[1], xb[2], xb[3], cb[0], cb[1], cb[2]};
        double scale = 0.0;
        for (int i = 0; i < 7; ++i) scale = std::max(scale, std::fabs(fd[i]));
        for (int i = 0; i < 7; ++i) worst = std::max(worst, rel(ad[i], fd[i], 1e-3 * scale));
        ++n;
    }
    std::printf("term %d %.3e\n", n, worst);
}

// The differentiable parameters (spin, theta_cam, eddington, r_cam) of a
// scene; a scene with the conical jet is re-normalised for a new spin and
// accretion rate, as the jet's power depends on both.
static bhx_scene with_param(const bhx_scene& sc, int p, double v) {
    bhx_scene s = sc;
    double* f = p == 0 ? &s.spin : p == 1 ? &s.cam_theta : p == 2 ? &s.eddington : &s.cam_r;
    *f = v;
    if (s.jet_enabled && s.jet_model == 1) bhx_setup_conical_jet(&s, 50.0, 0.02);
    return s;
}

// Full adaptive forward of one ray: its intensity (disc plus jet).
static double ray_I(const bhx_scene& sc, int K, int64_t ray, int* cls, int* n_acc) {
    Frame F;
    FrameCam G;
    make_frame(sc, K, F, G);
    std::vector<double> ck(static_cast<size_t>(F.max_steps / K + 1) * kCk);
    FlatCk store{ck.data()};
    Ray s;
    fwd_init(F, G, ray, s, store);
    double yo[6], se = 0.0;
    int term = -1;
    while (term < 0) term = fwd_step(F, s, store, yo, &se);
    *cls = term;
    *n_acc = s.nsteps;
    return (term == BHX_DISC ? disc_intensity(F, yo[1], s.pt, s.pph) : 0.0) + F.kappa * s.jet;
}

// Whole rays of scene sc (every `stride`-th ray with I > 0, up to n_max):
// the adjoint's d I / d(spin, theta_cam, eddington, r_cam) against central
// differences (step 1e-6 max(1, |p|)) of the full adaptive forward, over rays
// whose class does not change.
static void check_ray(const char* name, const bhx_scene& sc, int K, int64_t stride, int n_max) {
    Frame F;
    FrameCam G;
    if (!make_frame(sc, K, F, G)) {
        std::printf("%s 0 1 mismatch 0\n", name);
        return;
    }
    std::vector<double> ck(static_cast<size_t>(F.max_steps / K + 1) * kCk), seg(static_cast<size_t>(K + 1) * kSeg);
    FlatCk store{ck.data()};
    double worst = 0.0;
    int n = 0, mism = 0, n_jet = 0;
    for (int64_t ray = 0; ray < (int64_t)sc.width * sc.height && n < n_max; ray += stride) {
        Ray s;
        fwd_init(F, G, ray, s, store);
        double yo[6], se = 0.0;
        int term = -1;
        while (term < 0) term = fwd_step(F, s, store, yo, &se);
        const double I0 = (term == BHX_DISC ? disc_intensity(F, yo[1], s.pt, s.pph) : 0.0) + F.kappa * s.jet;
        if (!(I0 > 0.0)) continue;
        double grad[4] = {0, 0, 0, 0};
        mism += bwd_ray(F, G, ray, s.nsteps, term, 1.0, store, seg.data(), 1, grad);
        double fd[4];
        bool ok = true;
        for (int p = 0; p < 4; ++p) {
            const double v = p == 0 ? sc.spin : p == 1 ? sc.cam_theta : p == 2 ? sc.eddington : sc.cam_r;
            const double d = 1e-6 * std::max(1.0, std::fabs(v));   // theta_cam can be a few degrees
            int ca, cb_, na, nb;
            const double Ia = ray_I(with_param(sc, p, v + d), K, ray, &ca, &na);
            const double Ib = ray_I(with_param(sc, p, v - d), K, ray, &cb_, &nb);
            if (ca != term || cb_ != term) ok = false;
            fd[p] = (Ia - Ib) / (2 * d);
        }
        if (!ok) continue;
        double e = 0.0;
        for (int p = 0; p < 4; ++p) e = std::max(e, rel(grad[p], fd[p], 1e-30));
        if (e > 1e-3)
            std::printf("  ray %ld cls %d steps %d jet/I %.3e grad %.6e %.6e %.6e %.6e fd %.6e %.6e %.6e %.6e\n",
                        (long)ray, term, s.nsteps, F.kappa * s.jet / I0, grad[0], grad[1], grad[2], grad[3], fd[0],
                        fd[1], fd[2], fd[3]);
        worst = std::max(worst, e);
        if (s.jet > 0.0) ++n_jet;
        ++n;
    }
    std::printf("%s %d %.3e mismatch %d jet %d\n", name, n, worst, mism, n_jet);
}

// A record of every accepted step (K = 1, read back without replaying) and
// segments of K accepted steps replayed from checkpoints give the same
// adjoint bit for bit: the same rays (class, steps, intensity) and, per lit
// ray, the same four derivatives.  Prints the rays compared and how many
// differ.
static void check_log_equals_replay(const char* name, const bhx_scene& sc, int K, int64_t stride, int n_max) {
    Frame F1, FK;
    FrameCam G1, GK;
    if (!make_frame(sc, 1, F1, G1) || !make_frame(sc, K, FK, GK)) {
        std::printf("%s 0 1 mismatch 0\n", name);
        return;
    }
    std::vector<double> ck1(static_cast<size_t>(F1.max_steps + 1) * kCk);
    std::vector<double> ckK(static_cast<size_t>(FK.max_steps / K + 1) * kCk), seg(static_cast<size_t>(K + 1) * kSeg);
    FlatCk st1{ck1.data()}, stK{ckK.data()};
    int n = 0, differ = 0, mism = 0;
    for (int64_t ray = 0; ray < (int64_t)sc.width * sc.height && n < n_max; ray += stride) {
        Ray a, b;
        double ya[6], yb[6], sa = 0.0, sb = 0.0;
        fwd_init(F1, G1, ray, a, st1);
        fwd_init(FK, GK, ray, b, stK);
        int ta = -1, tb = -1;
        while (ta < 0) ta = fwd_step(F1, a, st1, ya, &sa);
        while (tb < 0) tb = fwd_step(FK, b, stK, yb, &sb);
        const double Ia = (ta == BHX_DISC ? disc_intensity(F1, ya[1], a.pt, a.pph) : 0.0) + F1.kappa * a.jet;
        const double Ib = (tb == BHX_DISC ? disc_intensity(FK, yb[1], b.pt, b.pph) : 0.0) + FK.kappa * b.jet;
        if (ta != tb || a.nsteps != b.nsteps || std::memcmp(&Ia, &Ib, sizeof Ia) != 0) {
            ++differ;
            continue;
        }
        if (!(Ia > 0.0)) continue;
        double ga[4] = {0, 0, 0, 0}, gb[4] = {0, 0, 0, 0};
        mism += bwd_ray(F1, G1, ray, a.nsteps, ta, 1.0, st1, seg.data(), 1, ga);
        mism += bwd_ray(FK, GK, ray, b.nsteps, tb, 1.0, stK, seg.data(), 1, gb);
        if (std::memcmp(ga, gb, sizeof ga) != 0) ++differ;
        ++n;
    }
    std::printf("%s %d %d mismatch %d\n", name, n, differ, mism);
}

// The differentiation scene, sampled: a = 0.7, i = 45 deg, r = 500 M, +-20 M.
static bhx_scene disc_scene(double rtol, double atol) {
    bhx_scene sc{};
    sc.M_kg = 10.0 * 1.98847e30; sc.spin = 0.7;
    sc.cam_r = 500.0; sc.cam_theta = 45.0 * 3.14159265358979323846 / 180.0; sc.cam_fov = 2.0 * std::atan(20.0 / 500.0);
    sc.width = 64; sc.height = 64; sc.sqrt_spp = 1; sc.max_steps = 20000;
    sc.disc_enabled = 1; sc.disc_r_out = 20.0; sc.eddington = 0.1; sc.disc_sense = 1;
    sc.rtol = rtol; sc.atol = atol; sc.r_escape = 3.0 * 500.0 + 100.0; sc.seed = 20240921ull;
    return sc;
}

// A quasar frame zoomed on the jets' base: a = 0.998, 1000 M disc at 0.5
// Eddington, conical jet (5 deg, Gamma 10, 6-500 M), camera at 2500 M and
// inclination incl_deg over +-half M.
static bhx_scene jet_scene(double incl_deg, double half, double rtol, double atol) {
    bhx_scene sc{};
    sc.M_kg = 10.0 * 1.98847e30; sc.spin = 0.998;
    sc.cam_r = 2500.0; sc.cam_theta = incl_deg * 3.14159265358979323846 / 180.0;
    sc.cam_fov = 2.0 * std::atan(half / 2500.0);
    sc.width = 48; sc.height = 48; sc.sqrt_spp = 1; sc.max_steps = 20000;
    sc.disc_enabled = 1; sc.disc_r_out = 1000.0; sc.eddington = 0.5; sc.disc_sense = 1;
    sc.jet_enabled = 1; sc.jet_model = 1;
    sc.jet_cone_angle = 5.0 * 3.14159265358979323846 / 180.0; sc.jet_cone_r0 = 6.0; sc.jet_cone_r1 = 500.0;
    sc.jet_gamma_max = 10.0; sc.jet_alpha = 0.7;
    bhx_setup_conical_jet(&sc, 50.0, 0.02);
    sc.rtol = rtol; sc.atol = atol; sc.r_escape = 3.0 * 2500.0 + 100.0; sc.seed = 20240921ull;
    return sc;
}

int main() {
    std::mt19937_64 rng(12345);
    check_rhs(rng);
    check_step(rng);
    check_terminal(rng);
    check_ray("ray_rtol1e-09_K16", disc_scene(1e-9, 1e-11), 16, 37, 40);
    check_ray("ray_rtol1e-12_K7", disc_scene(1e-12, 1e-14), 7, 37, 40);
    check_ray("jet_i10_K16", jet_scene(10.0, 120.0, 1e-10, 1e-12), 16, 7, 80);
    check_ray("jet_i3_K7", jet_scene(3.0, 120.0, 1e-12, 1e-14), 7, 13, 80);
    check_ray("ray_rtol1e-09_K1", disc_scene(1e-9, 1e-11), 1, 37, 40);
    check_ray("jet_i10_K1", jet_scene(10.0, 120.0, 1e-10, 1e-12), 1, 7, 80);
    check_log_equals_replay("log_vs_replay_disc", disc_scene(1e-10, 1e-12), 16, 29, 60);
    check_log_equals_replay("log_vs_replay_jet", jet_scene(10.0, 120.0, 1e-10, 1e-12), 7, 11, 60);
    return 0;
}
