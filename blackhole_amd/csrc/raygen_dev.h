// raygen_dev.h - the camera ray of one sample, shared by the tracer's ray
// generation and the adjoint (host-callable for the adjoint's self-checks).
#pragma once

#include "kerr_dev.h"
#include "params.h"

namespace bhx {

// Unit direction in the camera's ZAMO frame of sample k of pixel (px, py):
// the pixel jitter of render.cpp:440-446 (draws 2k+1, 2k+2 of the pixel's
// splitmix64 stream), the image-plane mapping of render.cpp:447-448 and
// Camera::ray_direction (scene.h:356-384).  Independent of spin and of the
// camera's position, so the adjoint treats it as a constant of the ray.
__host__ __device__ __forceinline__ void camera_dir(const RaygenArgs& P, int px, int py, int k, double d[3]) {
    const uint64_t sd = P.seed ^ ((uint64_t)py << 32) ^ ((uint64_t)px * 0x9E3779B97F4A7C15ull);
    const int sx = k % P.S, sy = k / P.S;
    const double jx = (sx + splitmix_u01(sd, 2 * (uint64_t)k + 1)) * P.inv_S;
    const double jy = (sy + splitmix_u01(sd, 2 * (uint64_t)k + 2)) * P.inv_S;
    const double u = (2.0 * (px + jx) / P.W - 1.0);
    const double v = (1.0 - 2.0 * (py + jy) / P.H);
    double cx = u * P.tanf, cy = v * P.tanf / P.aspect;
    const double cz = 1.0;
    if (P.has_roll) {
        const double xr = P.c_roll * cx - P.s_roll * cy;
        const double yr = P.s_roll * cx + P.c_roll * cy;
        cx = xr;
        cy = yr;
    }
    const double y2 = P.c_pitch * cy - P.s_pitch * cz;
    const double z2 = P.s_pitch * cy + P.c_pitch * cz;
    const double x2 = P.c_yaw * cx + P.s_yaw * z2;
    const double z3 = -P.s_yaw * cx + P.c_yaw * z2;
    const double v0 = -z3, v1 = -y2, v2 = x2;
    const double n = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    d[0] = v0 / n;
    d[1] = v1 / n;
    d[2] = v2 / n;
}

// make_photon (geodesic.h:215-231), past-directed: p^mu = -(e0 - d.e) in the
// ZAMO tetrad, lowered with the camera's metric.  T is double, or a dual
// number carrying derivatives of the frame constants.
template <class T>
__host__ __device__ __forceinline__ void camera_photon_t(const T& e00, const T& e03, const T& e11, const T& e22,
                                                         const T& e33, const T& gl_tt, const T& gl_tp,
                                                         const T& gl_rr, const T& gl_hh, const T& gl_pp,
                                                         const double d[3], T& pt, T& pr, T& pth, T& pph) {
    const T pu0 = -e00;
    const T pu1 = d[0] * e11;
    const T pu2 = d[1] * e22;
    const T pu3 = -(e03 - d[2] * e33);
    pt = gl_tt * pu0 + gl_tp * pu3;
    pr = gl_rr * pu1;
    pth = gl_hh * pu2;
    pph = gl_tp * pu0 + gl_pp * pu3;
}

__host__ __device__ __forceinline__ void camera_photon(const RaygenArgs& P, const double d[3], double& pt,
                                                       double& pr, double& pth, double& pph) {
    camera_photon_t<double>(P.e00, P.e03, P.e11, P.e22, P.e33, P.gl_tt, P.gl_tp, P.gl_rr, P.gl_hh, P.gl_pp,
                            d, pt, pr, pth, pph);
}

}  // namespace bhx
