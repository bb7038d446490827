// display.hip - veiling glare on the device: img::add_glare
// (src/image.cpp:47-106) behind bhx_add_glare_dev.
//
// The image plus a narrow (sigma = R) and a wide (sigma = 3R) Gaussian of
// itself, weighted 1 - s, 0.7 s, 0.3 s.  Each Gaussian is separable: taps
// -r..r (r = max(1, ceil(3 sigma))), normalised weights, clamp to the edge,
// a horizontal pass then a vertical pass.  One thread per pixel and channel
// sums its taps in the reference's order, and this file is compiled without
// FMA contraction (-ffp-contract=off), so every sum is the reference's
// sequence of IEEE products and additions: the result is bit-identical to the
// CPU's.  The weights are computed on the host with the C library's exp,
// exactly as the reference computes them.
//
// Memory: both passes read the image as 3 doubles per pixel, row-major; a
// wave covers 64 consecutive (pixel, channel) slots of one row, so the
// horizontal taps read contiguous bytes and the vertical taps read whole
// row segments.  The 2r + 1 re-reads of a pixel hit L2 (a 1024² XYZ frame is
// 25 MB); the pass is HBM-bound at about 2 x 25 MB per pass.
#include "context.h"

#include <cmath>
#include <vector>

namespace {

constexpr int kMaxHalfTaps = 1024;   // r of the wide blur, ceil(9 R), at most this

__global__ void __launch_bounds__(256) k_blur_rows(const double* __restrict__ in, double* __restrict__ out,
                                                   int W, int H, const double* __restrict__ k, int r) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t n3 = static_cast<int64_t>(W) * H * 3;
    if (i >= n3) return;
    const int c = static_cast<int>(i % 3);
    const int64_t pix = i / 3;
    const int x = static_cast<int>(pix % W);
    const int64_t row = pix - x;
    double a = 0.0;
    for (int t = -r; t <= r; ++t) {
        const int xs = min(max(x + t, 0), W - 1);
        a = a + k[t + r] * in[3 * (row + xs) + c];
    }
    out[i] = a;
}

__global__ void __launch_bounds__(256) k_blur_cols(const double* __restrict__ in, double* __restrict__ out,
                                                   int W, int H, const double* __restrict__ k, int r) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t n3 = static_cast<int64_t>(W) * H * 3;
    if (i >= n3) return;
    const int c = static_cast<int>(i % 3);
    const int64_t pix = i / 3;
    const int x = static_cast<int>(pix % W);
    const int y = static_cast<int>(pix / W);
    double a = 0.0;
    for (int t = -r; t <= r; ++t) {
        const int ys = min(max(y + t, 0), H - 1);
        a = a + k[t + r] * in[3 * (static_cast<int64_t>(ys) * W + x) + c];
    }
    out[i] = a;
}

__global__ void __launch_bounds__(256) k_glare_mix(double* __restrict__ xyz, const double* __restrict__ nearb,
                                                   const double* __restrict__ farb, int64_t n3, double core,
                                                   double wa, double wb) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    xyz[i] = xyz[i] * core + nearb[i] * wa + farb[i] * wb;
}

// Normalised Gaussian taps of the reference's blur (image.cpp:60-67).
int gauss_taps(double sigma, std::vector<double>& k) {
    const int r = std::max(1, static_cast<int>(std::ceil(3.0 * sigma)));
    k.assign(2 * r + 1, 0.0);
    double sum = 0.0;
    for (int i = -r; i <= r; ++i) {
        k[i + r] = std::exp(-0.5 * (i * i) / (sigma * sigma));
        sum += k[i + r];
    }
    for (double& v : k) v /= sum;
    return r;
}

}  // namespace

extern "C" {

int bhx_add_glare_dev(bhx_context* ctx, double* xyz, int32_t width, int32_t height, double strength,
                      double radius_px, void* stream) {
    if (!ctx) return BHX_E_INVALID;
    if (!xyz || width <= 0 || height <= 0) return bhx_fail(ctx, BHX_E_INVALID, "bad argument");
    if (strength <= 0.0 || radius_px <= 0.0) return BHX_OK;   // image.cpp:48
    if (std::ceil(9.0 * radius_px) > kMaxHalfTaps) return bhx_fail(ctx, BHX_E_TOO_LARGE, "glare radius above 113 px");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t n3 = static_cast<int64_t>(width) * height * 3;
    std::vector<double> kn, kf;
    const int rn = gauss_taps(radius_px, kn), rf = gauss_taps(3.0 * radius_px, kf);
    HIPCHK(ctx->glare_tmp.ensure(sizeof(double) * n3));
    HIPCHK(ctx->glare_near.ensure(sizeof(double) * n3));
    HIPCHK(ctx->glare_far.ensure(sizeof(double) * n3));
    HIPCHK(ctx->glare_taps.ensure(sizeof(double) * (kn.size() + kf.size())));
    // the taps buffer may still be read by an earlier call's kernels
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipMemcpy(ctx->glare_taps.p, kn.data(), sizeof(double) * kn.size(), hipMemcpyHostToDevice));
    double* d_kn = static_cast<double*>(ctx->glare_taps.p);
    double* d_kf = d_kn + kn.size();
    HIPCHK(hipMemcpy(d_kf, kf.data(), sizeof(double) * kf.size(), hipMemcpyHostToDevice));
    double* tmp = static_cast<double*>(ctx->glare_tmp.p);
    double* nearb = static_cast<double*>(ctx->glare_near.p);
    double* farb = static_cast<double*>(ctx->glare_far.p);
    const unsigned blocks = static_cast<unsigned>((n3 + 255) / 256);
    hipLaunchKernelGGL(k_blur_rows, dim3(blocks), dim3(256), 0, st, xyz, tmp, width, height, d_kn, rn);
    hipLaunchKernelGGL(k_blur_cols, dim3(blocks), dim3(256), 0, st, tmp, nearb, width, height, d_kn, rn);
    hipLaunchKernelGGL(k_blur_rows, dim3(blocks), dim3(256), 0, st, xyz, tmp, width, height, d_kf, rf);
    hipLaunchKernelGGL(k_blur_cols, dim3(blocks), dim3(256), 0, st, tmp, farb, width, height, d_kf, rf);
    const double core = 1.0 - strength, wa = 0.7 * strength, wb = 0.3 * strength;
    hipLaunchKernelGGL(k_glare_mix, dim3(blocks), dim3(256), 0, st, xyz, nearb, farb, n3, core, wa, wb);
    HIPCHK(hipGetLastError());
    return BHX_OK;
}

}  // extern "C"
