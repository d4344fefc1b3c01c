// surf.hip -- SURF descriptors of given keypoints on gfx950 (the reference computes them with
// OpenCV's cv::cuda::SURF_CUDA on the BIMP keypoints, bimp.cpp:32-59 and :188-199, with
// useProvidedKeypoints, upright off, 64 dimensions).
//
// Restated from the published method (Bay, Ess, Tuytelaars, Van Gool, "Speeded-Up Robust
// Features (SURF)", CVIU 110(3), 2008), not from OpenCV's code; byte parity with OpenCV is
// not pinned.  For a keypoint (x, y, size) the scale is s = 1.2 * size / 9:
//   - Haar wavelet responses come from the integral image: for a half side h, dx is the sum of
//     the h x 2h box right of the sample minus the one left of it, dy the box below minus the
//     box above (pixels outside the frame count as 0);
//   - orientation: samples at (round(x + i*s), round(y + j*s)) for integers i^2 + j^2 <= 36,
//     Haar half side max(1, round(2s)), Gaussian weight exp(-(i^2 + j^2) / 8); the weighted
//     response vectors whose angle falls in a window of pi/3 are summed for 72 window starts
//     (5 degree steps), and the longest sum gives the angle (the first on ties);
//   - descriptor: a 20s x 20s square rotated to that angle, 20 x 20 samples at offsets
//     ((u + 0.5)s, (v + 0.5)s), u, v in [-10, 10), Haar half side max(1, round(s)), responses
//     rotated into the keypoint's frame and weighted by exp(-((u+0.5)^2 + (v+0.5)^2) / 21.78)
//     (sigma 3.3s); each of the 4 x 4 subregions (5 x 5 samples) sums dx, dy, |dx|, |dy|; the
//     64 values are scaled to unit length.
// Rounding of sample positions is half away from zero, evaluated in FP32.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/bimp_hip.h"

namespace {

constexpr int kOriSamples = 113;  // integer (i, j) with i^2 + j^2 <= 36
constexpr int kWindows = 72;
constexpr float kPi = 3.14159265358979323846f;

// Integral image, rows then columns: II has (H+1) x (W+1) uint32 entries, II[0][*] = II[*][0] = 0.
__global__ void k_integral_rows(const uint8_t* __restrict__ img, size_t pitch, int W, int H,
                                unsigned* __restrict__ ii) {
    // one wavefront per row: a running sum over chunks of 64 pixels (wavefront prefix sums)
    const int y = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (y > H) return;
    unsigned* out = ii + (size_t)y * (W + 1);
    if (y == 0) {
        for (int x = lane; x <= W; x += 64) out[x] = 0;
        return;
    }
    const uint8_t* row = img + (size_t)(y - 1) * pitch;
    unsigned carry = 0;
    if (lane == 0) out[0] = 0;
    for (int x0 = 0; x0 < W; x0 += 64) {
        const int x = x0 + lane;
        unsigned v = x < W ? row[x] : 0u;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned t = __shfl_up(v, d, 64);
            if (lane >= d) v += t;
        }
        if (x < W) out[x + 1] = carry + v;
        carry += __shfl(v, 63, 64);
    }
}

// Column prefix sums: one thread per column, eight rows loaded ahead of their sums.
__global__ void k_integral_cols(unsigned* __restrict__ ii, int W, int H) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x > W) return;
    const size_t ld = (size_t)W + 1;
    unsigned acc = 0;
    int y = 1;
    for (; y + 7 <= H; y += 8) {
        unsigned v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = ii[(size_t)(y + j) * ld + x];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            acc += v[j];
            ii[(size_t)(y + j) * ld + x] = acc;
        }
    }
    for (; y <= H; ++y) {
        acc += ii[(size_t)y * ld + x];
        ii[(size_t)y * ld + x] = acc;
    }
}

__device__ __forceinline__ int round_away(float v) { return (int)roundf(v); }

// Sum of pixels in columns [c0, c1) x rows [r0, r1), clipped to the frame.
__device__ __forceinline__ int box(const unsigned* __restrict__ ii, int W, int H, int c0, int r0,
                                   int c1, int r1) {
    c0 = min(max(c0, 0), W);
    c1 = min(max(c1, 0), W);
    r0 = min(max(r0, 0), H);
    r1 = min(max(r1, 0), H);
    if (c1 <= c0 || r1 <= r0) return 0;
    const size_t ld = (size_t)W + 1;
    return (int)(ii[(size_t)r1 * ld + c1] - ii[(size_t)r0 * ld + c1] - ii[(size_t)r1 * ld + c0] +
                 ii[(size_t)r0 * ld + c0]);
}

__device__ __forceinline__ void haar(const unsigned* __restrict__ ii, int W, int H, int px, int py,
                                     int h, float* dx, float* dy) {
    *dx = (float)(box(ii, W, H, px, py - h, px + h, py + h) - box(ii, W, H, px - h, py - h, px, py + h));
    *dy = (float)(box(ii, W, H, px - h, py, px + h, py + h) - box(ii, W, H, px - h, py - h, px + h, py));
}

// One wavefront per keypoint.
__global__ void __launch_bounds__(64) k_surf(const unsigned* __restrict__ ii, int W, int H,
                                             const float* __restrict__ kp, int ld, int n,
                                             int upright, float* __restrict__ desc, int desc_ld,
                                             float* __restrict__ angle_out) {
    __shared__ float rx[kOriSamples], ry[kOriSamples], ra[kOriSamples];
    __shared__ float win[kWindows][2];
    __shared__ float smp[400][2];
    const int k = blockIdx.x, lane = threadIdx.x;
    if (k >= n) return;
    const float x = kp[k], y = kp[(size_t)ld + k], size = kp[4 * (size_t)ld + k];
    const float s = 1.2f * size / 9.f;
    float theta = 0.f;
    if (!upright) {
        const int ho = max(1, round_away(2.f * s));
        for (int t = lane; t < kOriSamples; t += 64) {
            // t-th integer point of the disc, rows j = -6..6 in order, columns i ascending
            int j = -6, cnt = t;
            for (; j <= 6; ++j) {
                const int w = (int)floorf(sqrtf((float)(36 - j * j)));
                if (cnt < 2 * w + 1) break;
                cnt -= 2 * w + 1;
            }
            const int w = (int)floorf(sqrtf((float)(36 - j * j)));
            const int i = cnt - w;
            float dx, dy;
            haar(ii, W, H, round_away(x + i * s), round_away(y + j * s), ho, &dx, &dy);
            const float g = expf(-(float)(i * i + j * j) / 8.f);
            rx[t] = g * dx;
            ry[t] = g * dy;
            float a = atan2f(dy, dx);
            if (a < 0) a += 2.f * kPi;
            ra[t] = a;
        }
        __syncthreads();
        for (int wv = lane; wv < kWindows; wv += 64) {
            const float a0 = wv * (2.f * kPi / kWindows);
            float sx = 0.f, sy = 0.f;
            for (int t = 0; t < kOriSamples; ++t) {
                float d = ra[t] - a0;
                if (d < 0) d += 2.f * kPi;
                if (d < kPi / 3.f) {
                    sx += rx[t];
                    sy += ry[t];
                }
            }
            win[wv][0] = sx;
            win[wv][1] = sy;
        }
        __syncthreads();
        float best = -1.f;
        for (int wv = 0; wv < kWindows; ++wv) {
            const float m = win[wv][0] * win[wv][0] + win[wv][1] * win[wv][1];
            if (m > best) {
                best = m;
                theta = atan2f(win[wv][1], win[wv][0]);
            }
        }
    }
    const float co = cosf(theta), si = sinf(theta);
    const int hd = max(1, round_away(s));
    for (int t = lane; t < 400; t += 64) {
        const int u = t % 20 - 10, v = t / 20 - 10;
        const float ou = (u + 0.5f) * s, ov = (v + 0.5f) * s;
        const int px = round_away(x + co * ou - si * ov);
        const int py = round_away(y + si * ou + co * ov);
        float dx, dy;
        haar(ii, W, H, px, py, hd, &dx, &dy);
        const float g = expf(-((u + 0.5f) * (u + 0.5f) + (v + 0.5f) * (v + 0.5f)) / (2.f * 3.3f * 3.3f));
        smp[t][0] = g * (dx * co + dy * si);
        smp[t][1] = g * (-dx * si + dy * co);
    }
    __syncthreads();
    // lane l holds feature l % 4 (dx, dy, |dx|, |dy|) of subregion l / 4 (4 x 4, row-major),
    // summed over its 5 x 5 samples in row-major order
    const int r = lane / 4, q = lane % 4;
    const int u0 = (r % 4) * 5, v0 = (r / 4) * 5;
    float val = 0.f;
    for (int vv = 0; vv < 5; ++vv)
        for (int uu = 0; uu < 5; ++uu) {
            const float f = smp[(v0 + vv) * 20 + u0 + uu][q & 1];
            val += q < 2 ? f : fabsf(f);
        }
    float sq = val * val;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) sq += __shfl_xor(sq, d, 64);
    const float inv = sq > 0.f ? 1.f / sqrtf(sq) : 0.f;
    desc[(size_t)k * desc_ld + lane] = val * inv;
    if (angle_out && lane == 0) angle_out[k] = theta;
}

}  // namespace

extern "C" {

int bimp_integral_image(const uint8_t* d_img, size_t pitch, int width, int height,
                        unsigned* d_integral, bimp_stream_t stream) {
    if (!d_img || !d_integral || width <= 0 || height <= 0 || pitch < (size_t)width ||
        (size_t)width * height * 255u > 0x7fffffffu)
        return BIMP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    k_integral_rows<<<(height + 1 + 3) / 4, 256, 0, s>>>(d_img, pitch, width, height, d_integral);
    k_integral_cols<<<(width + 1 + 255) / 256, 256, 0, s>>>(d_integral, width, height);
    return hipGetLastError() == hipSuccess ? BIMP_OK : BIMP_E_DEVICE;
}

int bimp_surf_descriptors(const unsigned* d_integral, int width, int height, const float* d_kp,
                          int ld, int n, int upright, float* d_desc, int desc_ld, float* d_angle,
                          bimp_stream_t stream) {
    if (n < 0 || (n > 0 && (!d_integral || !d_kp || !d_desc)) || ld < n || desc_ld < 64 ||
        width <= 0 || height <= 0)
        return BIMP_E_ARG;
    if (n == 0) return BIMP_OK;
    k_surf<<<n, 64, 0, (hipStream_t)stream>>>(d_integral, width, height, d_kp, ld, n, upright,
                                              d_desc, desc_ld, d_angle);
    return hipGetLastError() == hipSuccess ? BIMP_OK : BIMP_E_DEVICE;
}

}  // extern "C"
