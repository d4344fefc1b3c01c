// frame.hip -- getGPUImage's frame preparation on the device (cuda_bimp.cu:156-177): the
// minimum-size upscale (cv::resize, bilinear) and, for colour frames, BGR -> gray.
//
// The reference upscales a frame whose smaller relevant side is below MIN_SIZE = 350 with
// cv::resize(in, out, Size(), sf, sf) (INTER_LINEAR) on the host, then uploads it; a
// Context's CudaStuff stays sized from the original frame (bimp.cpp:8-10), so detection reads
// the top-left width x height of the upscaled frame.  OpenCV 3.3.1's 8-bit bilinear resize is
// restated here from its published definition: source coordinate (x + 0.5) / sf - 0.5 per
// axis, clamped to the frame, 11-bit fixed-point weights (each rounded to nearest from the
// float fraction), the two-axis sum rounded once, (sum + 2^21) >> 22.  OpenCV's vectorised
// paths round the intermediate differently in rare cases, so byte parity with it is not
// pinned.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/bimp_hip.h"

namespace {

constexpr int kMinSize = 350;  // MIN_SIZE, cuda_bimp.cu:157
constexpr int kCoefBits = 11;
constexpr int kCoefScale = 1 << kCoefBits;

struct Tap {
    int i0, i1;  // source indices (clamped)
    int w0, w1;  // fixed-point weights
};

// Source taps of output index d along an axis of n source samples, scale = 1 / sf.
__device__ __forceinline__ Tap axis_tap(int d, double scale, int n) {
    float f = (float)((d + 0.5) * scale - 0.5);
    int s = (int)floorf(f);
    f -= (float)s;
    if (s < 0) {
        f = 0.f;
        s = 0;
    }
    if (s >= n - 1) {
        f = 0.f;
        s = n - 1;
    }
    Tap t;
    t.i0 = s;
    t.i1 = s + 1 < n ? s + 1 : n - 1;
    t.w0 = (int)rintf((1.f - f) * kCoefScale);
    t.w1 = (int)rintf(f * kCoefScale);
    return t;
}

__device__ __forceinline__ uint8_t lerp2(const uint8_t* r0, const uint8_t* r1, const Tap& tx,
                                         const Tap& ty, int cn, int c) {
    const int h0 = r0[tx.i0 * cn + c] * tx.w0 + r0[tx.i1 * cn + c] * tx.w1;
    const int h1 = r1[tx.i0 * cn + c] * tx.w0 + r1[tx.i1 * cn + c] * tx.w1;
    const int v = (ty.w0 * h0 + ty.w1 * h1 + (1 << (2 * kCoefBits - 1))) >> (2 * kCoefBits);
    return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
}

// OpenCV's documented fixed-point BGR2GRAY weights (B 1868, G 9617, R 4899, 14 bits).
__device__ __forceinline__ uint8_t gray_of(unsigned b, unsigned g, unsigned r) {
    return (uint8_t)((1868u * b + 9617u * g + 4899u * r + (1u << 13)) >> 14);
}

// One thread per output byte (pixel x channel); rows by blockIdx.y.
__global__ void k_resize_linear(const uint8_t* __restrict__ src, size_t spitch, int sw, int sh,
                                int cn, uint8_t* __restrict__ dst, size_t dpitch, int dw, int dh,
                                double sx, double sy) {
    const int xc = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (xc >= dw * cn || y >= dh) return;
    const int x = xc / cn, c = xc - x * cn;
    const Tap tx = axis_tap(x, sx, sw), ty = axis_tap(y, sy, sh);
    dst[(size_t)y * dpitch + xc] =
        lerp2(src + (size_t)ty.i0 * spitch, src + (size_t)ty.i1 * spitch, tx, ty, cn, c);
}

// Resized BGR pixel (each channel rounded to a byte, as the resized image holds it), then
// converted to gray; sx = sy = 0 skips the resize.
__global__ void k_resize_gray(const uint8_t* __restrict__ src, size_t spitch, int sw, int sh,
                              uint8_t* __restrict__ dst, size_t dpitch, int dw, int dh,
                              double sx, double sy) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= dw || y >= dh) return;
    uint8_t b, g, r;
    if (sx > 0) {
        const Tap tx = axis_tap(x, sx, sw), ty = axis_tap(y, sy, sh);
        const uint8_t* r0 = src + (size_t)ty.i0 * spitch;
        const uint8_t* r1 = src + (size_t)ty.i1 * spitch;
        b = lerp2(r0, r1, tx, ty, 3, 0);
        g = lerp2(r0, r1, tx, ty, 3, 1);
        r = lerp2(r0, r1, tx, ty, 3, 2);
    } else {
        const uint8_t* p = src + (size_t)y * spitch + 3 * (size_t)x;
        b = p[0];
        g = p[1];
        r = p[2];
    }
    dst[(size_t)y * dpitch + x] = gray_of(b, g, r);
}

}  // namespace

extern "C" {

int bimp_min_size_shape(int width, int height, int* out_w, int* out_h, float* scale) {
    if (width <= 0 || height <= 0 || !out_w || !out_h) return BIMP_E_ARG;
    float sf = 1.f;
    // getGPUImage (cuda_bimp.cu:159-168): rows below 350 with cols >= 350 scale by 350/rows;
    // otherwise cols below 350 scale by 350/cols (the factor is a float)
    if (height < kMinSize && width >= kMinSize)
        sf = (float)(kMinSize / (1.0 * height));
    else if (width < kMinSize)
        sf = (float)(kMinSize / (1.0 * width));
    if (sf == 1.f) {
        *out_w = width;
        *out_h = height;
    } else {
        // cv::resize with dsize = Size(): round(side * factor), half to even
        *out_w = (int)std::nearbyint(width * (double)sf);
        *out_h = (int)std::nearbyint(height * (double)sf);
    }
    if (scale) *scale = sf;
    return BIMP_OK;
}

int bimp_resize_linear_u8(const uint8_t* d_src, size_t src_pitch, int sw, int sh, int channels,
                          uint8_t* d_dst, size_t dst_pitch, int dw, int dh, double fx, double fy,
                          bimp_stream_t stream) {
    if (!d_src || !d_dst || sw <= 0 || sh <= 0 || dw <= 0 || dh <= 0 ||
        (channels != 1 && channels != 3) || src_pitch < (size_t)sw * channels ||
        dst_pitch < (size_t)dw * channels)
        return BIMP_E_ARG;
    // cv::resize maps with 1 / factor when a factor is given, else with src / dst
    const double sx = fx > 0 ? 1.0 / fx : (double)sw / dw;
    const double sy = fy > 0 ? 1.0 / fy : (double)sh / dh;
    const dim3 grid((dw * channels + 255) / 256, dh);
    k_resize_linear<<<grid, 256, 0, (hipStream_t)stream>>>(d_src, src_pitch, sw, sh, channels,
                                                           d_dst, dst_pitch, dw, dh, sx, sy);
    return hipGetLastError() == hipSuccess ? BIMP_OK : BIMP_E_DEVICE;
}

int bimp_get_gpu_image(const uint8_t* d_src, size_t src_pitch, int width, int height,
                       int channels, int grayscale, uint8_t* d_dst, size_t dst_pitch,
                       bimp_stream_t stream) {
    if (!d_src || !d_dst || (channels != 1 && channels != 3) ||
        src_pitch < (size_t)width * channels)
        return BIMP_E_ARG;
    int ow = 0, oh = 0;
    float sf = 1.f;
    int st = bimp_min_size_shape(width, height, &ow, &oh, &sf);
    if (st) return st;
    const bool to_gray = grayscale && channels == 3;
    if (dst_pitch < (size_t)ow * (to_gray ? 1 : channels)) return BIMP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (to_gray) {
        const double sc = sf != 1.f ? 1.0 / (double)sf : 0.0;
        k_resize_gray<<<dim3((ow + 255) / 256, oh), 256, 0, s>>>(d_src, src_pitch, width, height,
                                                                 d_dst, dst_pitch, ow, oh, sc, sc);
        return hipGetLastError() == hipSuccess ? BIMP_OK : BIMP_E_DEVICE;
    }
    if (sf != 1.f)
        return bimp_resize_linear_u8(d_src, src_pitch, width, height, channels, d_dst, dst_pitch,
                                     ow, oh, sf, sf, stream);
    return hipMemcpy2DAsync(d_dst, dst_pitch, d_src, src_pitch, (size_t)width * channels, height,
                            hipMemcpyDeviceToDevice, s) == hipSuccess
               ? BIMP_OK
               : BIMP_E_DEVICE;
}

}  // extern "C"
