// features.hip -- the keypoints' downstream consumers on gfx950: the colour histogram
// descriptor and brute-force two-nearest-neighbour matching with the ratio test.
//
// Reference: Context::getDescriptors (bimp/src/bimp.cpp:32-59) appends a 24-float colour
// histogram (addColorFeatures_kernel, bimp/include/detection/cuda_bimp.cuh:8-71) to each
// keypoint's descriptor; matchContexts / downloadMatches (bimp.cpp:188-240) match two
// descriptor sets with OpenCV's CUDA brute-force matcher (L2, k = 2) and keep a match when
// the best distance is below 0.6 x the second best.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/bimp_hip.h"

namespace {

// ------------------------------------------------------------------ colour histogram
// One thread per keypoint (a window holds at most (2p)^2 pixels, p = (int)(log2(size)/2)).
// The histogram arithmetic is the reference's: a channel byte is a signed char v, bin k
// collects 31.875*k < v < 31.875*(k+1), so only bytes 1..127 count (bins 0..3); the counts
// are scaled by 0.5 / (3 * pixels) in double.  Pixels outside the image are skipped (the
// reference's inclusive bounds also read column `width` and row `height`).
__global__ void __launch_bounds__(64) k_color_features(const float* __restrict__ kp, int ld, int n,
                                                       const uint8_t* __restrict__ bgr,
                                                       size_t pitch, int width, int height,
                                                       float* __restrict__ desc, int desc_ld) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = kp[i], y = kp[(size_t)ld + i], size = kp[4 * (size_t)ld + i];
    const int cx = (int)x, cy = (int)y;
    const int p = (int)(log2((double)size) / 2.0);
    unsigned cnt[12] = {0};  // bins 0..3 of B, G, R (bins 4..7 cannot be reached)
    int found = 0;
    for (int iy = cy - p; iy < cy + p; ++iy) {
        if (iy < 0 || iy >= height) continue;
        const uint8_t* row = bgr + (size_t)iy * pitch;
        for (int ix = cx - p; ix < cx + p; ++ix) {
            if (ix < 0 || ix >= width) continue;
            found += 3;
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                const int v = row[3 * ix + ch];
                // 31.875*k < v < 31.875*(k+1) for integer v in 1..127 is k = (8*v) / 255
                if (v >= 1 && v <= 127) cnt[ch * 4 + (8 * v) / 255] += 1;
            }
        }
    }
    float* d = desc + (size_t)i * desc_ld;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const double c = k < 4 ? (double)cnt[ch * 4 + k] : 0.0;
            // (0.5 * c) / found, rounded once to float as the reference does
            d[ch * 8 + k] = (float)((0.5 * c) / (found * 1.0));
        }
    }
}

// ------------------------------------------------------------------ 2-NN matching
// A workgroup of 256 threads owns 64 query rows and walks every train row in tiles of 64.
// Thread (ty, tx) accumulates the 4 x 4 squared distances of queries 4ty.. and trains
// 4tx.. of a tile over the dimensions, staged through LDS 32 at a time (both tiles stored
// dimension-major, so a thread reads one float4 of each per dimension); it keeps the best two
// (distance, index) of its 4 queries over its train columns, and the 16 threads of a query
// row merge theirs at the end.  Ties go to the lower train index.
constexpr int KQ = 64, KT = 64, DC = 32, PADQ = KQ + 4;

struct Best2 {
    float d0, d1;
    int i0, i1;
};

__device__ __forceinline__ bool before(float d, int i, float e, int j) {
    return d < e || (d == e && (unsigned)i < (unsigned)j);
}

__device__ __forceinline__ void insert(Best2& b, float d, int i) {
    if (before(d, i, b.d0, b.i0)) {
        b.d1 = b.d0;
        b.i1 = b.i0;
        b.d0 = d;
        b.i0 = i;
    } else if (before(d, i, b.d1, b.i1)) {
        b.d1 = d;
        b.i1 = i;
    }
}

__global__ void __launch_bounds__(256) k_knn2(const float* __restrict__ q, int nq, int qld,
                                              const float* __restrict__ t, int nt, int tld,
                                              int dim, int* __restrict__ idx2,
                                              float* __restrict__ dist2) {
    __shared__ __align__(16) float qs[DC][PADQ];
    __shared__ __align__(16) float ts[DC][PADQ];
    __shared__ Best2 part[KQ][17];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int q0 = blockIdx.x * KQ;
    Best2 b[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) b[r] = Best2{INFINITY, INFINITY, -1, -1};
    for (int t0 = 0; t0 < nt; t0 += KT) {
        float acc[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[r][c] = 0.f;
        for (int d0 = 0; d0 < dim; d0 += DC) {
            __syncthreads();
            for (int e = threadIdx.x; e < KQ * DC; e += 256) {
                const int r = e / DC, d = e % DC;
                const bool dv = d0 + d < dim;
                qs[d][r] = (dv && q0 + r < nq) ? q[(size_t)(q0 + r) * qld + d0 + d] : 0.f;
                ts[d][r] = (dv && t0 + r < nt) ? t[(size_t)(t0 + r) * tld + d0 + d] : 0.f;
            }
            __syncthreads();
            const int dn = min(DC, dim - d0);
            for (int d = 0; d < dn; ++d) {
                const float4 qa = *reinterpret_cast<const float4*>(&qs[d][4 * ty]);
                const float4 ta = *reinterpret_cast<const float4*>(&ts[d][4 * tx]);
                const float qv[4] = {qa.x, qa.y, qa.z, qa.w};
                const float tv[4] = {ta.x, ta.y, ta.z, ta.w};
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const float df = qv[r] - tv[c];
                        acc[r][c] = fmaf(df, df, acc[r][c]);
                    }
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int ti = t0 + 4 * tx + c;
            if (ti >= nt) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) insert(b[r], acc[r][c], ti);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) part[4 * ty + r][tx] = b[r];
    __syncthreads();
    if (threadIdx.x < KQ && q0 + (int)threadIdx.x < nq) {
        const int r = threadIdx.x;
        Best2 m = part[r][0];
        for (int k = 1; k < 16; ++k) {
            const Best2 o = part[r][k];
            insert(m, o.d0, o.i0);
            insert(m, o.d1, o.i1);
        }
        const size_t qi = (size_t)(q0 + r);
        idx2[2 * qi] = m.i0;
        idx2[2 * qi + 1] = m.i1;
        dist2[2 * qi] = sqrtf(m.d0);
        dist2[2 * qi + 1] = sqrtf(m.d1);
    }
}

}  // namespace

extern "C" {

int bimp_color_features(const float* d_kp, int ld, int n, const uint8_t* d_bgr, size_t bgr_pitch,
                        int width, int height, float* d_desc, int desc_ld, bimp_stream_t stream) {
    if (n < 0 || (n > 0 && (!d_kp || !d_bgr || !d_desc)) || ld < n || desc_ld < BIMP_COLOR_DIMENSIONS ||
        width <= 0 || height <= 0 || bgr_pitch < (size_t)3 * width)
        return BIMP_E_ARG;
    if (n == 0) return BIMP_OK;
    k_color_features<<<(n + 63) / 64, 64, 0, (hipStream_t)stream>>>(d_kp, ld, n, d_bgr, bgr_pitch,
                                                                    width, height, d_desc, desc_ld);
    return hipGetLastError() == hipSuccess ? BIMP_OK : BIMP_E_DEVICE;
}

int bimp_match_knn2(const float* d_query, int nq, int q_ld, const float* d_train, int nt,
                    int t_ld, int dim, int* d_idx2, float* d_dist2, bimp_stream_t stream) {
    if (nq < 0 || nt < 0 || dim <= 0 || q_ld < dim || t_ld < dim ||
        (nq > 0 && (!d_query || !d_idx2 || !d_dist2)) || (nt > 0 && !d_train))
        return BIMP_E_ARG;
    if (nq == 0) return BIMP_OK;
    k_knn2<<<(nq + KQ - 1) / KQ, 256, 0, (hipStream_t)stream>>>(d_query, nq, q_ld, d_train, nt,
                                                                t_ld, dim, d_idx2, d_dist2);
    return hipGetLastError() == hipSuccess ? BIMP_OK : BIMP_E_DEVICE;
}

int bimp_ratio_matches(const int* idx2, const float* dist2, int nq, double ratio,
                       bimp_match* out, int cap, int* count) {
    if (nq < 0 || !count || (nq > 0 && (!idx2 || !dist2)) || cap < 0 || (cap > 0 && !out))
        return BIMP_E_ARG;
    int n = 0;
    for (int i = 0; i < nq; ++i) {
        const float d1 = dist2[2 * (size_t)i], d2 = dist2[2 * (size_t)i + 1];
        // downloadMatches (bimp.cpp:228): dist1 < 0.6 * dist2, the product in double
        if (idx2[2 * (size_t)i] >= 0 && (double)d1 < ratio * (double)d2) {
            if (n < cap) out[n] = bimp_match{i, idx2[2 * (size_t)i], 0, d1};
            ++n;
        }
    }
    *count = n;
    return n > cap ? BIMP_W_OVERFLOW : BIMP_OK;
}

}  // extern "C"
