// direct.hip -- precision form: small-wavelength scales by direct Gabor convolution on MFMA.
//
// For a scale marked "direct", the complex-cell maps C_o are computed without FFTs:
//   e_o + i*o_o = conv(Pd, Ke_o + i*Ko_o)   over the periodic padded image Pd (as the FFT
//                                            path, so wrap-around semantics match)
// (kernels: cuda_bimp.cu:304-331).  The 16 filters are convolved as an implicit GEMM in BF16
// or FP16 with FP32 accumulation (v_mfma_f32_16x16x32_{bf16,f16}): taps scaled by 1/fscale
// and rounded once, image samples entering as their uint8 values (exact in both formats).
// With truncated taps the even filter is the carrier Ke_o + dc_o inside the window.  Either
// way each filter's response is corrected by -dc[f] * B, B the kW x kW box sum of Pd computed
// exactly in integers (k_box_rows / k_box_cols), with dc[f] chosen so that a constant image
// gets the exact full kernel's response: rounding and truncation then leave no DC leak.  The
// 255/65536 input scale and fscale are applied to the FP32 accumulators.
//
// GEMM shape per workgroup (8 wavefronts, output tile 16 rows x 64 columns): M = pixels in
// groups of 16 consecutive columns, N = 16 filters (even/odd x 8 orientations), K = taps
// (u, v) with v padded to a multiple of 8 so that the 8 taps a lane holds lie in one image
// row.  The image tile sits in LDS as four copies shifted by 0..3 samples, which makes every
// lane's 8-sample window two aligned 8-byte reads.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace bimpk {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <class T>
struct Mfma;
template <>
struct Mfma<__bf16> {
    typedef bf16x8 v8;
    typedef bf16x4 v4;
    static __device__ __forceinline__ f32x4 run(v8 a, v8 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<_Float16> {
    typedef f16x8 v8;
    typedef f16x4 v4;
    static __device__ __forceinline__ f32x4 run(v8 a, v8 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

// Row of the periodic padded image (padDataClampToBorder rule, convolutionFFT2D.cuh:56-85)
__device__ __forceinline__ int pad_index(int i, int n, int F, int k) {
    i %= F;
    if (i < 0) i += F;
    return i < n ? i : (i < n + k ? n - 1 : 0);
}

constexpr int kDirTY = 16, kDirTX = 64, kDirWaves = 8;

// Horizontal box sums over kW = 2k+1 columns of the periodic padded uint8 image, for the
// rows y in [-k, Hc + k) (stored at y + k) and columns x < Wc.
__global__ void __launch_bounds__(256) k_box_rows(const uint8_t* __restrict__ img, size_t pitch,
                                                  int H, int W, int FH, int FW, int k, int Hc,
                                                  int Wc, int* __restrict__ hs) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;  // 0 .. Hc + 2k - 1
    if (x >= Wc) return;
    const uint8_t* row = img + (size_t)pad_index(r - k, H, FH, k) * pitch;
    int acc = 0;
    for (int v = -k; v <= k; ++v) acc += row[pad_index(x - v, W, FW, k)];
    hs[(size_t)r * Wc + x] = acc;
}

__global__ void __launch_bounds__(256) k_box_cols(const int* __restrict__ hs, int k, int Hc,
                                                  int Wc, float* __restrict__ box) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= Wc) return;
    int acc = 0;
    for (int u = -k; u <= k; ++u) acc += hs[(size_t)(y - u + k) * Wc + x];
    box[(size_t)y * Wc + x] = (float)acc;
}

template <class T>
__global__ void __launch_bounds__(kDirWaves * 64) k_direct(
        const uint8_t* __restrict__ img, size_t pitch, DirectScale ds,
        const typename Mfma<T>::v8* __restrict__ bfrag, const float* __restrict__ box,
        float* __restrict__ Cm) {
    typedef typename Mfma<T>::v8 v8;
    typedef typename Mfma<T>::v4 v4;
    extern __shared__ unsigned char smem[];
    T* copies = reinterpret_cast<T*>(smem);
    const int R = ds.R, extp = ds.extp;
    const int LH = kDirTY + 2 * R;
    const int LWP = ds.lwp;  // multiple of 4, >= TX + extp + 8
    const int csize = LH * LWP;
    const int x0 = blockIdx.x * kDirTX, y0 = blockIdx.y * kDirTY;
    const int tid = threadIdx.x;

    // stage the tile (rows y0-R.., columns x0-R..) as four shifted copies: tile sample c goes
    // to position c - s of copy s.  Taps past the truncated extent (v >= ext) meet zero
    // filter entries, so every position that can be read is finite (zero past the tile).
    const int wave = tid >> 6, lane = tid & 63;
    const int TW = kDirTX + 2 * R;
    const bool inside = y0 - R >= 0 && y0 + kDirTY + R <= ds.H && x0 - R >= 0 && x0 + TW - R <= ds.W;
    for (int r = wave; r < LH; r += kDirWaves) {
        const int gy = inside ? y0 - R + r : pad_index(y0 - R + r, ds.H, ds.FH, ds.kpad);
        const uint8_t* row = img + (size_t)gy * pitch;
        T* dst = copies + r * LWP;
        for (int c = lane; c < LWP; c += 64) {
            T v = T(0.f);
            if (c < TW) {
                const int gx = inside ? x0 - R + c : pad_index(x0 - R + c, ds.W, ds.FW, ds.kpad);
                v = T((float)row[gx]);
            }
#pragma unroll
            for (int sft = 0; sft < 4; ++sft)
                if (c >= sft) dst[sft * csize + c - sft] = v;
        }
    }
    __syncthreads();

    const int p = lane & 15, kb = lane >> 4;
    f32x4 acc[8];
#pragma unroll
    for (int g = 0; g < 8; ++g) acc[g] = f32x4{0.f, 0.f, 0.f, 0.f};
    // this lane's taps in K-step 0: k0 = 8*kb -> (u, v0)
    int u = 0, v0 = 8 * kb;
    while (v0 >= extp) {
        v0 -= extp;
        ++u;
    }
#pragma unroll 1
    for (int ks = 0; ks < ds.ksteps; ++ks) {
        const v8 b = bfrag[(size_t)ks * 64 + lane];
        const int uu = u < ds.ext ? u : ds.ext - 1;  // K padding past the last row: zero taps
#pragma unroll
        for (int g = 0; g < 8; ++g) {
            const int lr = 2 * wave + (g >> 2) + uu;          // tile row
            const int c = 16 * (g & 3) + p + v0;               // tile column of the window
            const int s = c & 3;
            const T* src = copies + s * csize + lr * LWP + (c - s);
            const v4 lo = *reinterpret_cast<const v4*>(src);
            const v4 hi = *reinterpret_cast<const v4*>(src + 4);
            v8 a;
            a[0] = lo[0]; a[1] = lo[1]; a[2] = lo[2]; a[3] = lo[3];
            a[4] = hi[0]; a[5] = hi[1]; a[6] = hi[2]; a[7] = hi[3];
            acc[g] = Mfma<T>::run(a, b, acc[g]);
        }
        v0 += 32;
        while (v0 >= extp) {
            v0 -= extp;
            ++u;
        }
    }

    // epilogue: lane holds filter j = lane%16 for pixels 4*(lane/16) .. +3 of each group;
    // the odd filter of orientation j/2 sits in the neighbouring lane
    const int j = lane & 15;
    const int o = j >> 1;
    const float dce = ds.dc[2 * o], dco = ds.dc[2 * o + 1];
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        const int y = y0 + 2 * wave + (g >> 2);
        const int xb = x0 + 16 * (g & 3) + 4 * (lane >> 4);
        float c4[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const float mine = acc[g][t];
            const float other = __shfl_xor(mine, 1, 64);
            const int x = xb + t;
            float bx = 0.f;
            if (y < ds.Hc && x < ds.Wc) bx = box[(size_t)y * ds.Wc + x];
            const float e = ds.fscale * ((j & 1) ? other : mine) - dce * bx;
            const float od = ds.fscale * ((j & 1) ? mine : other) - dco * bx;
            c4[t] = ds.iscale * __builtin_amdgcn_sqrtf(e * e + od * od);
        }
        if (!(j & 1) && y < ds.Hc) {
            float* out = Cm + ((size_t)o * ds.Hc + y) * ds.ldc + xb;
            if (xb + 3 < ds.Wc) {
                *reinterpret_cast<float4*>(out) = make_float4(c4[0], c4[1], c4[2], c4[3]);
            } else {
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    if (xb + t < ds.Wc) out[t] = c4[t];
            }
        }
    }
}

}  // namespace

size_t direct_lds_bytes(const DirectScale& ds) {
    return (size_t)4 * (kDirTY + 2 * ds.R) * ds.lwp * 2;
}

int direct_lwp(int R, int extp) { return (kDirTX + 2 * R + extp + 8 + 3) / 4 * 4; }

hipError_t prepare_direct() {
    hipError_t e = hipFuncSetAttribute((const void*)k_direct<__bf16>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kDirectMaxLds);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)k_direct<_Float16>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)kDirectMaxLds);
}

hipError_t launch_box(const uint8_t* img, size_t pitch, const DirectScale& ds, int kfull,
                      int* hs, float* box, hipStream_t s) {
    dim3 b(256);
    dim3 g1((ds.Wc + 255) / 256, ds.Hc + 2 * kfull);
    k_box_rows<<<g1, b, 0, s>>>(img, pitch, ds.H, ds.W, ds.FH, ds.FW, kfull, ds.Hc, ds.Wc, hs);
    dim3 g2((ds.Wc + 255) / 256, ds.Hc);
    k_box_cols<<<g2, b, 0, s>>>(hs, kfull, ds.Hc, ds.Wc, box);
    return hipGetLastError();
}

hipError_t launch_direct(const uint8_t* img, size_t pitch, const DirectScale& ds, bool fp16,
                         const void* bfrag, const float* box, float* C, hipStream_t s) {
    const size_t lds = direct_lds_bytes(ds);
    if (lds > kDirectMaxLds) return hipErrorInvalidValue;
    dim3 grid((ds.Wc + kDirTX - 1) / kDirTX, (ds.Hc + kDirTY - 1) / kDirTY);
    if (fp16)
        k_direct<_Float16><<<grid, kDirWaves * 64, lds, s>>>(img, pitch, ds,
                                                             (const f16x8*)bfrag, box, C);
    else
        k_direct<__bf16><<<grid, kDirWaves * 64, lds, s>>>(img, pitch, ds, (const bf16x8*)bfrag,
                                                           box, C);
    return hipGetLastError();
}

}  // namespace bimpk
