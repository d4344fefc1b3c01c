// kernels.hip -- gfx950 kernels of the BIMP keypoint path.
//
// Stage map (reference: DoFullGPUFiltering, bimp/src/cuda_bimp.cu:462-559):
//   k_pyr_down     pyramid level (cv::cuda::pyrDown, cuda_bimp.cu:473-479), once per level
//   k_fwd_rows     u8 -> float, clamp padding in x, real row FFTs (two rows per complex FFT)
//                  (convertTo + collectCVMemory + padDataClampToBorder + cufftExecR2C, :480-499)
//   k_fwd_cols     clamp padding in y, column FFTs -> image spectrum D, stored column-major
//   k_inv_cols     D * G_o (quadrature bank, 1/(FH*FW) folded in), inverse column FFTs
//                  (FilterSimpleCells + the column half of two cufftExecC2R, :505-515)
//   k_inv_rows     inverse row FFTs, |e + i*o| -> complex-cell map C_o (complexResponse, :518)
//   k_cells        end-stopping + tangential/radial inhibition + inhibit + max (:526-552, :720)
//   k_nms_*        threshold, strict 8-neighbour max, parabola, ordered compaction (:714-743)
#include "fft_device.hpp"
#include "kernels.hpp"

using namespace bimpdev;

namespace bimpk {

namespace {

// ------------------------------------------------------------------ pyramid
__device__ __forceinline__ int reflect101(int i, int n) {
    if (n == 1) return 0;
    for (int it = 0; it < 4 && (i < 0 || i >= n); ++it) {
        if (i < 0) i = -i;
        if (i >= n) i = 2 * (n - 1) - i;
    }
    return i;
}

// 5x5 binomial [1 4 6 4 1]^2 / 256, reflect-101 border, round half to even: the float sum is
// exact, so this equals OpenCV-CUDA pyrDown's saturate_cast of its FP32 sum.
__global__ void k_pyr_down(const uint8_t* __restrict__ src, int sw, int sh, size_t spitch,
                           uint8_t* __restrict__ dst, int dw, int dh, size_t dpitch) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= dw || y >= dh) return;
    const int wt[5] = {1, 4, 6, 4, 1};
    int acc = 0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const uint8_t* row = src + (size_t)reflect101(2 * y + i - 2, sh) * spitch;
        int racc = 0;
#pragma unroll
        for (int j = 0; j < 5; ++j) racc += wt[j] * row[reflect101(2 * x + j - 2, sw)];
        acc += wt[i] * racc;
    }
    dst[(size_t)y * dpitch + x] = (uint8_t)rintf((float)acc * (1.0f / 256.0f));
}

// ------------------------------------------------------------------ forward rows
// Two image rows (a, b) form one complex vector a + i*b; after the FFT the half spectra of
// both are separated with Z(k) and conj(Z(N-k)).  Columns beyond W repeat column W-1 for kX
// samples, then column 0 (padDataClampToBorder, convolutionFFT2D.cuh:56-85).
template <int FW>
__global__ void __launch_bounds__(256) k_fwd_rows(const uint8_t* __restrict__ img, size_t pitch,
                                                  int H, int W, int kX, float scale,
                                                  float2* __restrict__ R, int ldr,
                                                  const float2* __restrict__ tw) {
    using C = FFTCfg<FW>;
    extern __shared__ float2 lds[];
    const int g = threadIdx.x / C::TPV, lane = threadIdx.x % C::TPV;
    const int vpb = blockDim.x / C::TPV;
    const int ya = 2 * (blockIdx.x * vpb + g), yb = ya + 1;
    const bool haveA = ya < H, haveB = yb < H;
    float2* my = lds + (size_t)g * FW;
    const uint8_t* rowA = img + (size_t)(haveA ? ya : 0) * pitch;
    const uint8_t* rowB = img + (size_t)(haveB ? yb : 0) * pitch;
    float2 v[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const int x = in_index<FW>(e, lane);
        const int cx = x < W ? x : (x < W + kX ? W - 1 : 0);
        const float a = haveA ? (float)rowA[cx] * scale : 0.f;
        const float b = haveB ? (float)rowB[cx] * scale : 0.f;
        v[e] = make_float2(a, b);
    }
    fft<FW, false>(v, my, lane, tw);
#pragma unroll
    for (int e = 0; e < C::E; ++e) my[out_index<FW>(e, lane)] = v[e];
    __syncthreads();
    for (int k = lane; k <= FW / 2; k += C::TPV) {
        const float2 z1 = my[k];
        const float2 z2 = my[k == 0 ? 0 : FW - k];
        if (haveA) R[(size_t)ya * ldr + k] = make_float2(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y));
        if (haveB) R[(size_t)yb * ldr + k] = make_float2(0.5f * (z1.y + z2.y), -0.5f * (z1.x - z2.x));
    }
}

// ------------------------------------------------------------------ column kernels
// A workgroup owns CB adjacent columns; global row-major data moves through LDS (stride
// N + 8 float2 per column keeps the transposing accesses conflict-free) so that every global
// access covers CB contiguous complex values.
template <int N>
struct ColCfg {
    static constexpr int TPV = FFTCfg<N>::TPV;
    static constexpr int CB = TPV == 64 ? (N <= 2048 ? 8 : 4) : (512 / TPV > 16 ? 16 : 512 / TPV);
    static constexpr int THREADS = CB * TPV;
    static constexpr int S = N + 8;
};

// Column FFT of the row spectra R (H rows, ldr complex per row) with clamp padding in y:
// rows [H, H+kY) repeat row H-1, later rows repeat row 0.  Output column kx is written
// contiguously at DT + kx*FH.
template <int FH>
__global__ void __launch_bounds__(ColCfg<FH>::THREADS) k_fwd_cols(
        const float2* __restrict__ R, int ldr, int H, int kY, int ncols,
        float2* __restrict__ DT, const float2* __restrict__ tw) {
    using C = FFTCfg<FH>;
    using CC = ColCfg<FH>;
    extern __shared__ float2 lds[];
    const int kx0 = blockIdx.x * CC::CB;
    for (int t = threadIdx.x; t < FH * CC::CB; t += CC::THREADS) {
        const int b = t % CC::CB, y = t / CC::CB;
        const int kx = kx0 + b;
        const int ry = y < H ? y : (y < H + kY ? H - 1 : 0);
        lds[b * CC::S + y] = kx < ncols ? R[(size_t)ry * ldr + kx] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const int g = threadIdx.x / C::TPV, lane = threadIdx.x % C::TPV;
    float2* my = lds + (size_t)g * CC::S;
    float2 v[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) v[e] = my[in_index<FH>(e, lane)];
    __syncthreads();
    fft<FH, false>(v, my, lane, tw);
    const int kx = kx0 + g;
    if (kx < ncols) {
        float2* out = DT + (size_t)kx * FH;
#pragma unroll
        for (int e = 0; e < C::E; ++e) out[out_index<FH>(e, lane)] = v[e];
    }
}

// Inverse column FFT of Z_o = D * G_o for orientation o = blockIdx.y.  D holds columns
// kx <= FW/2; column kx > FW/2 is conj(D(-ky, FW-kx)) (Hermitian symmetry of a real image).
// The bank stores orientations 0..4; G_o for o = 5..7 is G_{8-o} mirrored in kx.  Rows
// y < Hc of the result are written row-major into T (Hc x FW per orientation).
template <int FH>
__global__ void __launch_bounds__(ColCfg<FH>::THREADS) k_inv_cols(
        const float2* __restrict__ DT, const float* __restrict__ G, int FW, int Hc,
        float2* __restrict__ T, const float2* __restrict__ tw) {
    using C = FFTCfg<FH>;
    using CC = ColCfg<FH>;
    extern __shared__ float2 lds[];
    const int o = blockIdx.y;
    const int kx0 = blockIdx.x * CC::CB;
    const int g = threadIdx.x / C::TPV, lane = threadIdx.x % C::TPV;
    const int kx = kx0 + g;
    const bool mirror = kx > FW / 2;
    const float2* dcol = DT + (size_t)(mirror ? FW - kx : kx) * FH;
    const int og = o <= 4 ? o : kNumOri - o;
    const int kxg = (o <= 4 || kx == 0) ? kx : FW - kx;
    const float* gcol = G + ((size_t)og * FW + kxg) * FH;
    float2 v[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const int ky = in_index<FH>(e, lane);
        float2 d;
        if (mirror) {
            const float2 t = dcol[ky == 0 ? 0 : FH - ky];
            d = make_float2(t.x, -t.y);
        } else {
            d = dcol[ky];
        }
        v[e] = cscale(d, gcol[ky]);
    }
    float2* my = lds + (size_t)g * CC::S;
    fft<FH, true>(v, my, lane, tw);
#pragma unroll
    for (int e = 0; e < C::E; ++e) my[out_index<FH>(e, lane)] = v[e];
    __syncthreads();
    float2* To = T + (size_t)o * Hc * FW;
    for (int t = threadIdx.x; t < Hc * CC::CB; t += CC::THREADS) {
        const int b = t % CC::CB, y = t / CC::CB;
        To[(size_t)y * FW + kx0 + b] = lds[b * CC::S + y];
    }
}

// ------------------------------------------------------------------ inverse rows
template <int N>
struct RowCfg {
    static constexpr int TPV = FFTCfg<N>::TPV;
    static constexpr int THREADS = 256;
    static constexpr int VPB = THREADS / TPV;
};

// Inverse row FFT of T_o row y and complex-cell magnitude sqrt(e^2 + o^2) for x < Wc
// (complexResponse_kernel, cuda_bimp.cuh:165-180).
template <int FW>
__global__ void __launch_bounds__(256) k_inv_rows(const float2* __restrict__ T, int Hc, int Wc,
                                                  float* __restrict__ Cm, int ldc,
                                                  const float2* __restrict__ tw) {
    using C = FFTCfg<FW>;
    extern __shared__ float2 lds[];
    const int o = blockIdx.y;
    const int g = threadIdx.x / C::TPV, lane = threadIdx.x % C::TPV;
    const int y = blockIdx.x * RowCfg<FW>::VPB + g;
    const bool active = y < Hc;
    const float2* row = T + ((size_t)o * Hc + (active ? y : 0)) * FW;
    float2 v[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) v[e] = row[in_index<FW>(e, lane)];
    fft<FW, true>(v, lds + (size_t)g * FW, lane, tw);
    if (active) {
        float* out = Cm + ((size_t)o * Hc + y) * ldc;
#pragma unroll
        for (int e = 0; e < C::E; ++e) {
            const int x = out_index<FW>(e, lane);
            if (x < Wc) out[x] = sqrtf(v[e].x * v[e].x + v[e].y * v[e].y);
        }
    }
}

// ------------------------------------------------------------------ cells
__device__ __forceinline__ unsigned float_key(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(unsigned k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Inclusive validity of cuda_bimp.cuh:446-449 / :479-482.
__device__ __forceinline__ bool stencil_valid(int x, int y, int o1, int o2, int W, int H) {
    if (x - o2 < 0 || x - o2 > W || x + o2 < 0 || x + o2 > W) return false;
    if (y - o1 < 0 || y - o1 > H || y + o1 < 0 || y + o1 > H) return false;
    if (x - 2 * o2 < 0 || x - 2 * o2 > W || x + 2 * o2 < 0 || x + 2 * o2 > W) return false;
    if (y - 2 * o1 < 0 || y - 2 * o1 > H || y + 2 * o1 < 0 || y + 2 * o1 > H) return false;
    return true;
}

// The reference addresses complex maps linearly (y*FW + x over an FH x FW array).  Row y <= H
// and column x <= W: x == FW continues on the next row, and a row index of FH wraps to row 0
// (one element past the reference's array, undefined there).
struct MapView {
    const float* base;
    int Hc, ldc, FH, FW;
    __device__ __forceinline__ float at(int o, int y, int x) const {
        if (x >= FW) { x -= FW; ++y; }
        if (y >= FH) y -= FH;
        return base[((size_t)o * Hc + y) * ldc + x];
    }
};

__global__ void __launch_bounds__(256) k_cells(MapView mv, int H, int W, CellParams p,
                                               float* __restrict__ Dm, float* __restrict__ Sm,
                                               unsigned* __restrict__ maxslot) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const bool in = x < W && y < H;
    float Dv = -INFINITY;
    if (in) {
        float S = 0.f, D = 0.f, Tn = 0.f, Rd = 0.f;
#pragma unroll
        for (int o = 0; o < kNumOri; ++o) {
            const int a = p.ea[o], b = p.eb[o];
            if (!stencil_valid(x, y, a, b, W, H)) continue;
            float v = mv.at(o, y - a, x + b) - mv.at(o, y + a, x - b);
            v = v > 0.f ? v : 0.f;
            if (v > 0.f) S += v;
            v = (float)((double)mv.at(o, y, x) - 0.5 * (double)mv.at(o, y - 2 * a, x + 2 * b) -
                        0.5 * (double)mv.at(o, y + 2 * a, x - 2 * b));
            v = v > 0.f ? v : 0.f;
            if (v > 0.f) D += v;
        }
#pragma unroll
        for (int o = 0; o < kNumOri; ++o) {
            const int a = p.ia[o], b = p.ib[o];
            if (!stencil_valid(x, y, a, b, W, H)) continue;
            const int o2 = (o + kNumOri / 2) % kNumOri;
            const float centre = mv.at(o, y, x);
            float v = mv.at(o, y - a, x - b) - centre;
            v = v < 0.f ? 0.f : v;
            Tn += v;
            const int a2 = p.ra[o], b2 = p.rb[o];
            v = centre - 16.f * mv.at(o2, y + a2, x + b2);
            v = v < 0.f ? 0.f : v;
            v += centre - 16.f * mv.at(o2, y - a2, x - b2);
            v = v < 0.f ? 0.f : v;
            Rd += v;
        }
        const float inhib = Tn + Rd;
        Dv = D + (-inhib);
        Dm[(size_t)y * W + x] = Dv;
        Sm[(size_t)y * W + x] = S + (-inhib);
    }
    // block maximum of D -> one atomic per workgroup
    unsigned key = in ? float_key(Dv) : 0u;
#pragma unroll
    for (int sh = 32; sh >= 1; sh >>= 1) {
        const unsigned other = __shfl_xor(key, sh, 64);
        key = other > key ? other : key;
    }
    __shared__ unsigned wmax[16];
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int nw = (blockDim.x * blockDim.y + 63) / 64;
    if ((tid & 63) == 0) wmax[tid >> 6] = key;
    __syncthreads();
    if (tid == 0) {
        unsigned m = 0;
        for (int i = 0; i < nw; ++i) m = wmax[i] > m ? wmax[i] : m;
        if (m) atomicMax(maxslot, m);
    }
}

// ------------------------------------------------------------------ NMS + compaction
struct KpOut {
    float x, y, size, angle, response;
    int octave, class_id;
};

__device__ __forceinline__ int find_scale(const NmsScale* sc, int nscales, int blk) {
    int s = 0;
    while (s + 1 < nscales && blk >= sc[s + 1].block0) ++s;
    return s;
}

// identifyKeypoints_kernel predicate (cuda_bimp.cuh:100-108): exclusive border `off`,
// D - max8(D) > 0 (dilate with the centre removed, cuda_bimp.cu:73-85, :733), D > thresh.
__device__ __forceinline__ bool nms_hit(const float* D, int H, int W, int off, float thresh,
                                        int y, int x) {
#pragma clang fp contract(off)
    if (!(x > off && x < W - off && y > off && y < H - off)) return false;
    const float* p = D + (size_t)y * W;
    const float v = p[x];
    if (!(v > thresh)) return false;
    const float* pu = p - W;
    const float* pd = p + W;
    float nb = fmaxf(pu[x - 1], pu[x]);
    nb = fmaxf(nb, pu[x + 1]);
    nb = fmaxf(nb, p[x - 1]);
    nb = fmaxf(nb, p[x + 1]);
    nb = fmaxf(nb, pd[x - 1]);
    nb = fmaxf(nb, pd[x]);
    nb = fmaxf(nb, pd[x + 1]);
    return v - nb > 0.f;
}

__device__ __forceinline__ float scale_thresh(const unsigned* maxslots, int slot, double thr) {
    const float mx = key_float(maxslots[slot]);
    return (float)((double)mx * thr);
}

__global__ void __launch_bounds__(kNmsThreads) k_nms_count(const NmsScale* __restrict__ sc,
                                                           int nscales,
                                                           const unsigned* __restrict__ maxslots,
                                                           double thr, int* __restrict__ counts) {
    const int blk = blockIdx.x;
    const int s = find_scale(sc, nscales, blk);
    const NmsScale d = sc[s];
    const float thresh = scale_thresh(maxslots, d.slot, thr);
    const int A = d.H * d.W;
    const int base = (blk - d.block0) * kNmsBlockPixels;
    int c = 0;
    for (int i = 0; i < kNmsPerThread; ++i) {
        const int pix = base + i * kNmsThreads + threadIdx.x;
        if (pix < A) c += nms_hit(d.D, d.H, d.W, d.off, thresh, pix / d.W, pix % d.W) ? 1 : 0;
    }
#pragma unroll
    for (int sh = 32; sh >= 1; sh >>= 1) c += __shfl_xor(c, sh, 64);
    __shared__ int wc[kNmsThreads / 64];
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int i = 0; i < kNmsThreads / 64; ++i) t += wc[i];
        counts[blk] = t;
    }
}

// Exclusive scan of the per-block counts in place (one workgroup); the total goes to counter.
__global__ void __launch_bounds__(1024) k_nms_scan(int* __restrict__ counts, int n,
                                                   int* __restrict__ counter) {
    __shared__ int part[1024];
    const int tid = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b0 = tid * per, b1 = min(n, b0 + per);
    int sum = 0;
    for (int i = b0; i < b1; ++i) sum += counts[i];
    part[tid] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - sum;  // exclusive prefix of this thread's range
    for (int i = b0; i < b1; ++i) {
        const int c = counts[i];
        counts[i] = run;
        run += c;
    }
    if (tid == 1023) *counter = part[1023];
}

// Re-evaluates the predicate and writes each keypoint at its row-major rank.  matrix_ld > 0
// selects the 7-row SURF_CUDA layout (cuda_bimp.cuh:132-140) instead of keypoint structs.
__global__ void __launch_bounds__(kNmsThreads) k_nms_write(
        const NmsScale* __restrict__ sc, int nscales, const unsigned* __restrict__ maxslots,
        double thr, const int* __restrict__ offsets, void* __restrict__ out, int cap,
        int matrix_ld) {
#pragma clang fp contract(off)
    const int blk = blockIdx.x;
    const int s = find_scale(sc, nscales, blk);
    const NmsScale d = sc[s];
    const float thresh = scale_thresh(maxslots, d.slot, thr);
    const int A = d.H * d.W;
    const int base = (blk - d.block0) * kNmsBlockPixels;
    __shared__ int wc[kNmsThreads / 64];
    int run = offsets[blk];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = 0; i < kNmsPerThread; ++i) {
        const int pix = base + i * kNmsThreads + threadIdx.x;
        const int y = pix / d.W, x = pix % d.W;
        const bool hit = pix < A && nms_hit(d.D, d.H, d.W, d.off, thresh, y, x);
        const unsigned long long bal = __ballot(hit);
        const int rank_in_wave = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wc[wave] = __popcll(bal);
        __syncthreads();
        int before = 0, total = 0;
        for (int w = 0; w < kNmsThreads / 64; ++w) {
            before += w < wave ? wc[w] : 0;
            total += wc[w];
        }
        __syncthreads();
        if (hit) {
            const int idx = run + before + rank_in_wave;
            if (idx < cap) {
                const float* p = d.D + (size_t)y * d.W;
                const float xv1 = p[x - 1], xv2 = p[x], xv3 = p[x + 1];
                const float yv1 = p[x - d.W], yv3 = p[x + d.W];
                const float denom_x = xv1 + xv3 - 2.f * xv2;
                const float denom_y = yv1 + yv3 - 2.f * xv2;
                const float num_x = (float)(0.5 * (double)(xv1 - xv3));
                const float num_y = (float)(0.5 * (double)(yv1 - yv3));
                float xpos = (float)x, ypos = (float)y;
                if (denom_x != 0.f) xpos += num_x / denom_x;
                if (denom_y != 0.f) ypos += num_y / denom_y;
                const float size = d.lambda * d.pyrstep;
                if (matrix_ld > 0) {
                    float* m = (float*)out;
                    m[0 * (size_t)matrix_ld + idx] = xpos * d.pyrstep;
                    m[1 * (size_t)matrix_ld + idx] = ypos * d.pyrstep;
                    ((int*)m)[2 * (size_t)matrix_ld + idx] = 1;
                    ((int*)m)[3 * (size_t)matrix_ld + idx] = (int)size;
                    m[4 * (size_t)matrix_ld + idx] = size;
                    m[5 * (size_t)matrix_ld + idx] = -1.f;
                    m[6 * (size_t)matrix_ld + idx] = xv2;
                } else {
                    KpOut k;
                    k.x = xpos * d.pyrstep;
                    k.y = ypos * d.pyrstep;
                    k.size = size;
                    k.angle = -1.f;
                    k.response = xv2;
                    k.octave = (int)size;
                    k.class_id = 1;
                    ((KpOut*)out)[idx] = k;
                }
            }
        }
        run += total;
    }
}

// Maximum of a response map (thrust::max_element, cuda_bimp.cu:960-966) when D was supplied
// by the caller instead of k_cells.
__global__ void __launch_bounds__(256) k_max(const float* __restrict__ D, int n,
                                             unsigned* __restrict__ maxslot) {
    unsigned key = 0u;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned k = float_key(D[i]);
        key = k > key ? k : key;
    }
#pragma unroll
    for (int sh = 32; sh >= 1; sh >>= 1) {
        const unsigned other = __shfl_xor(key, sh, 64);
        key = other > key ? other : key;
    }
    __shared__ unsigned wmax[4];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned m = 0;
        for (int i = 0; i < 4; ++i) m = wmax[i] > m ? wmax[i] : m;
        if (m) atomicMax(maxslot, m);
    }
}

__global__ void k_init(unsigned* maxslots, int nslots, int* counter) {
    const int i = threadIdx.x;
    if (i < nslots) maxslots[i] = 0u;
    if (i == 0) *counter = 0;
}

__global__ void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

// ------------------------------------------------------------------ dispatch
#define BIMP_FOR_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(1536) X(2048) X(2560) X(4096)

bool fft_size_supported(int n) {
#define X(N) if (n == N) return true;
    BIMP_FOR_SIZES(X)
#undef X
    return false;
}

int col_batch(int n) {
#define X(N) if (n == N) return ColCfg<N>::CB;
    BIMP_FOR_SIZES(X)
#undef X
    return 0;
}

int row_vectors_per_block(int n) {
#define X(N) if (n == N) return RowCfg<N>::VPB;
    BIMP_FOR_SIZES(X)
#undef X
    return 0;
}

template <int N>
static hipError_t prepare_n() {
    hipError_t e;
    const int rowLds = RowCfg<N>::VPB * N * (int)sizeof(float2);
    const int colLds = ColCfg<N>::CB * ColCfg<N>::S * (int)sizeof(float2);
    if ((e = hipFuncSetAttribute((const void*)k_fwd_rows<N>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, rowLds)) != hipSuccess)
        return e;
    if ((e = hipFuncSetAttribute((const void*)k_inv_rows<N>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, rowLds)) != hipSuccess)
        return e;
    if ((e = hipFuncSetAttribute((const void*)k_fwd_cols<N>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, colLds)) != hipSuccess)
        return e;
    return hipFuncSetAttribute((const void*)k_inv_cols<N>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, colLds);
}

hipError_t prepare_size(int n) {
#define X(N) if (n == N) return prepare_n<N>();
    BIMP_FOR_SIZES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_init(unsigned* maxslots, int nslots, int* counter, hipStream_t s) {
    k_init<<<1, 256, 0, s>>>(maxslots, nslots, counter);
    return hipGetLastError();
}

hipError_t launch_pyr_down(const uint8_t* src, int sw, int sh, size_t spitch, uint8_t* dst,
                           int dw, int dh, size_t dpitch, hipStream_t s) {
    dim3 grid((dw + 127) / 128, dh);
    k_pyr_down<<<grid, 128, 0, s>>>(src, sw, sh, spitch, dst, dw, dh, dpitch);
    return hipGetLastError();
}

hipError_t launch_fwd_rows(int FW, const uint8_t* img, size_t pitch, int H, int W, int kX,
                           float scale, float2* R, int ldr, const float2* tw, hipStream_t s) {
#define X(N)                                                                                 \
    if (FW == N) {                                                                           \
        const int vpb = RowCfg<N>::VPB;                                                      \
        const int pairs = (H + 1) / 2;                                                       \
        k_fwd_rows<N><<<(pairs + vpb - 1) / vpb, 256, vpb * N * sizeof(float2), s>>>(        \
            img, pitch, H, W, kX, scale, R, ldr, tw);                                        \
        return hipGetLastError();                                                            \
    }
    BIMP_FOR_SIZES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_fwd_cols(int FH, const float2* R, int ldr, int H, int kY, int ncols,
                           float2* DT, const float2* tw, hipStream_t s) {
#define X(N)                                                                                 \
    if (FH == N) {                                                                           \
        using CC = ColCfg<N>;                                                                \
        k_fwd_cols<N><<<(ncols + CC::CB - 1) / CC::CB, CC::THREADS,                          \
                        CC::CB * CC::S * sizeof(float2), s>>>(R, ldr, H, kY, ncols, DT, tw); \
        return hipGetLastError();                                                            \
    }
    BIMP_FOR_SIZES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_inv_cols(int FH, const float2* DT, const float* G, int FW, int Hc, float2* T,
                           const float2* tw, hipStream_t s) {
#define X(N)                                                                                 \
    if (FH == N) {                                                                           \
        using CC = ColCfg<N>;                                                                \
        if (FW % CC::CB) return hipErrorInvalidValue;                                        \
        dim3 grid(FW / CC::CB, kNumOri);                                                     \
        k_inv_cols<N><<<grid, CC::THREADS, CC::CB * CC::S * sizeof(float2), s>>>(            \
            DT, G, FW, Hc, T, tw);                                                           \
        return hipGetLastError();                                                            \
    }
    BIMP_FOR_SIZES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_inv_rows(int FW, const float2* T, int Hc, int Wc, float* C, int ldc,
                           const float2* tw, hipStream_t s) {
#define X(N)                                                                                 \
    if (FW == N) {                                                                           \
        const int vpb = RowCfg<N>::VPB;                                                      \
        dim3 grid((Hc + vpb - 1) / vpb, kNumOri);                                            \
        k_inv_rows<N><<<grid, 256, vpb * N * sizeof(float2), s>>>(T, Hc, Wc, C, ldc, tw);    \
        return hipGetLastError();                                                            \
    }
    BIMP_FOR_SIZES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_cells(const float* C, int Hc, int ldc, int FH, int FW, int H, int W,
                        const CellParams& p, float* Dm, float* Sm, unsigned* maxslot,
                        hipStream_t s) {
    MapView mv{C, Hc, ldc, FH, FW};
    dim3 block(64, 4);
    dim3 grid((W + 63) / 64, (H + 3) / 4);
    k_cells<<<grid, block, 0, s>>>(mv, H, W, p, Dm, Sm, maxslot);
    return hipGetLastError();
}

hipError_t launch_max(const float* D, int n, unsigned* maxslot, hipStream_t s) {
    int blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    k_max<<<blocks, 256, 0, s>>>(D, n, maxslot);
    return hipGetLastError();
}

hipError_t launch_nms_count(const NmsScale* scales, int nscales, int total_blocks,
                            const unsigned* maxslots, double thresh, int* block_counts,
                            hipStream_t s) {
    k_nms_count<<<total_blocks, kNmsThreads, 0, s>>>(scales, nscales, maxslots, thresh,
                                                     block_counts);
    return hipGetLastError();
}

hipError_t launch_nms_scan(int* block_counts, int total_blocks, int* counter, hipStream_t s) {
    k_nms_scan<<<1, 1024, 0, s>>>(block_counts, total_blocks, counter);
    return hipGetLastError();
}

hipError_t launch_nms_write(const NmsScale* scales, int nscales, int total_blocks,
                            const unsigned* maxslots, double thresh, const int* block_offsets,
                            void* out, int cap, int matrix_ld, hipStream_t s) {
    k_nms_write<<<total_blocks, kNmsThreads, 0, s>>>(scales, nscales, maxslots, thresh,
                                                     block_offsets, out, cap, matrix_ld);
    return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes % 16) return hipErrorInvalidValue;
    const size_t n = bytes / 16;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    k_copy<<<blocks, 256, 0, s>>>((const uint4*)src, (uint4*)dst, n);
    return hipGetLastError();
}

}  // namespace bimpk
