// kernels.hip -- gfx950 kernels of the BIMP keypoint path.
//
// Stage map (reference: DoFullGPUFiltering, bimp/src/cuda_bimp.cu:462-559):
//   k_pyr_down     pyramid level (cv::cuda::pyrDown, cuda_bimp.cu:473-479), once per level
//   k_fwd_rows     u8 (or float) -> float, clamp padding in x, real row FFTs (two rows per complex FFT)
//                  (convertTo + collectCVMemory + padDataClampToBorder + cufftExecR2C, :480-499)
//   k_fwd_cols     clamp padding in y, column FFTs -> image spectrum D, stored column-major
//   k_inv_cols     D * G_o (quadrature bank, 1/(FH*FW) folded in), inverse column FFTs
//                  (FilterSimpleCells + the column half of two cufftExecC2R, :505-515)
//   k_inv_rows     inverse row FFTs, |e + i*o| -> complex-cell map C_o (complexResponse, :518)
//   k_cells        end-stopping + tangential/radial inhibition + inhibit + max (:526-552, :720)
//   k_nms_*        threshold, strict 8-neighbour max, parabola, ordered compaction (:714-743)
#include <algorithm>
#include <cmath>
#include <vector>

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

// 5x5 binomial [1 4 6 4 1]^2 / 256, reflect-101 border.  Round half to even: the float sum is
// exact, so this equals OpenCV-CUDA pyrDown's saturate_cast of its FP32 sum
// (cuda_bimp.cu:476).  half_up: (S + 128) >> 8, the fixed-point rounding of OpenCV's host
// pyrDown, which the reference's resize_cpu option uses (cuda_bimp.cu:660-665).
__global__ void k_pyr_down(const uint8_t* __restrict__ src, int sw, int sh, size_t spitch,
                           uint8_t* __restrict__ dst, int dw, int dh, size_t dpitch,
                           int half_up) {
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
    dst[(size_t)y * dpitch + x] =
        half_up ? (uint8_t)((acc + 128) >> 8) : (uint8_t)rintf((float)acc * (1.0f / 256.0f));
}

// ------------------------------------------------------------------ forward rows
template <int N>
struct RowCfg {
    static constexpr int TPV = FFTCfg<N>::TPV;
    static constexpr int THREADS = TPV > 64 ? TPV : 64;  // one wavefront, or one 128-lane group
    static constexpr int VPB = THREADS / TPV;
};

// Two image rows (a, b) form one complex vector a + i*b; after the FFT the half spectra of
// both are separated with Z(k) and conj(Z(N-k)).  Columns beyond W repeat column W-1 for kX
// samples, then column 0 (padDataClampToBorder, convolutionFFT2D.cuh:56-85).  blockIdx.y
// selects the scale of the batch (scales of one pyramid level differ only in kX).  Pixels are
// uint8 (detection) or float (the differentiation form's images); pitch counts pixels.
template <int FW, class Pix>
__global__ void __launch_bounds__(RowCfg<FW>::THREADS) k_fwd_rows(const Pix* __restrict__ img, size_t pitch,
                                                 int H, int W, ScaleSet ss, float scale,
                                                 float2* __restrict__ Rbase, int ldr,
                                                 size_t rstride, const float2* __restrict__ tw) {
    using C = FFTCfg<FW>;
    extern __shared__ float lds[];
    const int sidx = blockIdx.y;
    const int kX = ss.kX[sidx];
    float2* R = Rbase + (size_t)sidx * rstride;
    const int g = threadIdx.x / C::TPV, lane = threadIdx.x % C::TPV;
    const int ya = 2 * (blockIdx.x * RowCfg<FW>::VPB + g), yb = ya + 1;
    const bool haveA = ya < H, haveB = yb < H;
    float* my = lds + (size_t)g * 2 * FW;
    const Pix* rowA = img + (size_t)(haveA ? ya : 0) * pitch;
    const Pix* rowB = img + (size_t)(haveB ? yb : 0) * pitch;
    // unconditional loads (rows clamped above) so that all of them are in flight at once
    const float sa = haveA ? scale : 0.f, sb = haveB ? scale : 0.f;
    float2 v[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const int x = in_index<FW>(e, lane);
        const int cx = x < W ? x : (x < W + kX ? W - 1 : 0);
        v[e] = make_float2((float)rowA[cx] * sa, (float)rowB[cx] * sb);
    }
    fft<FW, false>(v, my, lane, tw);
    float2* z = reinterpret_cast<float2*>(my);
#pragma unroll
    for (int e = 0; e < C::E; ++e) z[out_index<FW>(e, lane)] = v[e];
    stage_sync<C::TPV>();
    for (int k = lane; k <= FW / 2; k += C::TPV) {
        const float2 z1 = z[k];
        const float2 z2 = z[k == 0 ? 0 : FW - k];
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
    // The forward column pass stages its columns in LDS in two halves of N/2 rows when whole
    // columns would take more than 64 KB, so that several workgroups share a CU.
    static constexpr int HALVES = (size_t)CB * (N + 2) * sizeof(float2) > 64u * 1024u ? 2 : 1;
    // column stride (complex): 2 (mod 16) puts the 8 columns of a 16-lane access 4 banks
    // apart, so staging writes (column index fastest) do not collide
    static constexpr int S = N / HALVES + 2;
};

// Column FFT of the row spectra R (H rows, ldr complex per row) with clamp padding in y:
// rows [H, H+kY) repeat row H-1, later rows repeat row 0.  Output column kx is written
// contiguously at DT + kx*FH.  blockIdx.y selects the scale of the batch.
template <int FH>
__global__ void __launch_bounds__(ColCfg<FH>::THREADS) k_fwd_cols(
        const float2* __restrict__ Rbase, int ldr, size_t rstride, int H, ScaleSet ss, int ncols,
        const float2* __restrict__ tw) {
    using C = FFTCfg<FH>;
    using CC = ColCfg<FH>;
    extern __shared__ float2 lds2[];
    // Workgroups are dealt round-robin to the 8 XCDs by linear id; pairs of adjacent column
    // blocks go to one XCD, so the two halves of each 128-byte row segment of R are fetched
    // into the same L2 (full groups of 16 blocks; the remainder keeps its order).
    int sidx, blk;
    {
        const int nb = gridDim.x, total = nb * gridDim.y;
        const int L = blockIdx.x + blockIdx.y * nb, full = total / 16 * 16;
        const int p = L < full ? 16 * (L / 16) + 2 * (L % 8) + ((L / 8) & 1) : L;
        sidx = p / nb;
        blk = p % nb;
    }
    const int kY = ss.kX[sidx];
    const float2* R = Rbase + (size_t)sidx * rstride;
    float2* DT = ss.DT[sidx];
    const int kx0 = blk * CC::CB;
    const int g = threadIdx.x / C::TPV, lane = threadIdx.x % C::TPV;
    float2* my = lds2 + (size_t)g * CC::S;
    float2 v[C::E];
    constexpr int HN = FH / CC::HALVES;  // rows staged per half
    // register slot e holds row in_index(e) in [(e % R) N/R, (e % R + 1) N/R), R the first
    // radix: with R even its half is known at compile time
    constexpr int R1 = RadInfo<typename C::Rad>::first;
    static_assert(CC::HALVES == 1 || R1 % 2 == 0, "halves follow the first radix");
#pragma unroll
    for (int h = 0; h < CC::HALVES; ++h) {
        if (h > 0) __syncthreads();  // every wavefront has read the previous half
        // a thread stages column b of the block, rows h*HN + y0 + i*YS; eight loads go out
        // before their LDS stores
        static_assert(CC::THREADS % CC::CB == 0 && (HN * CC::CB) % CC::THREADS == 0, "layout");
        constexpr int YS = CC::THREADS / CC::CB, IT = HN * CC::CB / CC::THREADS;
        const int b = threadIdx.x % CC::CB, y0 = threadIdx.x / CC::CB;
        const int kx = kx0 + b;
        const bool in = kx < ncols;
#pragma unroll
        for (int i0 = 0; i0 < IT; i0 += 8) {
            float2 val[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int y = h * HN + y0 + (i0 + j) * YS;
                const int ry = y < H ? y : (y < H + kY ? H - 1 : 0);
                val[j] = (i0 + j < IT && in) ? R[(size_t)ry * ldr + kx] : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (i0 + j < IT) lds2[b * CC::S + y0 + (i0 + j) * YS] = val[j];
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < C::E; ++e)
            if (CC::HALVES == 1 || ((e % R1) >= R1 / 2) == (h == 1))
                v[e] = my[in_index<FH>(e, lane) - h * HN];
    }
    stage_sync<C::TPV>();
    fft<FH, false>(v, reinterpret_cast<float*>(my), lane, tw);
    const int kx = kx0 + g;
    if (kx < ncols) {
        float2* out = DT + (size_t)kx * FH;
#pragma unroll
        for (int e = 0; e < C::E; ++e) out[out_index<FH>(e, lane)] = v[e];
    }
}

// Inverse column pass.  A workgroup owns one orientation o and CB stored spectrum columns
// kx in [k0, k0+CB) (kx <= FW/2) plus the Hermitian partners FW-kx of the columns
// [k0+1, k0+CB+1), so that both output blocks start on a CB-aligned column.  A partner
// column's spectrum is conj(D(-ky, kx)); its inverse transform is taken as
// conj(IFFT[D(ky, kx) * G(-ky, FW-kx)]) (G is real).  The bank stores orientations 0..4; G_o
// for o = 5..7 is G_{8-o} mirrored in kx.  Workgroups are dealt round-robin to the 8 XCDs;
// the numbering puts the eight orientations of one column block back to back on one XCD, so
// D and the shared bank columns come from HBM once and from that XCD's L2 afterwards.  Rows
// y < Hc go row-major into T (Hc x FW per orientation) through LDS, half at a time.
template <int FH>
struct InvColCfg {
    static constexpr int CB = ColCfg<FH>::CB;
    static constexpr int THREADS = ColCfg<FH>::THREADS;
    // rows leave through LDS in the blocks of the last radix: register slot e holds row
    // lane + TPV*(e/RL) + (e%RL)*BR, so block membership is known at compile time
    static constexpr int RL = RadInfo<typename FFTCfg<FH>::Rad>::last;
    static constexpr int BR = FH / RL;
    static constexpr size_t EXCH = (size_t)CB * fft_lds_floats<FH>() * sizeof(float);
    // GB consecutive blocks (contiguous rows) leave per staging round, each round costing two
    // workgroup barriers: the most that divide RL and stay within half a column, so the
    // staging region is the exchange region plus the 2-element pad per column (3840x2160:
    // two rounds instead of four cut k_inv_cols<2560> from 413 to 292 us per launch)
    static constexpr int pick_gb() {
        for (int g = RL; g > 1; --g)
            if (RL % g == 0 && g * BR <= FH / 2) return g;
        return 1;
    }
    static constexpr int GB = pick_gb();
    static constexpr int BRG = GB * BR;
    static constexpr int SS = BRG + 2;  // staging stride (float2); 2 (mod 16): conflict-free reads
    static constexpr size_t STAGE = (size_t)CB * SS * sizeof(float2);
    static constexpr size_t MAIN = STAGE > EXCH ? STAGE : EXCH;
    // The persistent workgroup keeps the twiddle table in LDS (after the exchange/staging
    // region) when that still leaves room for two workgroups per CU: LDS reads instead of
    // cache round trips inside each stage.
    // 4-wavefront workgroups (CB = 4) need three per CU for enough wavefronts: the twiddles
    // then stay in global memory (L1/L2 hits) when keeping them in LDS would leave room for
    // only two (2560 points: 516 -> 422 us per launch at 3840x2160)
    static constexpr bool TW_COSTS_A_WG = CB == 4 && 3 * MAIN <= kMaxLds &&
                                          3 * (MAIN + tw_lds_bytes<FH>()) > kMaxLds;
    static constexpr bool LDS_TW = MAIN + tw_lds_bytes<FH>() <= 80u * 1024u && !TW_COSTS_A_WG;
    static constexpr size_t LDS = MAIN + (LDS_TW ? tw_lds_bytes<FH>() : 0);
};

template <int FH>
__global__ void __launch_bounds__(ColCfg<FH>::THREADS) k_inv_cols(
        ScaleSet ss, int FW, int Hc, int nblk, int units, float2* __restrict__ Tbase,
        size_t tstride, const float2* __restrict__ tw) {
    using C = FFTCfg<FH>;
    using IC = InvColCfg<FH>;
    constexpr int CB = IC::CB;
    extern __shared__ float2 lds2[];
    const int half = FW / 2;
    // with one transform per wavefront the group index is uniform per wavefront: kept
    // scalar, so every column pointer is (smaller transforms put several groups in a wavefront)
    const int g = C::TPV == 64 ? __builtin_amdgcn_readfirstlane(threadIdx.x / 64)
                               : (int)(threadIdx.x / C::TPV);
    const int lane0 = threadIdx.x % C::TPV;
    float* my = reinterpret_cast<float*>(lds2) + (size_t)g * fft_lds_floats<FH>();
    float4* twl = reinterpret_cast<float4*>(reinterpret_cast<char*>(lds2) + IC::MAIN);
    if constexpr (IC::LDS_TW) {
        tw_to_lds<FH>(tw, twl, threadIdx.x, IC::THREADS);
        __syncthreads();
    }
    // A unit is one pass (direct or partner) of one orientation of one column block.  The
    // grid is at most one resident round of workgroups and walks the units with stride
    // gridDim.x (a multiple of 8), so the work splits evenly in half-block units instead of
    // leaving most of a last round idle.  XCD-aware numbering: unit u runs on XCD u % 8 and
    // its index i = u / 8 there gives orientation i % 8, pass (i / 8) % 2 and column block
    // 8 * (i / 16) + u % 8, so the orientations and both passes of a column block stay on one
    // XCD and share its D and bank columns through that XCD's L2.
#pragma unroll 1
    for (int u = blockIdx.x; u < units; u += gridDim.x) {
        const int xcd = u % 8, i = u / 8;
        const int o = i % kNumOri, partner = (i / kNumOri) % 2, gcb = (i / (2 * kNumOri)) * 8 + xcd;
        if (gcb >= nblk * ss.n) continue;  // uniform per workgroup: no barrier is skipped
        const int sidx = gcb / nblk, k0 = (gcb % nblk) * CB;
        const float2* DT = ss.DT[sidx];
        const float* G = ss.G[sidx];
        float2* To = Tbase + (size_t)sidx * tstride + (size_t)o * Hc * FW;
        const int og = o <= 4 ? o : kNumOri - o;
        // lane-derived indices are recomputed per unit rather than kept live across units
        int lane = lane0;
        asm volatile("" : "+v"(lane));
        // direct: column kx = k0 + g; partner: stored column kx = k0 + 1 + g, output FW - kx
        const int kx = k0 + partner + g;
        const bool ok = partner ? (kx < half) : (kx <= half);
        const float2* dcol = DT + (size_t)(ok ? kx : 0) * FH;
        // bank column: direct o <= 4 -> kx, o >= 5 -> FW-kx; partner o <= 4 -> FW-kx, else kx
        const int cg = ((o <= 4) != (partner != 0)) ? kx : (kx == 0 ? 0 : FW - kx);
        const float* gcol = G + ((size_t)og * FW + (ok ? cg : 0)) * FH;
        // D * G_o: per-lane base pointers, element offsets in_index(e, 0) known at compile
        // time; the partner pass reads G at row FH - ky, except ky = 0 (lane 0 of e = 0)
        float2 v[C::E];
        const float2* dl = dcol + lane;
        if (partner) {
            const float* gl = gcol + (FH - lane);
#pragma unroll
            for (int e = 0; e < C::E; ++e) {
                const int ce = in_index<FH>(e, 0);
                const float gv = e == 0 ? gcol[lane ? FH - lane : 0] : gl[-ce];
                v[e] = cscale(dl[ce], gv);
            }
        } else {
            const float* gl = gcol + lane;
#pragma unroll
            for (int e = 0; e < C::E; ++e) {
                const int ce = in_index<FH>(e, 0);
                v[e] = cscale(dl[ce], gl[ce]);
            }
        }
        if constexpr (IC::LDS_TW)
            fft<FH, true>(v, my, lane,
                          LdsTw{(const __attribute__((address_space(3))) float4*)twl});
        else
            fft<FH, true>(v, my, lane, tw);
        if (partner) {
#pragma unroll
            for (int e = 0; e < C::E; ++e) v[e].y = -v[e].y;
        }
        // output column of staging slot b: direct k0 + b; partner FW - (k0 + 1 + (CB-1-b))
        const int slot = partner ? CB - 1 - g : g;
        const int base = partner ? FW - k0 - CB : k0;
        // a thread stores column b of the block, rows yy0 + i*YS (IC::THREADS is a multiple
        // of CB); where a round takes 8 or more such rows, four LDS reads go out before their
        // four stores instead of one read waiting before each store
        static_assert(IC::THREADS % CB == 0, "whole rows of the block per pass");
        constexpr int YS = IC::THREADS / CB, IT = (IC::BRG + YS - 1) / YS;
        constexpr int J = IT % 4 == 0 ? 4 : IT % 3 == 0 ? 3 : IT % 2 == 0 ? 2 : 1;
        constexpr bool BATCH = IT >= 6 && J > 1 && IC::BRG % YS == 0;
        const int b = threadIdx.x % CB, yy0 = threadIdx.x / CB;
        const int ks = partner ? k0 + CB - b : k0 + b;  // stored column
        const bool wcol = partner ? (ks < half) : (ks <= half);
#pragma unroll
        for (int rq = 0; rq < IC::RL / IC::GB; ++rq) {
            const int ylo = rq * IC::BRG;
            if (ylo >= Hc) break;  // uniform
            const int rows = min(Hc - ylo, IC::BRG);
            __syncthreads();  // transforms and earlier readers of the region are done
#pragma unroll
            for (int m = 0; m < C::E / IC::RL; ++m)
#pragma unroll
                for (int gb = 0; gb < IC::GB; ++gb)
                    lds2[slot * IC::SS + gb * IC::BR + lane + C::TPV * m] =
                            v[m * IC::RL + rq * IC::GB + gb];
            __syncthreads();
            // T is column-block-major: block base / CB holds rows 0..Hc-1 of its CB columns
            // contiguously, so a unit's rows go out as one contiguous run
            float2* const blk = To + (size_t)(base / CB) * Hc * CB;
            if constexpr (BATCH) {
#pragma unroll 1
                for (int i0 = 0; i0 < IT; i0 += J) {
                    float2 val[J];
#pragma unroll
                    for (int j = 0; j < J; ++j) val[j] = lds2[b * IC::SS + yy0 + (i0 + j) * YS];
                    float2* dst = blk + (size_t)(ylo + yy0 + i0 * YS) * CB + b;
#pragma unroll
                    for (int j = 0; j < J; ++j)
                        if (wcol && yy0 + (i0 + j) * YS < rows) dst[(size_t)j * YS * CB] = val[j];
                }
            } else {
                for (int t = threadIdx.x; t < rows * CB; t += IC::THREADS) {
                    const int bt = t % CB, yy = t / CB;
                    const int kst = partner ? k0 + CB - bt : k0 + bt;  // stored column
                    const bool w = partner ? (kst < half) : (kst <= half);
                    if (w) blk[(size_t)(ylo + yy) * CB + bt] = lds2[bt * IC::SS + yy];
                }
            }
        }
        __syncthreads();  // the exchange region is reused by the next transform
    }
}

// ------------------------------------------------------------------ inverse rows
// Inverse row FFT of T_o row y and complex-cell magnitude sqrt(e^2 + o^2) for x < Wc
// (complexResponse_kernel, cuda_bimp.cuh:165-180).  blockIdx.y = orientation,
// blockIdx.z = scale of the batch.
// Row y of T_o in the column-block-major layout the column pass writes (blocks of 2^tsh
// columns, rows contiguous within a block): element x sits at
// ((x >> tsh) * Hc + y) << tsh | (x & (2^tsh - 1)) from the orientation's base.
template <int FW>
__device__ __forceinline__ void load_t_row(float2 (&v)[FFTCfg<FW>::E], const float2* To, int Hc,
                                           int y, int tsh, int lane) {
    using C = FFTCfg<FW>;
    if constexpr (C::TPV == 64) {
        // in_index(e, lane) = lane + c_e with c_e a multiple of 64 >= the block width
        const float2* p = To + ((((size_t)(lane >> tsh) * Hc + y) << tsh) | (lane & ((1 << tsh) - 1)));
#pragma unroll
        for (int e = 0; e < C::E; ++e) v[e] = p[(size_t)in_index<FW>(e, 0) * Hc];
    } else {
#pragma unroll
        for (int e = 0; e < C::E; ++e) {
            const int x = in_index<FW>(e, lane);
            v[e] = To[(((size_t)(x >> tsh) * Hc + y) << tsh) | (x & ((1 << tsh) - 1))];
        }
    }
}

template <int FW>
__global__ void __launch_bounds__(RowCfg<FW>::THREADS) k_inv_rows(ScaleSet ss, const float2* __restrict__ Tbase,
                                                 size_t tstride, int Hc, int Wc, int ldc, int tsh,
                                                 const float2* __restrict__ tw) {
    using C = FFTCfg<FW>;
    extern __shared__ float lds[];
    const int o = blockIdx.y, sidx = blockIdx.z;
    const int g = threadIdx.x / C::TPV, lane = threadIdx.x % C::TPV;
    const int y = blockIdx.x * RowCfg<FW>::VPB + g;
    const bool active = y < Hc;
    float2 v[C::E];
    load_t_row<FW>(v, Tbase + (size_t)sidx * tstride + (size_t)o * Hc * FW, Hc, active ? y : 0,
                   tsh, lane);
    fft<FW, true, 8>(v, lds + (size_t)g * fft_lds_floats<FW>(), lane, tw);
    if (active) {
        float* out = ss.C[sidx] + ((size_t)o * Hc + y) * ldc;
#pragma unroll
        for (int e = 0; e < C::E; ++e) {
            const int x = out_index<FW>(e, lane);
            // native v_sqrt_f32 (1 ulp): the map tolerance is 1e-4 relative
            if (x < Wc) out[x] = __builtin_amdgcn_sqrtf(v[e].x * v[e].x + v[e].y * v[e].y);
        }
    }
}

// Persistent form for transforms of one or two full wavefronts (TPV = 64 or 128): each
// workgroup copies the twiddle table into LDS once and its RPU groups then walk units of RPU
// consecutive rows (one row per group) of one orientation of one scale, stride gridDim.x.
// Same arithmetic as k_inv_rows.
template <int FW>
struct InvRowCfg {
    static constexpr int WPB = 4;
    static constexpr int THREADS = 64 * WPB;
    static constexpr int RPU = THREADS / FFTCfg<FW>::TPV;  // rows per unit
    static constexpr size_t MAIN = (size_t)RPU * fft_lds_floats<FW>() * sizeof(float);
    static constexpr size_t LDS = MAIN + tw_lds_bytes<FW>();
    static constexpr bool PERSISTENT = FFTCfg<FW>::TPV >= 64;
};

template <int FW>
__global__ void __launch_bounds__(InvRowCfg<FW>::THREADS) k_inv_rows_p(
        ScaleSet ss, const float2* __restrict__ Tbase, size_t tstride, int Hc, int Wc, int ldc,
        int tsh, int units, const float2* __restrict__ tw) {
    using C = FFTCfg<FW>;
    using RC = InvRowCfg<FW>;
    static_assert(C::TPV >= 64 && RC::THREADS % C::TPV == 0, "whole groups of wavefronts");
    extern __shared__ float lds[];
    const int w = threadIdx.x / C::TPV, lane0 = threadIdx.x % C::TPV;
    float* my = lds + (size_t)w * fft_lds_floats<FW>();
    float4* twl = reinterpret_cast<float4*>(reinterpret_cast<char*>(lds) + RC::MAIN);
    tw_to_lds<FW>(tw, twl, threadIdx.x, RC::THREADS);
    __syncthreads();
    const int nrb = (Hc + RC::RPU - 1) / RC::RPU;
#pragma unroll 1
    for (int u = blockIdx.x; u < units; u += gridDim.x) {
        const int rb = u % nrb, o = (u / nrb) % kNumOri, sidx = u / (nrb * kNumOri);
        // lane-derived indices are recomputed per unit rather than kept live across units
        int lane = lane0;
        asm volatile("" : "+v"(lane));
        const int y = rb * RC::RPU + w;
        const bool active = y < Hc;
        float2 v[C::E];
        load_t_row<FW>(v, Tbase + (size_t)sidx * tstride + (size_t)o * Hc * FW, Hc,
                       active ? y : 0, tsh, lane);
        fft<FW, true, 8>(v, my, lane, LdsTw{(const __attribute__((address_space(3))) float4*)twl});
        if (active) {
            float* out = ss.C[sidx] + ((size_t)o * Hc + y) * ldc;
#pragma unroll
            for (int e = 0; e < C::E; ++e) {
                const int x = out_index<FW>(e, lane);
                // native v_sqrt_f32 (1 ulp): the map tolerance is 1e-4 relative
                if (x < Wc) out[x] = __builtin_amdgcn_sqrtf(v[e].x * v[e].x + v[e].y * v[e].y);
            }
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

// Output tile of the cells kernel: 64 x 16 pixels, 512 threads, 2 rows per thread.  The
// radial term couples orientation o with o+4, so the maps are staged in LDS in the pairs
// (o, o+4), o = 0..3, each tile with a halo of HS pixels (HS >= twice the largest offset, a
// multiple of 4 so that the staging moves aligned float4).  While one pair is evaluated the
// next pair is already in flight into registers.  The contributions of o+4 wait in
// registers so that every sum still runs in the order o = 0..7 of the reference's serialised
// launches; an invalid orientation contributes +0, which leaves a sum of non-negative terms
// unchanged (so does a -0 from the clamps).
constexpr int kCellTW = 64, kCellTH = 16, kCellTY = 8, kCellRows = kCellTH / kCellTY;
// dynamic LDS limit per workgroup (160 KiB per CU, minus room for static LDS)

struct CellTileOffsets {
    // LDS element offsets relative to the pixel's own position in a staged tile.  The
    // end-stop taps are +-es and +-2 es, es = (-a,b); the radial taps are +-rp on the partner
    // map.  Only the independent ones are passed: fewer scalar registers.
    int es[kNumOri];  // (-a,b)
    int tn[kNumOri];  // (-a',-b')
    int rp[kNumOri];  // (a'',b'')
    // validity: 2|b| <= x <= W - 2|b| and 2|a| <= y <= H - 2|a| (end-stop, inhibition), the
    // x bound in the low and the y bound in the high 16 bits
    int exy[kNumOri], ixy[kNumOri];
};

struct Contrib {
    float s, d, t, r;
};

// Contributions of orientation o at a thread's kCellRows pixels (tile elements li + rr*RS):
// m = its own staged tile, m2 = the tile of (o+4)%8.  Each tap's pointer is uniform, so one
// address add serves every row and the row step folds into the LDS instruction's offset.
template <int RS>
__device__ __forceinline__ void cell_terms(const float* __restrict__ m,
                                           const float* __restrict__ m2,
                                           const CellTileOffsets& q, int o, int li,
                                           Contrib (&c)[kCellRows]) {
#pragma clang fp contract(off)
    const int es = q.es[o], r = q.rp[o];
    const float* __restrict__ s1 = m + es;
    const float* __restrict__ s2 = m - es;
    const float* __restrict__ d1p = m + 2 * es;
    const float* __restrict__ d2p = m - 2 * es;
    const float* __restrict__ tn = m + q.tn[o];
    const float* __restrict__ rp = m2 + r;
    const float* __restrict__ rm = m2 - r;
#pragma unroll
    for (int rr = 0; rr < kCellRows; ++rr) {
        const int l = li + rr * RS;
        // Products by 0.5 and 16 are exact, so fma(-0.5, a, c) and fma(-16, m, c) round once
        // exactly where the reference's separate multiply and subtract do: same bits, fewer ops.
        const float centre = m[l];
        c[rr].s = fmaxf(s1[l] - s2[l], 0.f);
        const double d1 = __builtin_fma(-0.5, (double)d1p[l], (double)centre);
        c[rr].d = fmaxf((float)__builtin_fma(-0.5, (double)d2p[l], d1), 0.f);
        c[rr].t = fmaxf(tn[l] - centre, 0.f);
        float v = fmaxf(__builtin_fmaf(-16.f, rp[l], centre), 0.f);
        v += __builtin_fmaf(-16.f, rm[l], centre);
        c[rr].r = fmaxf(v, 0.f);
    }
}

template <int HS>
struct CellGeom {
    static constexpr int LW = kCellTW + 2 * HS, LH = kCellTH + 2 * HS, TS = LW * LH;
    static constexpr int LW4 = LW / 4;
    static constexpr int NV4 = 2 * LH * LW4;                       // float4 per staged pair
    static constexpr int PER = (NV4 + kCellTW * kCellTY - 1) / (kCellTW * kCellTY);
};

// Loads (INTERIOR: directly; else with the linear-index semantics and zero outside
// [0,H] x [0,W]) the float4s of pair pr that this thread stages.
template <int HS, bool INTERIOR>
__device__ __forceinline__ void cells_fetch(const MapView& mv, int H, int W, int x0, int y0,
                                            int pr, int tid, float4 (&r)[CellGeom<HS>::PER]) {
    using CG = CellGeom<HS>;
#pragma unroll
    for (int k = 0; k < CG::PER; ++k) {
        const int i = tid + k * kCellTW * kCellTY;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if ((k + 1) * kCellTW * kCellTY <= CG::NV4 || i < CG::NV4) {
            const int row = i / CG::LW4, c4 = i - row * CG::LW4;
            const int m = row >= CG::LH ? 1 : 0;
            const int ry = row - m * CG::LH;
            const int o = pr + 4 * m;
            const int gy = y0 - HS + ry, gx = x0 - HS + 4 * c4;
            if (INTERIOR) {
                v = *reinterpret_cast<const float4*>(mv.base + ((size_t)o * mv.Hc + gy) * mv.ldc + gx);
            } else {
                float e[4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    e[j] = (gy >= 0 && gy <= H && gx + j >= 0 && gx + j <= W) ? mv.at(o, gy, gx + j)
                                                                                : 0.f;
                v = make_float4(e[0], e[1], e[2], e[3]);
            }
        }
        r[k] = v;
    }
}

// Interior tiles: byte offsets, from the map of the pair's first orientation, of the float4s
// this thread stages (the pair's second orientation lies 4 maps further on).  Computed once
// per tile; a pair then adds only a uniform 64-bit base, and the loads take the
// scalar-base + 32-bit-offset form.  The launcher keeps 8 maps' bytes below 2^32.
template <int HS>
__device__ __forceinline__ void cells_offsets(const MapView& mv, int x0, int y0, int tid,
                                              unsigned (&off)[CellGeom<HS>::PER]) {
    using CG = CellGeom<HS>;
#pragma unroll
    for (int k = 0; k < CG::PER; ++k) {
        const int i = tid + k * kCellTW * kCellTY;
        const int row = i / CG::LW4, c4 = i - row * CG::LW4;
        const int m = row >= CG::LH ? 1 : 0;
        const int gy = y0 - HS + row - m * CG::LH, gx = x0 - HS + 4 * c4;
        // interior tiles: gy, gx >= 0
        off[k] = i < CG::NV4 ? ((unsigned)(4 * m * mv.Hc + gy) * (unsigned)mv.ldc + (unsigned)gx) *
                                   (unsigned)sizeof(float)
                             : 0u;
    }
}

template <int HS>
__device__ __forceinline__ void cells_fetch_interior(const MapView& mv, int pr, int tid,
                                                     const unsigned (&off)[CellGeom<HS>::PER],
                                                     float4 (&r)[CellGeom<HS>::PER]) {
    using CG = CellGeom<HS>;
    const char* base = reinterpret_cast<const char*>(mv.base + (size_t)pr * mv.Hc * mv.ldc);
#pragma unroll
    for (int k = 0; k < CG::PER; ++k) {
        const int i = tid + k * kCellTW * kCellTY;
        // every thread stages a float4 for k with (k+1) * NT <= NV4
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if ((k + 1) * kCellTW * kCellTY <= CG::NV4 || i < CG::NV4)
            v = *reinterpret_cast<const float4*>(base + off[k]);
        r[k] = v;
    }
}

template <int HS>
__device__ __forceinline__ void cells_put(float* __restrict__ tile, int tid,
                                          const float4 (&r)[CellGeom<HS>::PER]) {
    using CG = CellGeom<HS>;
#pragma unroll
    for (int k = 0; k < CG::PER; ++k) {
        const int i = tid + k * kCellTW * kCellTY;
        if ((k + 1) * kCellTW * kCellTY <= CG::NV4 || i < CG::NV4)
            reinterpret_cast<float4*>(tile)[i] = r[k];
    }
}

// INTERIOR: every pixel of the tile is valid for every orientation and the staged tile with
// its halo lies inside [0,H] x [0,W] without wrap-around, so no test is needed anywhere.
template <int HS, bool INTERIOR>
__device__ __forceinline__ void cells_tile(const MapView& mv, int H, int W,
                                           const CellTileOffsets& q, int x0, int y0,
                                           float* __restrict__ tile, float* __restrict__ Dm,
                                           float* __restrict__ Sm, float& Dmax, bool& any) {
#pragma clang fp contract(off)
    using CG = CellGeom<HS>;
    const int tid = threadIdx.y * kCellTW + threadIdx.x;
    const int x = x0 + threadIdx.x;
    float accS[kCellRows], accD[kCellRows], accT[kCellRows], accR[kCellRows];
    Contrib hold[kCellRows][4];
#pragma unroll
    for (int rr = 0; rr < kCellRows; ++rr) accS[rr] = accD[rr] = accT[rr] = accR[rr] = 0.f;
    float4 pre[CG::PER];
    unsigned off[CG::PER];
    if (INTERIOR) {
        cells_offsets<HS>(mv, x0, y0, tid, off);
        cells_fetch_interior<HS>(mv, 0, tid, off, pre);
    } else {
        cells_fetch<HS, false>(mv, H, W, x0, y0, 0, tid, pre);
    }
    cells_put<HS>(tile, tid, pre);
    __syncthreads();
    // opaque to the compiler, so that its constant part is not split off into LDS offsets
    // (which the row-pair reads cannot hold)
    int li = (threadIdx.y + HS) * CG::LW + (threadIdx.x + HS);
    asm volatile("" : "+v"(li));
#pragma unroll
    for (int pr = 0; pr < 4; ++pr) {
        if (pr < 3) {
            if (INTERIOR)
                cells_fetch_interior<HS>(mv, pr + 1, tid, off, pre);
            else
                cells_fetch<HS, false>(mv, H, W, x0, y0, pr + 1, tid, pre);
        }
        const int o1 = pr, o2 = pr + 4;
        Contrib cc1[kCellRows], cc2[kCellRows];
        cell_terms<kCellTY * CG::LW>(tile, tile + CG::TS, q, o1, li, cc1);
        cell_terms<kCellTY * CG::LW>(tile + CG::TS, tile, q, o2, li, cc2);
#pragma unroll
        for (int rr = 0; rr < kCellRows; ++rr) {
            const int y = y0 + threadIdx.y + kCellTY * rr;
            Contrib c1 = cc1[rr], c2 = cc2[rr];
            if (!INTERIOR) {
                auto inside = [&](int b) {
                    const int bx = b & 0xffff, by = b >> 16;
                    return x >= bx && x <= W - bx && y >= by && y <= H - by;
                };
                const bool e1 = inside(q.exy[o1]), i1 = inside(q.ixy[o1]);
                const bool e2 = inside(q.exy[o2]), i2 = inside(q.ixy[o2]);
                if (!e1) c1.s = c1.d = 0.f;
                if (!i1) c1.t = c1.r = 0.f;
                if (!e2) c2.s = c2.d = 0.f;
                if (!i2) c2.t = c2.r = 0.f;
            }
            hold[rr][pr] = c2;
            accS[rr] += c1.s;
            accD[rr] += c1.d;
            accT[rr] += c1.t;
            accR[rr] += c1.r;
        }
        if (pr < 3) {
            __syncthreads();
            cells_put<HS>(tile, tid, pre);
            __syncthreads();
        }
    }
#pragma unroll
    for (int rr = 0; rr < kCellRows; ++rr) {
        const int y = y0 + threadIdx.y + kCellTY * rr;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            accS[rr] += hold[rr][k].s;
            accD[rr] += hold[rr][k].d;
            accT[rr] += hold[rr][k].t;
            accR[rr] += hold[rr][k].r;
        }
        if (INTERIOR || (x < W && y < H)) {
            const float inhib = accT[rr] + accR[rr];
            const float Dv = accD[rr] + (-inhib);
            Dm[(size_t)y * W + x] = Dv;
            Sm[(size_t)y * W + x] = accS[rr] + (-inhib);
            Dmax = any ? fmaxf(Dmax, Dv) : Dv;
            any = true;
        }
    }
}

// Workgroups are dealt round-robin to the 8 XCDs; tiles are renumbered so that each XCD
// walks a contiguous raster band and the halos it re-reads stay in its own L2.
__device__ __forceinline__ int xcd_band_index(int b, int n) {
    const int per = (n + 7) / 8;
    const int t = (b % 8) * per + b / 8;
    return t < n ? t : -1;
}

// Tile geometry of one scale: tiles_x x tiles_y tiles, the interior ones (INTERIOR above)
// forming the rectangle [ix0, ix1) x [iy0, iy1).
struct CellTiles {
    int tiles_x, tiles_y, ix0, ix1, iy0, iy1;
    __device__ __forceinline__ int n_interior() const { return (ix1 - ix0) * (iy1 - iy0); }
    __device__ __forceinline__ int n_border() const {
        return tiles_x * tiles_y - n_interior();
    }
    // k-th border tile: rows above, rows below, then the left/right columns in between
    __device__ __forceinline__ void border_tile(int k, int& tx, int& ty) const {
        const int ntop = iy0 * tiles_x, nbot = (tiles_y - iy1) * tiles_x;
        if (k < ntop) {
            tx = k % tiles_x;
            ty = k / tiles_x;
        } else if (k < ntop + nbot) {
            k -= ntop;
            tx = k % tiles_x;
            ty = iy1 + k / tiles_x;
        } else {
            k -= ntop + nbot;
            const int side = ix0 + (tiles_x - ix1);
            ty = iy0 + k / side;
            const int c = k % side;
            tx = c < ix0 ? c : ix1 + (c - ix0);
        }
    }
};

// One scale of a cells launch: workgroups [block0, block0 + grid) of the launch, the first
// n_border of them on the border tiles (checked path), the rest on the interior rectangle
// (unchecked path, XCD-banded).
struct CellJob {
    MapView mv;
    int H, W;
    CellTileOffsets q;
    CellTiles tg;
    float* Dm;
    float* Sm;
    unsigned* maxslot;
    int block0, grid;
};
struct CellJobs {
    CellJob j[kCellMaxScales];
    int n;
};

// Scales that are ready together share one launch (a job per scale, block0 ascending), so
// that the small ones fill the last round of the large ones instead of leaving CUs idle.
template <int HS>
__global__ void __launch_bounds__(512) k_cells(const CellJobs jobs) {
    extern __shared__ float4 tile4[];
    float* tile = reinterpret_cast<float*>(tile4);
    const int tid = threadIdx.y * kCellTW + threadIdx.x;
    int js = jobs.n - 1;
    while (js > 0 && (int)blockIdx.x < jobs.j[js].block0) --js;  // uniform
    const CellJob& jb = jobs.j[js];
    const int b = (int)blockIdx.x - jb.block0;
    if (b >= jb.grid) return;  // alignment padding between jobs: uniform, before any barrier
    const MapView& mv = jb.mv;
    const CellTiles& tg = jb.tg;
    const int H = jb.H, W = jb.W;
    float Dmax = -INFINITY;
    bool any = false;
    const int nb = tg.n_border();
    if (b < nb) {
        int tx, ty;
        tg.border_tile(b, tx, ty);
        cells_tile<HS, false>(mv, H, W, jb.q, tx * kCellTW, ty * kCellTH, tile, jb.Dm, jb.Sm,
                              Dmax, any);
    } else {
        const int t = xcd_band_index(b - nb, jb.grid - nb);
        const int nx = tg.ix1 - tg.ix0;
        if (t >= 0 && t < tg.n_interior())
            cells_tile<HS, true>(mv, H, W, jb.q, (tg.ix0 + t % nx) * kCellTW,
                                 (tg.iy0 + t / nx) * kCellTH, tile, jb.Dm, jb.Sm, Dmax, any);
    }
    unsigned* __restrict__ maxslot = jb.maxslot;
    // block maximum of D -> one atomic per workgroup
    unsigned key = any ? float_key(Dmax) : 0u;
#pragma unroll
    for (int sh = 32; sh >= 1; sh >>= 1) {
        const unsigned other = __shfl_xor(key, sh, 64);
        key = other > key ? other : key;
    }
    __shared__ unsigned wmax[kCellTY];
    if ((tid & 63) == 0) wmax[tid >> 6] = key;
    __syncthreads();
    if (tid == 0) {
        unsigned m = 0;
        for (int i = 0; i < kCellTY; ++i) m = wmax[i] > m ? wmax[i] : m;
        if (m) atomicMax(maxslot, m);
    }
}

int cells_halo(const CellParams& p) {
    int h = 0;
    for (int o = 0; o < kNumOri; ++o) {
        h = max(h, 2 * max(abs(p.ea[o]), abs(p.eb[o])));
        h = max(h, 2 * max(abs(p.ia[o]), abs(p.ib[o])));
    }
    return h;
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
// Evaluated as D > thresh first (one load, rarely true), then the rest:
__device__ __forceinline__ bool nms_rest(const BIMP_GLOBAL float* D, int H, int W, int off,
                                         float v, int y, int x) {
#pragma clang fp contract(off)
    if (!(x > off && x < W - off && y > off && y < H - off)) return false;
    const BIMP_GLOBAL float* p = D + (size_t)y * W;
    const BIMP_GLOBAL float* pu = p - W;
    const BIMP_GLOBAL float* pd = p + W;
    float nb = fmaxf(pu[x - 1], pu[x]);
    nb = fmaxf(nb, pu[x + 1]);
    nb = fmaxf(nb, p[x - 1]);
    nb = fmaxf(nb, p[x + 1]);
    nb = fmaxf(nb, pd[x - 1]);
    nb = fmaxf(nb, pd[x]);
    nb = fmaxf(nb, pd[x + 1]);
    return v - nb > 0.f;
}

// Hit mask of a thread's kNmsPerThread pixels (bit i: pixel base + i*kNmsThreads + tid).
// All the pixels' loads go out first; the row/column split and the neighbourhood are only
// computed for the few above the threshold (about 2% of the pixels), each lane walking its
// own candidates, so that a wavefront takes as many neighbourhood rounds as its busiest lane
// has candidates rather than one per pixel slot that any lane needs.
__device__ __forceinline__ unsigned nms_mask(const NmsScale& d, float thresh, int base) {
    const int A = d.H * d.W;
    const BIMP_GLOBAL float* D = (const BIMP_GLOBAL float*)d.D;
    float v[kNmsPerThread];
#pragma unroll
    for (int i = 0; i < kNmsPerThread; ++i) {
        const int pix = base + i * kNmsThreads + (int)threadIdx.x;
        v[i] = pix < A ? D[pix] : 0.f;
    }
    unsigned cand = 0;
#pragma unroll
    for (int i = 0; i < kNmsPerThread; ++i) {
        const int pix = base + i * kNmsThreads + (int)threadIdx.x;
        if (pix < A && v[i] > thresh) cand |= 1u << i;
    }
    unsigned mask = 0;
    while (cand) {
        const int i = __builtin_ctz(cand);
        cand &= cand - 1u;
        const int pix = base + i * kNmsThreads + (int)threadIdx.x;
        const int y = pix / d.W, x = pix - y * d.W;
        if (nms_rest(D, d.H, d.W, d.off, D[pix], y, x)) mask |= 1u << i;
    }
    return mask;
}

__device__ __forceinline__ float scale_thresh(const unsigned* maxslots, int slot, double thr) {
    const float mx = key_float(maxslots[slot]);
    return (float)((double)mx * thr);
}

__global__ void __launch_bounds__(kNmsThreads) k_nms_count(const NmsScale* __restrict__ sc,
                                                           int nscales,
                                                           const unsigned* __restrict__ maxslots,
                                                           double thr, int* __restrict__ counts,
                                                           unsigned* __restrict__ masks) {
    const int blk = blockIdx.x;
    const int s = find_scale(sc, nscales, blk);
    const NmsScale d = sc[s];
    const float thresh = scale_thresh(maxslots, d.slot, thr);
    const int base = (blk - d.block0) * kNmsBlockPixels;
    const unsigned mask = nms_mask(d, thresh, base);
    ((BIMP_GLOBAL unsigned*)masks)[(size_t)blk * kNmsThreads + threadIdx.x] = mask;
    int c = __popc(mask);
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

// Writes each keypoint of k_nms_count's hit masks at its row-major rank.  A block's first
// rank is the sum of the counts of the blocks before it, summed here (a few thousand
// cache-resident ints per block with keypoints) rather than by a scan launch; the last block
// also stores the frame's total in counter and clears the per-scale max slots (the count pass,
// their only reader, is done), so that the next frame starts from zero.  matrix_ld > 0
// selects the 7-row SURF_CUDA layout (cuda_bimp.cuh:132-140) instead of keypoint structs.
__global__ void __launch_bounds__(kNmsThreads) k_nms_write(
        const NmsScale* __restrict__ sc, int nscales, const int* __restrict__ counts,
        int total_blocks, int* __restrict__ counter, unsigned* __restrict__ maxslots,
        const unsigned* __restrict__ masks, void* __restrict__ out, int cap, int matrix_ld) {
#pragma clang fp contract(off)
    const int blk = blockIdx.x;
    const bool last = blk == total_blocks - 1;
    const int mine = counts[blk];
    if (mine == 0 && !last) return;  // uniform: no keypoints, no total to store
    constexpr int NW = kNmsThreads / 64;
    __shared__ int ws[NW];
    int pre = 0;
    for (int i = (int)threadIdx.x; i < blk; i += kNmsThreads) pre += counts[i];
#pragma unroll
    for (int sh = 32; sh >= 1; sh >>= 1) pre += __shfl_xor(pre, sh, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = pre;
    __syncthreads();
    int start = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) start += ws[i];
    if (last) {
        if (threadIdx.x == 0) *counter = start + mine;
        for (int i = (int)threadIdx.x; i < nscales; i += kNmsThreads) maxslots[i] = 0u;
    }
    // a block without keypoints, or whose first rank is past the capacity, writes nothing
    if (mine == 0 || start >= cap) return;  // uniform
    const int s = find_scale(sc, nscales, blk);
    const NmsScale d = sc[s];
    const int base = (blk - d.block0) * kNmsBlockPixels;
    // rank of pixel (i, thread) = keypoints of earlier iterations, then of earlier waves in
    // iteration i, then of earlier lanes: the row-major order of the pixels
    __shared__ int wc[kNmsPerThread * NW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned mask =
        ((const BIMP_GLOBAL unsigned*)masks)[(size_t)blk * kNmsThreads + threadIdx.x];
    unsigned rank_lo[kNmsPerThread];
#pragma unroll
    for (int i = 0; i < kNmsPerThread; ++i) {
        const unsigned long long bal = __ballot((mask >> i) & 1u);
        rank_lo[i] = (unsigned)__popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wc[i * NW + wave] = __popcll(bal);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = start;
        for (int j = 0; j < kNmsPerThread * NW; ++j) {
            const int c = wc[j];
            wc[j] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kNmsPerThread; ++i) {
        if ((mask >> i) & 1u) {
            const int pix = base + i * kNmsThreads + (int)threadIdx.x;
            const int y = pix / d.W, x = pix - y * d.W;
            const int idx = wc[i * NW + wave] + (int)rank_lo[i];
            if (idx < cap) {
                const BIMP_GLOBAL float* p = (const BIMP_GLOBAL float*)d.D + (size_t)y * d.W;
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

// 8-bit BGR -> gray with the fixed-point weights of OpenCV's COLOR_BGR2GRAY (0.114, 0.587,
// 0.299 in Q14: 1868, 9617, 4899; rounding by 2^13), as getGPUImage uses it
// (cuda_bimp.cu:170-174).
__global__ void k_bgr_to_gray(const uint8_t* __restrict__ bgr, size_t bpitch, int w, int h,
                              uint8_t* __restrict__ gray, size_t gpitch) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const uint8_t* p = bgr + (size_t)y * bpitch + 3 * (size_t)x;
    const unsigned v = 1868u * p[0] + 9617u * p[1] + 4899u * p[2] + (1u << 13);
    gray[(size_t)y * gpitch + x] = (uint8_t)(v >> 14);
}

__global__ void k_store_count(const int* __restrict__ counter, int* __restrict__ dst) {
    if (threadIdx.x == 0) dst[0] = counter[0];
}

__global__ void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

// ------------------------------------------------------------------ dispatch
#define BIMP_CELL_HALOS(X) X(4) X(8) X(12) X(16) X(20)

// Host construction of the packed twiddle table of FFTCfg<N> (layout: TwLayout).
template <class L, int Ns>
struct TwBuild;
template <int R, int... Rest, int Ns>
struct TwBuild<RadList<R, Rest...>, Ns> {
    static void run(std::vector<float2>& out, bool first) {
        if (!first) {
            for (int k = 0; k < Ns; ++k) {
                for (int r = 1; r < R; ++r) {
                    const double a = -2.0 * 3.14159265358979323846 * (double)r * k / ((double)Ns * R);
                    out.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
                }
                for (int r = R - 1; r < tw_row(R); ++r) out.push_back(make_float2(0.f, 0.f));
            }
        }
        TwBuild<RadList<Rest...>, Ns * R>::run(out, false);
    }
};
template <int Ns>
struct TwBuild<RadList<>, Ns> {
    static void run(std::vector<float2>&, bool) {}
};

template <int N>
static void twiddles_n(std::vector<float2>& out) {
    out.clear();
    TwBuild<typename FFTCfg<N>::Rad, 1>::run(out, true);
    if (out.empty()) out.push_back(make_float2(1.f, 0.f));
}

bool twiddle_table(int n, std::vector<float2>& out) {
#define X(N) if (n == N) { twiddles_n<N>(out); return true; }
    BIMP_FOR_SIZES(X)
#undef X
    return false;
}

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

// Resident workgroups of k_inv_cols<N> on the whole device (one round), a multiple of 8 so
// that the unit stride keeps each unit on its XCD; computed once per size.
template <int N>
static int inv_cols_slots() {
    static int slots = [] {
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_inv_cols<N>, InvColCfg<N>::THREADS,
                                                         InvColCfg<N>::LDS) != hipSuccess)
            return 0;
        return std::max(1, cus * per / 8) * 8;
    }();
    return slots;
}

// Resident workgroups of k_inv_rows_p<N> on the whole device (one round), a multiple of 8.
template <int N>
static int inv_rows_slots() {
    static int slots = [] {
        if constexpr (!InvRowCfg<N>::PERSISTENT) {
            return 0;
        } else {
            int dev = 0, cus = 0, per = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipFuncSetAttribute((const void*)k_inv_rows_p<N>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)InvRowCfg<N>::LDS) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_inv_rows_p<N>,
                                                             InvRowCfg<N>::THREADS,
                                                             InvRowCfg<N>::LDS) != hipSuccess)
                return 0;
            return std::max(1, cus * per / 8) * 8;
        }
    }();
    return slots;
}

template <int N>
static hipError_t prepare_n() {
    hipError_t e;
    const int rowLds = RowCfg<N>::VPB * N * (int)sizeof(float2);  // fwd rows stage complex output
    const int colLds = ColCfg<N>::CB * ColCfg<N>::S * (int)sizeof(float2);
    if ((e = hipFuncSetAttribute((const void*)k_fwd_rows<N, uint8_t>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, rowLds)) != hipSuccess)
        return e;
    if ((e = hipFuncSetAttribute((const void*)k_fwd_rows<N, float>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, rowLds)) != hipSuccess)
        return e;
    if ((e = hipFuncSetAttribute((const void*)k_inv_rows<N>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, rowLds)) != hipSuccess)
        return e;
    if ((e = hipFuncSetAttribute((const void*)k_fwd_cols<N>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, colLds)) != hipSuccess)
        return e;
    if ((e = hipFuncSetAttribute((const void*)k_inv_cols<N>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds)) !=
        hipSuccess)
        return e;
    if (InvRowCfg<N>::PERSISTENT && inv_rows_slots<N>() <= 0) return hipErrorInvalidValue;
    return inv_cols_slots<N>() > 0 ? hipSuccess : hipErrorInvalidValue;
}

hipError_t prepare_common() {
    hipError_t e = hipSuccess;
#define X(HS)                                                                                \
    if (e == hipSuccess)                                                                     \
        e = hipFuncSetAttribute((const void*)k_cells<HS>,                                    \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
    BIMP_CELL_HALOS(X)
#undef X
    return e;
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
                           int dw, int dh, size_t dpitch, bool half_up, hipStream_t s) {
    dim3 grid((dw + 127) / 128, dh);
    k_pyr_down<<<grid, 128, 0, s>>>(src, sw, sh, spitch, dst, dw, dh, dpitch, half_up ? 1 : 0);
    return hipGetLastError();
}

template <class Pix>
static hipError_t fwd_rows(int FW, const Pix* img, size_t pitch, int H, int W, const ScaleSet& ss,
                           float scale, float2* R, int ldr, size_t rstride, const float2* tw,
                           hipStream_t s) {
#define X(N)                                                                                 \
    if (FW == N) {                                                                           \
        const int vpb = RowCfg<N>::VPB;                                                      \
        const int pairs = (H + 1) / 2;                                                       \
        dim3 grid((pairs + vpb - 1) / vpb, ss.n);                                            \
        k_fwd_rows<N, Pix><<<grid, RowCfg<N>::THREADS, vpb * N * sizeof(float2), s>>>(       \
            img, pitch, H, W, ss, scale, R, ldr, rstride, tw);                               \
        return hipGetLastError();                                                            \
    }
    BIMP_FOR_SIZES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_fwd_rows(int FW, const uint8_t* img, size_t pitch, int H, int W,
                           const ScaleSet& ss, float scale, float2* R, int ldr, size_t rstride,
                           const float2* tw, hipStream_t s) {
    return fwd_rows(FW, img, pitch, H, W, ss, scale, R, ldr, rstride, tw, s);
}

hipError_t launch_fwd_rows_f32(int FW, const float* img, size_t pitch, int H, int W,
                               const ScaleSet& ss, float scale, float2* R, int ldr,
                               size_t rstride, const float2* tw, hipStream_t s) {
    return fwd_rows(FW, img, pitch, H, W, ss, scale, R, ldr, rstride, tw, s);
}

hipError_t launch_fwd_cols(int FH, const float2* R, int ldr, size_t rstride, int H,
                           const ScaleSet& ss, int ncols, const float2* tw, hipStream_t s) {
#define X(N)                                                                                 \
    if (FH == N) {                                                                           \
        using CC = ColCfg<N>;                                                                \
        dim3 grid((ncols + CC::CB - 1) / CC::CB, ss.n);                                      \
        k_fwd_cols<N><<<grid, CC::THREADS, CC::CB * CC::S * sizeof(float2), s>>>(            \
            R, ldr, rstride, H, ss, ncols, tw);                                              \
        return hipGetLastError();                                                            \
    }
    BIMP_FOR_SIZES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_inv_cols(int FH, const ScaleSet& ss, int FW, int Hc, float2* T,
                           size_t tstride, const float2* tw, hipStream_t s) {
#define X(N)                                                                                 \
    if (FH == N) {                                                                           \
        using IC = InvColCfg<N>;                                                             \
        const size_t lds = IC::LDS;                                                          \
        if (lds > kMaxLds || Hc > N || FW % IC::CB) return hipErrorInvalidValue;            \
        const int nblk = (FW / 2 + IC::CB) / IC::CB;                                         \
        const int units = (nblk * ss.n + 7) / 8 * 8 * 2 * kNumOri;                           \
        const int grid = std::min(units, inv_cols_slots<N>());                               \
        if (grid <= 0) return hipErrorInvalidValue;                                          \
        k_inv_cols<N><<<grid, IC::THREADS, lds, s>>>(ss, FW, Hc, nblk, units, T, tstride,    \
                                                     tw);                                    \
        return hipGetLastError();                                                            \
    }
    BIMP_FOR_SIZES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_inv_rows(int FW, int FH, const ScaleSet& ss, const float2* T, size_t tstride,
                           int Hc, int Wc, int ldc, const float2* tw, hipStream_t s) {
    // the column pass's block width (columns per T block) for transforms of length FH
    const int cb = col_batch(FH);
    int tsh = 0;
    while ((1 << tsh) < cb) ++tsh;
    if (cb <= 0 || (1 << tsh) != cb || FW % cb) return hipErrorInvalidValue;
#define X(N)                                                                                 \
    if (FW == N) {                                                                           \
        if constexpr (InvRowCfg<N>::PERSISTENT) {                                            \
            using RC = InvRowCfg<N>;                                                         \
            const int units = (Hc + RC::RPU - 1) / RC::RPU * kNumOri * ss.n;                 \
            const int grid = std::min(units, inv_rows_slots<N>());                           \
            if (grid <= 0) return hipErrorInvalidValue;                                      \
            k_inv_rows_p<N><<<grid, RC::THREADS, RC::LDS, s>>>(ss, T, tstride, Hc, Wc, ldc,  \
                                                               tsh, units, tw);              \
            return hipGetLastError();                                                        \
        }                                                                                    \
        const int vpb = RowCfg<N>::VPB;                                                      \
        dim3 grid((Hc + vpb - 1) / vpb, kNumOri, ss.n);                                      \
        k_inv_rows<N><<<grid, RowCfg<N>::THREADS, vpb * fft_lds_floats<N>() * sizeof(float), s>>>( \
            ss, T, tstride, Hc, Wc, ldc, tsh, tw);                                           \
        return hipGetLastError();                                                            \
    }
    BIMP_FOR_SIZES(X)
#undef X
    return hipErrorInvalidValue;
}

template <int HS>
static void make_cell_job(const MapView& mv, int FH, int FW, int H, int W, const CellParams& p,
                          float* Dm, float* Sm, unsigned* maxslot, CellJob& job) {
    using CG = CellGeom<HS>;
    CellTileOffsets q;
    for (int o = 0; o < kNumOri; ++o) {
        const int a = p.ea[o], b = p.eb[o];
        q.es[o] = -a * CG::LW + b;
        q.tn[o] = -p.ia[o] * CG::LW - p.ib[o];
        q.rp[o] = p.ra[o] * CG::LW + p.rb[o];
        // (bounds are at most twice an offset within the halo: far below 2^15)
        q.exy[o] = 2 * abs(b) | (2 * abs(a)) << 16;
        q.ixy[o] = 2 * abs(p.ib[o]) | (2 * abs(p.ia[o])) << 16;
    }
    CellTiles tg;
    tg.tiles_x = (W + kCellTW - 1) / kCellTW;
    tg.tiles_y = (H + kCellTH - 1) / kCellTH;
    // interior tiles: x0 >= HS, x0 + TW - 1 <= W - HS (same in y), no wrap-around reads
    auto first = [](int hh, int t) { return (hh + t - 1) / t; };
    auto end = [](int lim, int t) { return lim >= 0 ? lim / t + 1 : 0; };  // x0 <= lim
    tg.ix0 = first(HS, kCellTW);
    tg.ix1 = std::max(tg.ix0, std::min(tg.tiles_x, end(W - HS - (kCellTW - 1), kCellTW)));
    tg.iy0 = first(HS, kCellTH);
    tg.iy1 = std::max(tg.iy0, std::min(tg.tiles_y, end(H - HS - (kCellTH - 1), kCellTH)));
    // (interior staging offsets span 5 maps from the pair's base: 8 maps must fit 32 bits)
    if (W >= FW || H >= FH || tg.ix0 >= tg.tiles_x || tg.iy0 >= tg.tiles_y ||
        (size_t)kNumOri * mv.Hc * mv.ldc * sizeof(float) > 0xffffffffu)
        tg.ix0 = tg.ix1 = tg.iy0 = tg.iy1 = 0;  // everything takes the checked path
    const int nin = (tg.ix1 - tg.ix0) * (tg.iy1 - tg.iy0);
    const int nbd = tg.tiles_x * tg.tiles_y - nin;
    job.mv = mv;
    job.H = H;
    job.W = W;
    job.q = q;
    job.tg = tg;
    job.Dm = Dm;
    job.Sm = Sm;
    job.maxslot = maxslot;
    job.grid = nbd + (nin + 7) / 8 * 8;
}

template <int HS>
static hipError_t launch_cells_hs(const CellScale* sc, int n, hipStream_t s) {
    using CG = CellGeom<HS>;
    const size_t lds = (size_t)2 * CG::TS * sizeof(float);
    CellJobs jobs;
    jobs.n = n;
    int blocks = 0;
    for (int k = 0; k < n; ++k) {
        const CellScale& c = sc[k];
        make_cell_job<HS>(MapView{c.C, c.Hc, c.ldc, c.FH, c.FW}, c.FH, c.FW, c.H, c.W, c.p,
                          c.Dm, c.Sm, c.maxslot, jobs.j[k]);
        jobs.j[k].block0 = blocks;
        blocks += (jobs.j[k].grid + 7) / 8 * 8;  // each job starts on a multiple of 8
    }
    k_cells<HS><<<blocks, dim3(kCellTW, kCellTY), lds, s>>>(jobs);
    return hipGetLastError();
}

hipError_t launch_cells(const CellScale* sc, int n, hipStream_t s) {
    if (n < 1 || n > kCellMaxScales) return hipErrorInvalidValue;
    // one halo for the launch: the largest the scales need (a larger one stages more, with
    // the same results)
    int h = 0;
    for (int k = 0; k < n; ++k) h = std::max(h, cells_halo(sc[k].p));
    const int hs = (h + 3) / 4 * 4;
#define X(HS) \
    if (hs <= HS) return launch_cells_hs<(HS > 0 ? HS : 4)>(sc, n, s);
    BIMP_CELL_HALOS(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_cells(const float* C, int Hc, int ldc, int FH, int FW, int H, int W,
                        const CellParams& p, float* Dm, float* Sm, unsigned* maxslot,
                        hipStream_t s) {
    const CellScale one{C, Hc, ldc, FH, FW, H, W, p, Dm, Sm, maxslot};
    return launch_cells(&one, 1, s);
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
                            unsigned* masks, hipStream_t s) {
    k_nms_count<<<total_blocks, kNmsThreads, 0, s>>>(scales, nscales, maxslots, thresh,
                                                     block_counts, masks);
    return hipGetLastError();
}

hipError_t launch_nms_write(const NmsScale* scales, int nscales, int total_blocks,
                            const int* block_counts, int* counter, unsigned* maxslots,
                            const unsigned* masks, void* out, int cap, int matrix_ld,
                            hipStream_t s) {
    if (total_blocks < 1) return hipErrorInvalidValue;
    k_nms_write<<<total_blocks, kNmsThreads, 0, s>>>(scales, nscales, block_counts, total_blocks,
                                                     counter, maxslots, masks, out, cap,
                                                     matrix_ld);
    return hipGetLastError();
}

hipError_t launch_bgr_to_gray(const uint8_t* bgr, size_t bpitch, int w, int h, uint8_t* gray,
                              size_t gpitch, hipStream_t s) {
    k_bgr_to_gray<<<dim3((w + 255) / 256, h), 256, 0, s>>>(bgr, bpitch, w, h, gray, gpitch);
    return hipGetLastError();
}

hipError_t launch_store_count(const int* counter, int* dst, hipStream_t s) {
    k_store_count<<<1, 64, 0, s>>>(counter, dst);
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
