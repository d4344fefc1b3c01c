// direct.hip -- precision form: small-wavelength scales by direct Gabor convolution on MFMA.
//
// For a scale marked "direct", the complex-cell maps C_o are computed without FFTs:
//   e_o + i*o_o = conv(Pd, Ke_o + i*Ko_o)   over the periodic padded image Pd (as the FFT
//                                            path, so wrap-around semantics match)
// (kernels: cuda_bimp.cu:304-331).  The 16 filters are convolved as an implicit GEMM in BF16
// or FP16 with FP32 accumulation (v_mfma_f32_16x16x32_{bf16,f16}): taps scaled by 1/fscale
// and rounded once, image samples entering as their uint8 values (exact in both formats).
// With truncated taps the even filter is the carrier Ke_o + dc_o inside the window.  Either
// way each filter's response is corrected by -dc[f] * B, B the kW x kW box sum of Pd, with
// dc[f] chosen so that a constant image gets the exact full kernel's response: rounding and
// truncation then leave no DC leak.  B is summed in FP32 from the staged tile; every partial
// sum is an integer below 2^24, so it is exact.  The 255/65536 input scale and fscale are
// applied to the FP32 accumulators.
//
// GEMM shape: M = 16 pixels (consecutive columns), N = 16 filters (even/odd x 8
// orientations), K = 32 taps = 4 tap rows x 8 tap columns, one tap row per 16-lane group.
// A K-step is keyed by an aligned block of 4 image rows: the image fragment of rows
// 4*rho .. 4*rho+3 serves every output row j of the wavefront, with the tap block
// u0 = 4*rho - j (filter fragments precomputed for every u0).  So each fragment read from
// LDS feeds kDirRows x kDirGroups MFMAs and each filter fragment kDirGroups of them.
// The image tile sits in LDS as two copies shifted by 0 and 1 sample: every lane's 8-sample
// window then starts on a 4-byte boundary in one of them and is four 4-byte reads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

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

// Row (or column) of the periodic padded image (padDataClampToBorder rule,
// convolutionFFT2D.cuh:56-85): index i of the F-periodic domain, data size n, border k.
__device__ __forceinline__ int pad_index(int i, int n, int F, int k) {
    if (i < 0 && i >= -F)
        i += F;
    else if (i >= F && i < 2 * F)
        i -= F;
    else if (i < 0 || i >= F) {
        i %= F;
        if (i < 0) i += F;
    }
    return i < n ? i : (i < n + k ? n - 1 : 0);
}

// One launch covers every scale of a job table (scales of one pyramid level, heaviest first):
// workgroup b works on the scale whose tile range [tile0, tile0 + tiles) holds b.
// Two wavefronts per SIMD whatever the workgroup size: 4-wave workgroups run two to a CU.
template <class T, int NW>
__global__ void __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(2, 2))) k_direct(
        const DirectLevels lv, const DirectJob* __restrict__ jobs, int njobs) {
    typedef typename Mfma<T>::v8 v8;
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    constexpr int M = kDirRows, NCG = kDirGroups, TY = kDirTY, TX = kDirTX, NT = NW * 64;
    // output rows per wavefront, in passes of M rows each
    constexpr int PASSES = TY / (M * NW);
    static_assert(PASSES * M * NW == TY, "tile rows split evenly over the wavefronts");
    int js = njobs - 1;
    while (js > 0 && (int)blockIdx.x < jobs[js].tile0) --js;  // uniform
    const DirectJob& job = jobs[js];
    const BIMP_GLOBAL uint8_t* __restrict__ img = (const BIMP_GLOBAL uint8_t*)lv.img[job.level];
    const size_t pitch = lv.pitch[job.level];
    const DirectScale& ds = job.ds;
    const BIMP_GLOBAL v8* __restrict__ bfrag = (const BIMP_GLOBAL v8*)job.frag;
    BIMP_GLOBAL float* __restrict__ Cm = (BIMP_GLOBAL float*)job.C;
    const int tb = (int)blockIdx.x - job.tile0;
    const int tby = tb / job.tiles_x, tbx = tb - tby * job.tiles_x;
    extern __shared__ unsigned char smem[];
    T* copies = reinterpret_cast<T*>(smem);
    const int LWP = ds.lwp, csize = ds.cstride;
    float* vs = reinterpret_cast<float*>(smem + (size_t)2 * csize * sizeof(T));
    const int k = ds.kpad, kW = ds.kW;
    const int TWi = TX + 2 * k;  // columns of the image window the tile needs
    const int x0 = tbx * TX, y0 = tby * TY;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;

    // Stage rows y0-k .. and columns x0-k .. of Pd as two copies, the second shifted by one
    // sample (tile sample c at position c - s of copy s), so that every lane's 8-sample window
    // starts on a 4-byte boundary in one of them.  Positions past the image window hold zero;
    // they only meet zero taps.  Wave w takes rows w, w + NW, ..., lane l the four columns
    // 4l .. 4l+3, written as one 8-byte store per copy (the shifted copy's fourth sample is
    // the next lane's first); the byte loads of kStageRows rows are all issued before any is
    // written, so that their latencies overlap instead of adding up row by row.
    const bool inside = y0 - k >= 0 && y0 + TY + k <= ds.H && x0 - k >= 0 && x0 + TX + k <= ds.W;
    {
        constexpr int kStageRows = 4;
        static_assert(kDirMaxLwp == 4 * 64, "one group of four staged columns per lane");
        typedef typename Mfma<T>::v4 v4;
        // source columns of staged columns c0 .. c0+3 (-1: zero), the same for every row
        const int c0 = 4 * lane;
        int gxs[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = c0 + j;
            gxs[j] = c >= TWi ? -1 : (inside ? x0 - k + c : pad_index(x0 - k + c, ds.W, ds.FW, k));
        }
        for (int r0 = wave; r0 < ds.LH; r0 += kStageRows * NW) {
            float v[kStageRows][4];
#pragma unroll
            for (int q = 0; q < kStageRows; ++q) {
                const int r = r0 + q * NW;
                const bool rowin = r < TY + 2 * k;
                const int gy = !rowin ? 0
                                      : (inside ? y0 - k + r : pad_index(y0 - k + r, ds.H, ds.FH, k));
                const BIMP_GLOBAL uint8_t* row = img + (size_t)gy * pitch;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    v[q][j] = (rowin && gxs[j] >= 0) ? (float)row[gxs[j]] : 0.f;
            }
#pragma unroll
            for (int q = 0; q < kStageRows; ++q) {
                const int r = r0 + q * NW;
                if (r >= ds.LH) break;  // uniform
                // sample c0 + 4: the next lane's first (column 256 and beyond: zero)
                const float next = __shfl_down(v[q][0], 1, 64);
                const float nx = lane == 63 ? 0.f : next;
                if (c0 < LWP) {  // LWP is a multiple of 16: the lane's four positions fit
                    T* dst = copies + r * LWP + c0;
                    const v4 a = {T(v[q][0]), T(v[q][1]), T(v[q][2]), T(v[q][3])};
                    const v4 b = {T(v[q][1]), T(v[q][2]), T(v[q][3]), T(nx)};
                    *reinterpret_cast<v4*>(dst) = a;          // copy 0: position c = sample c
                    *reinterpret_cast<v4*>(dst + csize) = b;  // copy 1: position c = sample c+1
                }
            }
        }
    }
    __syncthreads();

    // kW x kW box sums of the tile's output pixels: vertical running sums over kW rows into
    // vs[r][c] (two segments of TY/2 rows per column when the workgroup has a thread for each,
    // else one of TY rows), then horizontal ones, stored back into vs as a TY x TX array.
    {
        const int nseg = 2 * TWi <= NT ? 2 : 1;
        const int rows = TY / nseg;
        for (int t = tid; t < nseg * TWi; t += NT) {
            const int seg = t >= TWi ? 1 : 0;
            const int c = t - seg * TWi, r0 = seg * rows;
            const T* col = copies + c;
            // integer-valued partial sums below 2^24: exact in FP32 in any order
            float s4[4] = {0.f, 0.f, 0.f, 0.f};
            int u = 0;
            for (; u + 4 <= kW; u += 4) {
#pragma unroll
                for (int i = 0; i < 4; ++i) s4[i] += (float)col[(r0 + u + i) * LWP];
            }
            for (; u < kW; ++u) s4[0] += (float)col[(r0 + u) * LWP];
            float sum = (s4[0] + s4[1]) + (s4[2] + s4[3]);
            vs[r0 * ds.vsp + c] = sum;
#pragma unroll 4
            for (int r = 1; r < rows; ++r) {
                sum += (float)col[(r0 + r + kW - 1) * LWP] - (float)col[(r0 + r - 1) * LWP];
                vs[(r0 + r) * ds.vsp + c] = sum;
            }
        }
    }
    __syncthreads();
    constexpr int HP = TY * (TX / 8) / NT;  // horizontal-sum items (8 columns each) per thread
    float bsum[HP][8];
#pragma unroll
    for (int h = 0; h < HP; ++h) {
        const int t = tid + h * NT, r = t >> 4, xs = (t & 15) * 8;
        const float* vr = vs + r * ds.vsp + xs;
        float s4[4] = {0.f, 0.f, 0.f, 0.f};
        int v = 0;
        for (; v + 4 <= kW; v += 4) {
#pragma unroll
            for (int i = 0; i < 4; ++i) s4[i] += vr[v + i];
        }
        for (; v < kW; ++v) s4[0] += vr[v];
        float sum = (s4[0] + s4[1]) + (s4[2] + s4[3]);
        bsum[h][0] = sum;
#pragma unroll
        for (int i = 1; i < 8; ++i) {
            sum += vr[i + kW - 1] - vr[i - 1];
            bsum[h][i] = sum;
        }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < HP; ++h) {
        const int t = tid + h * NT, r = t >> 4, xs = (t & 15) * 8;
        float4* dstb = reinterpret_cast<float4*>(vs + r * TX + xs);
        dstb[0] = make_float4(bsum[h][0], bsum[h][1], bsum[h][2], bsum[h][3]);
        dstb[1] = make_float4(bsum[h][4], bsum[h][5], bsum[h][6], bsum[h][7]);
    }
    __syncthreads();  // the box sums are read by the epilogues

    // Main loop.  Lane (p, kb): pixel column p of each 16-column group, tap row kb of the
    // K-step.  The copy s = (off + p) & 1 is fixed per lane; its window starts on an even
    // position, so a fragment is four 4-byte LDS reads.
    const int p = lane & 15, kb = lane >> 4;
    const int s = (ds.off + p) & 1;
    const int nvb = ds.nvb, nsteps = ds.nrho * nvb;
    // step cursor: (rho, vb) and the element offsets of its image and filter fragments
    struct Cursor {
        int vb, aoff, boff;
    };
    // filter fragments of u0 = 4*rho - j sit at row index u0 + M - 1
    const BIMP_GLOBAL v8* __restrict__ bbase = bfrag + (size_t)(M - 1) * nvb * 64 + lane;
    auto advance = [&](Cursor& c) {
        if (++c.vb == nvb) {
            c.vb = 0;
            c.aoff += 4 * LWP - 8 * (nvb - 1);
            c.boff += (3 * nvb + 1) * 64;
        } else {
            c.aoff += 8;
            c.boff += 64;
        }
    };
    auto load_b = [&](const Cursor& c, v8* b) {
        const BIMP_GLOBAL v8* bro = bbase + c.boff;
#pragma unroll
        for (int j = 0; j < M; ++j) b[j] = bro[-j * nvb * 64];
    };
    const int jf = lane & 15, o = jf >> 1;
    const float dce = ds.dc[2 * o], dco = ds.dc[2 * o + 1];
#pragma unroll 1
    for (int pass = 0; pass < PASSES; ++pass) {
        const int row0 = (wave * PASSES + pass) * M;  // first tile row of this pass
        const T* lanebase = copies + s * csize + (ds.off + p - s) + (row0 + ds.off + kb) * LWP;
        auto load_a = [&](const Cursor& c, v8* a) {
            const unsigned* rowp = reinterpret_cast<const unsigned*>(lanebase + c.aoff);
#pragma unroll
            for (int g = 0; g < NCG; ++g) {
                const u32x4 w = {rowp[8 * g], rowp[8 * g + 1], rowp[8 * g + 2], rowp[8 * g + 3]};
                a[g] = __builtin_bit_cast(v8, w);
            }
        };
        f32x4 acc[M][NCG];
#pragma unroll
        for (int j = 0; j < M; ++j)
#pragma unroll
            for (int g = 0; g < NCG; ++g) acc[j][g] = f32x4{0.f, 0.f, 0.f, 0.f};
        // Software-pipelined over the K-steps it = rho * nvb + vb with two fragment sets used
        // in turn (the loop body is two steps): while the MFMAs of one step run, the image
        // (LDS) and filter (L2) fragments of the next are loaded into the other set.  No
        // fragment register is copied, so no wait falls on loads still in flight for a later
        // step.  Fragment offsets advance incrementally (no division per step); the load after
        // the last step repeats the last step's (clamped cursor, unused).
        auto mfma_step = [&](const v8* a, const v8* b) {
#pragma unroll
            for (int j = 0; j < M; ++j)
#pragma unroll
                for (int g = 0; g < NCG; ++g) acc[j][g] = Mfma<T>::run(a[g], b[j], acc[j][g]);
        };
        v8 a0[NCG], b0[M], a1[NCG], b1[M];
        Cursor cl{0, 0, 0};
        load_a(cl, a0);
        load_b(cl, b0);
        int it = 0;
#pragma unroll 1
        for (; it + 1 < nsteps; it += 2) {
            advance(cl);  // step it + 1
            load_a(cl, a1);
            load_b(cl, b1);
            mfma_step(a0, b0);
            if (it + 2 < nsteps) advance(cl);  // step it + 2
            load_a(cl, a0);
            load_b(cl, b0);
            mfma_step(a1, b1);
        }
        if (it < nsteps) mfma_step(a0, b0);

        // Epilogue: lane holds filter jf = lane%16 for pixels 4*kb .. 4*kb+3 of each group;
        // the odd filter of orientation jf/2 sits in the neighbouring lane.  Over a pair of
        // groups (g2, g2+1) the even lane finishes all four pixels of g2 and the odd lane those
        // of g2+1, after swapping the other group's values, so that each lane stores a float4
        // and an orientation's 32 pixels of the pair go out together.
        static_assert(NCG % 2 == 0, "groups are finished in pairs");
        const bool oddf = (jf & 1) != 0;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const int yl = row0 + j, y = y0 + yl;
#pragma unroll
            for (int g2 = 0; g2 < NCG; g2 += 2) {
                const int gm = g2 + (oddf ? 1 : 0);  // the group this lane finishes
                const int xl = 16 * gm + 4 * kb, x = x0 + xl;
                const float4 bx = *reinterpret_cast<const float4*>(vs + yl * TX + xl);
                const float bxs[4] = {bx.x, bx.y, bx.z, bx.w};
                float c4[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float mine = oddf ? acc[j][g2 + 1][q] : acc[j][g2][q];
                    const float send = oddf ? acc[j][g2][q] : acc[j][g2 + 1][q];
                    // the neighbouring lane's value: DPP quad_perm [1, 0, 3, 2] (a VALU move,
                    // no LDS crossbar round trip)
                    const float other = __int_as_float(__builtin_amdgcn_update_dpp(
                        0, __float_as_int(send), 0xB1, 0xF, 0xF, false));
                    const float e = ds.fscale * (oddf ? other : mine) - dce * bxs[q];
                    const float od = ds.fscale * (oddf ? mine : other) - dco * bxs[q];
                    c4[q] = ds.iscale * __builtin_amdgcn_sqrtf(e * e + od * od);
                }
                if (y < ds.Hc) {
                    BIMP_GLOBAL float* out = Cm + ((size_t)o * ds.Hc + y) * ds.ldc + x;
                    if (x + 3 < ds.Wc) {
                        *reinterpret_cast<BIMP_GLOBAL f32x4*>(out) =
                            f32x4{c4[0], c4[1], c4[2], c4[3]};
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (x + q < ds.Wc) out[q] = c4[q];
                    }
                }
            }
        }
    }
}

}  // namespace

size_t direct_geometry(DirectScale& d) {
    const int R = d.ext / 2;
    const int extp = (d.ext + 7) / 8 * 8;
    d.off = d.kpad - R;
    d.nvb = extp / 8;
    d.nrho = (kDirRows - 1 + d.ext + 3) / 4;
    const int k = d.kpad;
    d.LH = std::max(kDirTY + 2 * k, (kDirTY - kDirRows) + d.off + 4 * d.nrho);
    const int twi = kDirTX + 2 * k;
    // widest position read (off + TX + extp - 2 in copy 0) plus one sample of slack, since
    // the shifted copy leaves its last position unwritten
    const int tws = std::max(twi, d.off + kDirTX + extp - 1) + 2;
    // row pitch 16 (mod 32) elements: the 4 tap rows of a read start 8 x (odd) banks apart
    d.lwp = tws + ((16 - tws % 32) + 32) % 32;
    d.vsp = twi;
    // copy stride 32 (mod 64) elements: the two copies sit 16 banks apart, so each 16-lane
    // group of a fragment read (8 consecutive words in each copy) hits distinct banks
    const int cs = d.LH * d.lwp;
    d.cstride = cs + ((32 - cs % 64) + 64) % 64;
    return (size_t)2 * d.cstride * 2 + (size_t)kDirTY * d.vsp * sizeof(float);
}

// Workgroup size of a launch: 4 wavefronts (two passes of 4 rows each) when two workgroups
// fit one CU's LDS, so that one's staging and epilogue overlap the other's MFMA loop;
// otherwise 8 (one pass).
static int direct_waves(size_t lds) { return 2 * lds + 2048 <= 160 * 1024 ? 4 : 8; }

hipError_t prepare_direct() {
    const void* fns[] = {(const void*)k_direct<__bf16, 4>, (const void*)k_direct<__bf16, 8>,
                         (const void*)k_direct<_Float16, 4>, (const void*)k_direct<_Float16, 8>};
    for (const void* f : fns) {
        const hipError_t e =
            hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

size_t direct_jobs(const DirectScale* ds, const void* const* frag, float* const* C,
                   const int* level, int n, DirectJob* jobs, int* total_tiles) {
    // heaviest scales first, so that the short ones fill the tail of the launch
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (long)ds[a].ext * ds[a].ext > (long)ds[b].ext * ds[b].ext;
    });
    size_t lds = 0;
    int t = 0;
    for (int k = 0; k < n; ++k) {
        const int i = order[k];
        DirectScale d = ds[i];
        const size_t l = direct_geometry(d);
        if (l > kMaxLds || d.off != ds[i].off || d.nvb != ds[i].nvb || d.nrho != ds[i].nrho ||
            d.LH != ds[i].LH || d.lwp != ds[i].lwp || d.vsp != ds[i].vsp ||
            d.cstride != ds[i].cstride || ds[i].kW != 2 * ds[i].kpad + 1 || ds[i].off < 0 ||
            d.lwp > kDirMaxLwp)
            return 0;
        lds = std::max(lds, l);
        jobs[k].ds = ds[i];
        jobs[k].frag = frag[i];
        jobs[k].C = C[i];
        jobs[k].level = level[i];
        if (level[i] < 0 || level[i] >= kDirMaxLevels) return 0;
        jobs[k].tiles_x = (ds[i].Wc + kDirTX - 1) / kDirTX;
        jobs[k].tile0 = t;
        t += jobs[k].tiles_x * ((ds[i].Hc + kDirTY - 1) / kDirTY);
    }
    *total_tiles = t;
    return lds;
}

hipError_t launch_direct(const DirectLevels& lv, const DirectJob* d_jobs, int njobs,
                         int total_tiles, size_t lds, bool fp16, hipStream_t s) {
    if (lds > kMaxLds || njobs < 1 || total_tiles < 1) return hipErrorInvalidValue;
    const bool w4 = direct_waves(lds) == 4;
    if (fp16 && w4)
        k_direct<_Float16, 4><<<total_tiles, 4 * 64, lds, s>>>(lv, d_jobs, njobs);
    else if (fp16)
        k_direct<_Float16, 8><<<total_tiles, 8 * 64, lds, s>>>(lv, d_jobs, njobs);
    else if (w4)
        k_direct<__bf16, 4><<<total_tiles, 4 * 64, lds, s>>>(lv, d_jobs, njobs);
    else
        k_direct<__bf16, 8><<<total_tiles, 8 * 64, lds, s>>>(lv, d_jobs, njobs);
    return hipGetLastError();
}

}  // namespace bimpk
