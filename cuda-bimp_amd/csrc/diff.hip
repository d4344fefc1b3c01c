// diff.hip -- differentiation form: the filter / complex-cell / end-stop / inhibition chain
// on float images, and its hand-written adjoint.
//
// Forward (per scale, per image; FP32): level image L (float pyramid), padded image
//   Pd = s * pad(L)                                 (clamp padding, convolutionFFT2D.cuh:56-85)
//   Z_o = IFFT(FFT(Pd) * Qhat_o) / N,  Qhat_o = FFT(pad(Ke_o + i*Ko_o))   (cuda_bimp.cu:486-507)
//   C_o = |Z_o|,  D = end-stop - inhibition          (cuda_bimp.cuh:436-517)
// Adjoint, for an upstream gradient gD of D:
//   gC_o from the stencil masks (two deterministic passes, no atomics),
//   gZ_o = gC_o * Z_o / |Z_o|,  Ghat_o = FFT(gZ_o),
//   gPd  = Re IFFT(sum_o conj(Qhat_o) * Ghat_o) / N,
//   Y_o  = IFFT(Ghat_o * conj(FFT(Pd))) / N = dL/dKe_o + i dL/dKo_o at the tap positions,
//   then gL = s * pad^T(gPd) and the float-pyramid adjoint down to the image.
// All transforms are full complex 2D FFTs: batched row transforms, then column transforms
// staged through LDS.
#include <hip/hip_runtime.h>

#include "fft_device.hpp"
#include "kernels.hpp"

namespace bimpk {

namespace {

using bimpdev::FFTCfg;
using bimpdev::in_index;
using bimpdev::out_index;

constexpr int kRowThreads = 256;

// Batched in-place complex transforms of rows of length N: `rows` rows in all, the first
// `per_map` rows of each map of `map_stride` elements (per_map = rows: contiguous rows).
template <int N, bool INV>
__global__ void __launch_bounds__(kRowThreads) k_c2c_rows(float2* __restrict__ data, int rows,
                                                          int per_map, size_t map_stride,
                                                          const float2* __restrict__ tw) {
    using C = FFTCfg<N>;
    constexpr int GPB = kRowThreads / C::TPV;
    extern __shared__ float lds[];
    const int grp = threadIdx.x / C::TPV, lane = threadIdx.x % C::TPV;
    // every group transforms (a spare one row 0, unstored): a 128-lane group's stages take
    // the workgroup barrier
    const int row = blockIdx.x * GPB + grp;
    const bool act = row < rows;
    const int r = act ? row : 0;
    float2* p = data + (size_t)(r / per_map) * map_stride + (size_t)(r % per_map) * N;
    float2 v[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) v[e] = p[in_index<N>(e, lane)];
    bimpdev::fft<N, INV>(v, lds + grp * bimpdev::fft_lds_floats<N>(), lane, tw);
    if (act) {
#pragma unroll
        for (int e = 0; e < C::E; ++e) p[out_index<N>(e, lane)] = v[e];
    }
}

// In-place complex transforms of the columns of `nb` N x W maps ([b][r][c], r < N): a
// workgroup stages CB adjacent columns through LDS (each row contributes CB contiguous
// values), one group of lanes transforms each column, and the results leave the same way.
// It replaces transpose + row transforms + transpose and moves each value once each way.
template <int N>
struct CColCfg {
    static constexpr int TPV = FFTCfg<N>::TPV;
    static constexpr int CB = TPV == 64 ? (N <= 2048 ? 8 : 4) : (512 / TPV > 16 ? 16 : 512 / TPV);
    static constexpr int THREADS = CB * TPV;
    static constexpr int S = N + 2;  // column stride (complex): 2 (mod 16) spreads the staging
};

template <int N, bool INV>
__global__ void __launch_bounds__(CColCfg<N>::THREADS) k_c2c_cols(float2* __restrict__ data,
                                                                   int W,
                                                                   const float2* __restrict__ tw) {
    using C = FFTCfg<N>;
    using CC = CColCfg<N>;
    extern __shared__ float2 lds2[];
    float2* map = data + (size_t)blockIdx.y * N * W;
    const int c0 = blockIdx.x * CC::CB;
    for (int t = threadIdx.x; t < N * CC::CB; t += CC::THREADS) {
        const int b = t % CC::CB, r = t / CC::CB;
        lds2[b * CC::S + r] = c0 + b < W ? map[(size_t)r * W + c0 + b] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const int g = threadIdx.x / C::TPV, lane = threadIdx.x % C::TPV;
    float2* my = lds2 + (size_t)g * CC::S;
    float2 v[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) v[e] = my[in_index<N>(e, lane)];
    bimpdev::stage_sync<C::TPV>();
    bimpdev::fft<N, INV>(v, reinterpret_cast<float*>(my), lane, tw);
#pragma unroll
    for (int e = 0; e < C::E; ++e) my[out_index<N>(e, lane)] = v[e];
    __syncthreads();
    for (int t = threadIdx.x; t < N * CC::CB; t += CC::THREADS) {
        const int b = t % CC::CB, r = t / CC::CB;
        if (c0 + b < W) map[(size_t)r * W + c0 + b] = lds2[b * CC::S + r];
    }
}

__device__ __forceinline__ int reflect101(int t, int n) {
    if (t < 0) t = -t;
    if (t >= n) t = 2 * (n - 1) - t;
    return t;
}

// Float pyramid level (the 5x5 binomial / 256 of pyrDown with reflect-101 borders, no rounding).
// blockIdx.z: image of the batch (levels of one size back to back)
__global__ void k_pyr_down_f(const float* __restrict__ src, int sw, int sh,
                             float* __restrict__ dst, int dw, int dh) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= dw) return;
    src += (size_t)blockIdx.z * sw * sh;
    dst += (size_t)blockIdx.z * dw * dh;
    const float k[5] = {1.f, 4.f, 6.f, 4.f, 1.f};
    float acc = 0.f;
    for (int i = -2; i <= 2; ++i) {
        const float* row = src + (size_t)reflect101(2 * y + i, sh) * sw;
        float r = 0.f;
        for (int j = -2; j <= 2; ++j) r += k[j + 2] * row[reflect101(2 * x + j, sw)];
        acc += k[i + 2] * r;
    }
    dst[(size_t)y * dw + x] = acc * (1.f / 256.f);
}

// Weights of the pyramid's 1D stencil seen from source index p: every (d, w) with
// src index p receiving weight w/16 from destination d.  Preimages of p under reflect-101
// are p, -p (p > 0) and 2(n-1) - p (p < n-1).
__device__ __forceinline__ int pyr_taps(int p, int n, int dn, int* d, float* w) {
    const float k[5] = {1.f, 4.f, 6.f, 4.f, 1.f};
    int cnt = 0;
    int pre[3], np = 0;
    pre[np++] = p;
    if (p > 0) pre[np++] = -p;
    if (p < n - 1) pre[np++] = 2 * (n - 1) - p;
    for (int a = 0; a < np; ++a) {
        const int t = pre[a];
        for (int i = -2; i <= 2; ++i) {
            const int q = t - i;
            if (q & 1) continue;
            const int dd = q / 2;
            if (dd < 0 || dd >= dn) continue;
            d[cnt] = dd;
            w[cnt] = k[i + 2] * (1.f / 16.f);
            ++cnt;
        }
    }
    return cnt;
}

// gsrc += pyrDown^T(gdst)
__global__ void k_pyr_down_adj(const float* __restrict__ gdst, int dw, int dh,
                               float* __restrict__ gsrc, int sw, int sh) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= sw) return;
    gdst += (size_t)blockIdx.z * dw * dh;
    gsrc += (size_t)blockIdx.z * sw * sh;
    int dy[9], dx[9];
    float wy[9], wx[9];
    const int ny = pyr_taps(y, sh, dh, dy, wy);
    const int nx = pyr_taps(x, sw, dw, dx, wx);
    float acc = 0.f;
    for (int a = 0; a < ny; ++a) {
        float r = 0.f;
        for (int b = 0; b < nx; ++b) r += wx[b] * gdst[(size_t)dy[a] * dw + dx[b]];
        acc += wy[a] * r;
    }
    gsrc[(size_t)y * sw + x] += acc;
}

__device__ __forceinline__ int clamp_src(int y, int H, int k) { return y < H ? y : (y < H + k ? H - 1 : 0); }

// Pd = s * pad(L) as complex (imaginary part 0).  L is the top-left H x W region of a level
// image with `pitch` floats per row (a level is (n+1)/2 wide, its data n/2).
// blockIdx.z: image of the batch (level images lstride floats apart, outputs FH*FW apart)
__global__ void k_pad_fwd(const float* __restrict__ L, int pitch, size_t lstride, int H, int W,
                          int FH, int FW, int k, float s, float2* __restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= FW) return;
    L += blockIdx.z * lstride;
    out += (size_t)blockIdx.z * FH * FW;
    out[(size_t)y * FW + x] =
            make_float2(s * L[(size_t)clamp_src(y, H, k) * pitch + clamp_src(x, W, k)], 0.f);
}

// Source rows of the padded image that read data row r: r itself, the band [H, H+k) for
// r = H-1 and the band [H+k, F) for r = 0.
__device__ __forceinline__ void pad_band(int r, int H, int F, int k, int* lo1, int* hi1) {
    *lo1 = 0;
    *hi1 = 0;
    if (r == H - 1) {
        *lo1 = H;
        *hi1 = min(H + k, F);
    } else if (r == 0) {
        *lo1 = min(H + k, F);
        *hi1 = F;
    }
}

// gL += s * pad^T(Re gPd) over the top-left H x W region of gL (`pitch` floats per row)
__global__ void k_pad_adj(const float2* __restrict__ g, int H, int W, int FH, int FW, int k,
                          float s, float* __restrict__ gL, int pitch, size_t lstride) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    g += (size_t)blockIdx.z * FH * FW;
    gL += blockIdx.z * lstride;
    int ylo, yhi, xlo, xhi;
    pad_band(y, H, FH, k, &ylo, &yhi);
    pad_band(x, W, FW, k, &xlo, &xhi);
    float acc = 0.f;
    for (int yy = -1; yy < yhi - ylo; ++yy) {
        const int sy = yy < 0 ? y : ylo + yy;
        const float2* row = g + (size_t)sy * FW;
        acc += row[x].x;
        for (int xx = xlo; xx < xhi; ++xx) acc += row[xx].x;
    }
    gL[(size_t)y * pitch + x] += s * acc;
}

// G_o(ky, kx) = Re Q_o(ky, kx) / (FH*FW) for o = 0..4, stored kx-major (G[o][kx][ky]): the
// quadrature bank layout of the detection path's inverse column pass, from the spectra Q of
// the current parameters.  32 x 32 tiles transposed through LDS.
__global__ void __launch_bounds__(256) k_q_to_g(const float2* __restrict__ Q, int FH, int FW,
                                               float invN, float* __restrict__ G) {
    __shared__ float tile[32][33];
    const int o = blockIdx.z, kx0 = blockIdx.x * 32, ky0 = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    const float2* q = Q + (size_t)o * FH * FW;
    for (int r = ty; r < 32; r += 8) {
        const int ky = ky0 + r, kx = kx0 + tx;
        if (ky < FH && kx < FW) tile[r][tx] = q[(size_t)ky * FW + kx].x * invN;
    }
    __syncthreads();
    float* g = G + (size_t)o * FW * FH;
    for (int r = ty; r < 32; r += 8) {
        const int kx = kx0 + r, ky = ky0 + tx;
        if (ky < FH && kx < FW) g[(size_t)kx * FH + ky] = tile[tx][r];
    }
}

// Padded complex kernels of the 8 orientations: tap (r, c) sits at ((r-kX) mod FH,
// (c-kX) mod FW) (requires kW <= FH, FW, so positions are distinct).
__global__ void k_bank_pad(const float* __restrict__ ke, const float* __restrict__ ko, int kW,
                           int kX, int FH, int FW, float2* __restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, o = blockIdx.z;
    if (x >= FW) return;
    const int r = (y + kX) % FH, c = (x + kX) % FW;
    float2 v = make_float2(0.f, 0.f);
    if (r < kW && c < kW) {
        const size_t m = ((size_t)o * kW + r) * kW + c;
        v = make_float2(ke[m], ko[m]);
    }
    out[((size_t)o * FH + y) * FW + x] = v;
}

// Z_o = D * Q_o * invN for the 8 orientations; nb images (D: nb maps of n, Z: nb x 8 maps).
__global__ void k_spec_mul(const float2* __restrict__ D, const float2* __restrict__ Q,
                           float2* __restrict__ Z, size_t n, int nb, float invN) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nb) return;
    const size_t b = t / n, i = t - b * n;
    const float2 d = D[t];
    float2* z = Z + b * kNumOri * n;
#pragma unroll
    for (int o = 0; o < kNumOri; ++o) {
        const float2 q = Q[(size_t)o * n + i];
        z[(size_t)o * n + i] =
            make_float2((d.x * q.x - d.y * q.y) * invN, (d.x * q.y + d.y * q.x) * invN);
    }
}

__global__ void k_mag(const float2* __restrict__ Z, float* __restrict__ C, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 z = Z[i];
    C[i] = sqrtf(z.x * z.x + z.y * z.y);
}

// Complex-cell map read with the reference's linear index y*FW + x over the FH x FW array,
// wrapping past the end (the forward stencil's semantics).
__device__ __forceinline__ float cat(const float* C, int o, long idx, long P) {
    if (idx >= P) idx -= P;
    if (idx < 0) idx += P;
    return C[(size_t)o * P + idx];
}

__device__ __forceinline__ bool valid(int x, int y, int a, int b, int W, int H) {
    const int aa = a < 0 ? -a : a, bb = b < 0 ? -b : b;
    return 2 * bb <= x && x <= W - 2 * bb && 2 * aa <= y && y <= H - 2 * aa;
}

// Pass 1 of the stencil adjoint: per pixel and orientation, upstream gradient times the
// active-branch masks of the forward (cuda_bimp.cuh:436-502):
//   Wd = g [ds > 0], Wt = g [tangential > 0], W1 = g [v2 > 0][v1 > 0], W2 = g [v2 > 0]
// stored as 4 x 8 maps over the FH x FW grid (zero outside the data region).
// blockIdx.z: image of the batch (C: 8 maps, Wm: 32 maps per image; gD gstride apart)
__global__ void k_cells_adj_w(const float* __restrict__ C, int FH, int FW, int H, int W,
                              CellParams p, const float* __restrict__ gD, size_t gstride,
                              float* __restrict__ Wm) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= FW) return;
    const long P = (long)FH * FW, g0 = (long)y * FW + x;
    C += (size_t)blockIdx.z * kNumOri * P;
    gD += blockIdx.z * gstride;
    Wm += (size_t)blockIdx.z * 4 * kNumOri * P;
    const float g = (y < H && x < W) ? gD[(size_t)y * W + x] : 0.f;
#pragma unroll
    for (int o = 0; o < kNumOri; ++o) {
        float wd = 0.f, wt = 0.f, w1 = 0.f, w2 = 0.f;
        if (g != 0.f) {
            const int a = p.ea[o], b = p.eb[o];
            if (valid(x, y, a, b, W, H)) {
                const float v = (float)((double)cat(C, o, g0, P) -
                                        0.5 * (double)cat(C, o, g0 - 2L * a * FW + 2 * b, P) -
                                        0.5 * (double)cat(C, o, g0 + 2L * a * FW - 2 * b, P));
                if (v > 0.f) wd = g;
            }
            const int ia = p.ia[o], ib = p.ib[o];
            if (valid(x, y, ia, ib, W, H)) {
                const int o2 = (o + kNumOri / 2) % kNumOri;
                const float c = cat(C, o, g0, P);
                const float vt = cat(C, o, g0 - (long)ia * FW - ib, P) - c;
                if (vt > 0.f) wt = g;
                const long rr = (long)p.ra[o] * FW + p.rb[o];
                float v1 = c - 16.f * cat(C, o2, g0 + rr, P);
                const bool m1 = v1 > 0.f;
                v1 = m1 ? v1 : 0.f;
                const float v2 = v1 + c - 16.f * cat(C, o2, g0 - rr, P);
                if (v2 > 0.f) {
                    w2 = g;
                    w1 = m1 ? g : 0.f;
                }
            }
        }
        Wm[((size_t)0 * kNumOri + o) * P + g0] = wd;
        Wm[((size_t)1 * kNumOri + o) * P + g0] = wt;
        Wm[((size_t)2 * kNumOri + o) * P + g0] = w1;
        Wm[((size_t)3 * kNumOri + o) * P + g0] = w2;
    }
}

// Pass 2: gather, for each map position p and orientation o, every forward read of C_o(p):
//   D = sum_o relu(C(x) - C(x+d1)/2 - C(x+d2)/2) - sum_o [relu(C(x+t) - C(x)) + R_o(x)],
//   R_o(x) = relu(relu(C_o(x) - 16 C_o'(x+r)) + C_o(x) - 16 C_o'(x-r)), o' = o+4.
// blockIdx.z = image * 8 + orientation
__global__ void k_cells_adj_g(const float* __restrict__ Wm, int FH, int FW, CellParams p,
                              float* __restrict__ gC) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    const int o = blockIdx.z % kNumOri, img = blockIdx.z / kNumOri;
    if (x >= FW) return;
    const long P = (long)FH * FW, q = (long)y * FW + x;
    Wm += (size_t)img * 4 * kNumOri * P;
    gC += (size_t)img * kNumOri * P;
    const float* Wd = Wm + ((size_t)0 * kNumOri + o) * P;
    const float* Wt = Wm + ((size_t)1 * kNumOri + o) * P;
    const int a = p.ea[o], b = p.eb[o];
    const long d1 = -2L * a * FW + 2 * b;
    float acc = cat(Wd, 0, q, P) - 0.5f * cat(Wd, 0, q - d1, P) - 0.5f * cat(Wd, 0, q + d1, P);
    const long t = -(long)p.ia[o] * FW - p.ib[o];
    acc += cat(Wt, 0, q, P) - cat(Wt, 0, q - t, P);
    // own radial term (centre)
    const float* W1 = Wm + ((size_t)2 * kNumOri + o) * P;
    const float* W2 = Wm + ((size_t)3 * kNumOri + o) * P;
    acc -= cat(W1, 0, q, P) + cat(W2, 0, q, P);
    // partner terms: C_o is the second map of orientation o' = o + 4 (mod 8)
    const int op = (o + kNumOri / 2) % kNumOri;
    const float* W1p = Wm + ((size_t)2 * kNumOri + op) * P;
    const float* W2p = Wm + ((size_t)3 * kNumOri + op) * P;
    const long rr = (long)p.ra[op] * FW + p.rb[op];
    acc += 16.f * cat(W1p, 0, q - rr, P) + 16.f * cat(W2p, 0, q + rr, P);
    gC[(size_t)o * P + q] = acc;
}

// gZ = gC * Z / |Z| in place over Z (zero where |Z| = 0).
__global__ void k_grad_z(const float* __restrict__ gC, float2* __restrict__ Z, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 z = Z[i];
    const float m = sqrtf(z.x * z.x + z.y * z.y);
    const float s = m > 0.f ? gC[i] / m : 0.f;
    Z[i] = make_float2(z.x * s, z.y * s);
}

// out = invN * sum_o conj(Q_o) * G_o per image (G: nb x 8 maps of n, out: nb maps)
__global__ void k_accum_conjq(const float2* __restrict__ Q, const float2* __restrict__ G,
                              float2* __restrict__ out, size_t n, int nb, float invN) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nb) return;
    const size_t b = t / n, i = t - b * n;
    const float2* gb = G + b * kNumOri * n;
    float re = 0.f, im = 0.f;
#pragma unroll
    for (int o = 0; o < kNumOri; ++o) {
        const float2 q = Q[(size_t)o * n + i], g = gb[(size_t)o * n + i];
        re += q.x * g.x + q.y * g.y;
        im += q.x * g.y - q.y * g.x;
    }
    out[t] = make_float2(re * invN, im * invN);
}

// G_o = invN * G_o * conj(D) in place, per image (G: nb x 8 maps of n, D: nb maps).
__global__ void k_mul_conjd(float2* __restrict__ G, const float2* __restrict__ D, size_t n,
                            int nb, float invN) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nb) return;
    const size_t b = t / n, i = t - b * n;
    const float2 d = D[t];
    float2* gb = G + b * kNumOri * n;
#pragma unroll
    for (int o = 0; o < kNumOri; ++o) {
        const float2 g = gb[(size_t)o * n + i];
        gb[(size_t)o * n + i] =
            make_float2((g.x * d.x + g.y * d.y) * invN, (g.y * d.x - g.x * d.y) * invN);
    }
}

// Rows (r - kX) mod FH, r = 0..kW-1, of the 8 column-transformed maps Z (FH x FW each) into
// Yc (8 x kW rows of FW): the only rows of Y_o the parameter partials read.
__global__ void k_gather_tap_rows(const float2* __restrict__ Z, int kW, int kX, int FH, int FW,
                                  float2* __restrict__ Yc) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, o = blockIdx.z;
    if (x >= FW) return;
    const int y = (r - kX + FH) % FH;
    Yc[((size_t)o * kW + r) * FW + x] = Z[((size_t)o * FH + y) * FW + x];
}

// Parameter gradient partials: sum over orientations and taps of
//   Re Y_o(tap) dKe/dtheta + Im Y_o(tap) dKo/dtheta  for theta = (sigma ratio, gamma),
// with Y_o's tap rows in Yc (k_gather_tap_rows, then row transforms); one pair of partial
// sums per block (fixed-order tree), summed on the host in block order.
__global__ void __launch_bounds__(256) k_param_partials(const float2* __restrict__ Yc, int kW,
                                                        int kX, int FW,
                                                        const float* __restrict__ dK,
                                                        float* __restrict__ partials,
                                                        size_t pstride) {
    __shared__ float sr[256], sg[256];
    const int n = kNumOri * kW * kW;
    const int i = blockIdx.x * 256 + threadIdx.x;
    // blockIdx.y: image of the batch (its tap rows and its partials)
    Yc += (size_t)blockIdx.y * kNumOri * kW * FW;
    partials += (size_t)blockIdx.y * pstride;
    float vr = 0.f, vg = 0.f;
    if (i < n) {
        const int o = i / (kW * kW), m = i % (kW * kW), r = m / kW, c = m % kW;
        const int x = (c - kX + FW) % FW;
        const float2 v = Yc[((size_t)o * kW + r) * FW + x];
        const size_t stride = (size_t)n;
        vr = v.x * dK[0 * stride + i] + v.y * dK[1 * stride + i];
        vg = v.x * dK[2 * stride + i] + v.y * dK[3 * stride + i];
    }
    sr[threadIdx.x] = vr;
    sg[threadIdx.x] = vg;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            sr[threadIdx.x] += sr[threadIdx.x + h];
            sg[threadIdx.x] += sg[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sr[0];
        partials[2 * blockIdx.x + 1] = sg[0];
    }
}

__global__ void k_fill(float* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

inline dim3 ew(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

hipError_t diff_prepare(int n) {
#define X(N)                                                                                 \
    if (n == N) {                                                                            \
        const int lds = (kRowThreads / FFTCfg<N>::TPV) * bimpdev::fft_lds_floats<N>() * (int)sizeof(float); \
        const int clds = CColCfg<N>::CB * CColCfg<N>::S * (int)sizeof(float2);               \
        hipError_t e = hipFuncSetAttribute((const void*)k_c2c_rows<N, false>,                \
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds); \
        if (e == hipSuccess)                                                                 \
            e = hipFuncSetAttribute((const void*)k_c2c_rows<N, true>,                        \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds);        \
        if (e == hipSuccess)                                                                 \
            e = hipFuncSetAttribute((const void*)k_c2c_cols<N, false>,                       \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, clds);       \
        if (e == hipSuccess)                                                                 \
            e = hipFuncSetAttribute((const void*)k_c2c_cols<N, true>,                        \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, clds);       \
        return e;                                                                            \
    }
    BIMP_FOR_SIZES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t diff_rows_maps(float2* data, int maps, int per_map, size_t map_stride, int n,
                          bool inv, const float2* tw, hipStream_t s) {
    const int rows = maps * per_map;
    if (maps < 1 || per_map < 1) return hipErrorInvalidValue;
#define X(N)                                                                                 \
    if (n == N) {                                                                            \
        constexpr int GPB = kRowThreads / FFTCfg<N>::TPV;                                    \
        const int lds = GPB * bimpdev::fft_lds_floats<N>() * (int)sizeof(float);             \
        dim3 grid((rows + GPB - 1) / GPB);                                                   \
        if (inv)                                                                             \
            k_c2c_rows<N, true><<<grid, kRowThreads, lds, s>>>(data, rows, per_map, map_stride, \
                                                               tw);                          \
        else                                                                                 \
            k_c2c_rows<N, false><<<grid, kRowThreads, lds, s>>>(data, rows, per_map,         \
                                                                map_stride, tw);             \
        return hipGetLastError();                                                            \
    }
    BIMP_FOR_SIZES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t diff_rows(float2* data, int rows, int n, bool inv, const float2* tw, hipStream_t s) {
    return diff_rows_maps(data, 1, rows, 0, n, inv, tw, s);
}

hipError_t diff_cols(float2* data, int n, int w, int batch, bool inv, const float2* tw,
                     hipStream_t s) {
#define X(N)                                                                                 \
    if (n == N) {                                                                            \
        using CC = CColCfg<N>;                                                               \
        const int lds = CC::CB * CC::S * (int)sizeof(float2);                                \
        dim3 grid((w + CC::CB - 1) / CC::CB, batch);                                         \
        if (inv)                                                                             \
            k_c2c_cols<N, true><<<grid, CC::THREADS, lds, s>>>(data, w, tw);                 \
        else                                                                                 \
            k_c2c_cols<N, false><<<grid, CC::THREADS, lds, s>>>(data, w, tw);                \
        return hipGetLastError();                                                            \
    }
    BIMP_FOR_SIZES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t diff_pyr_down(const float* src, int sw, int sh, float* dst, int dw, int dh, int nb,
                         hipStream_t s) {
    k_pyr_down_f<<<dim3((dw + 127) / 128, dh, nb), 128, 0, s>>>(src, sw, sh, dst, dw, dh);
    return hipGetLastError();
}

hipError_t diff_pyr_down_adj(const float* gdst, int dw, int dh, float* gsrc, int sw, int sh,
                             int nb, hipStream_t s) {
    k_pyr_down_adj<<<dim3((sw + 127) / 128, sh, nb), 128, 0, s>>>(gdst, dw, dh, gsrc, sw, sh);
    return hipGetLastError();
}

hipError_t diff_pad(const float* L, int pitch, size_t lstride, int H, int W, int FH, int FW,
                    int k, float sc, float2* out, int nb, hipStream_t s) {
    k_pad_fwd<<<dim3((FW + 127) / 128, FH, nb), 128, 0, s>>>(L, pitch, lstride, H, W, FH, FW, k,
                                                             sc, out);
    return hipGetLastError();
}

hipError_t diff_pad_adj(const float2* g, int H, int W, int FH, int FW, int k, float sc,
                        float* gL, int pitch, size_t lstride, int nb, hipStream_t s) {
    k_pad_adj<<<dim3((W + 127) / 128, H, nb), 128, 0, s>>>(g, H, W, FH, FW, k, sc, gL, pitch,
                                                           lstride);
    return hipGetLastError();
}

hipError_t diff_q_to_g(const float2* Q, int FH, int FW, float invN, float* G, hipStream_t s) {
    k_q_to_g<<<dim3((FW + 31) / 32, (FH + 31) / 32, kUniqueOri), 256, 0, s>>>(Q, FH, FW, invN, G);
    return hipGetLastError();
}

hipError_t diff_bank_pad(const float* ke, const float* ko, int kW, int kX, int FH, int FW,
                         float2* out, hipStream_t s) {
    k_bank_pad<<<dim3((FW + 127) / 128, FH, kNumOri), 128, 0, s>>>(ke, ko, kW, kX, FH, FW, out);
    return hipGetLastError();
}

hipError_t diff_spec_mul(const float2* D, const float2* Q, float2* Z, size_t n, int nb,
                         float invN, hipStream_t s) {
    k_spec_mul<<<ew(n * nb), 256, 0, s>>>(D, Q, Z, n, nb, invN);
    return hipGetLastError();
}

hipError_t diff_mag(const float2* Z, float* C, size_t n, hipStream_t s) {
    k_mag<<<ew(n), 256, 0, s>>>(Z, C, n);
    return hipGetLastError();
}

hipError_t diff_cells_adj(const float* C, int FH, int FW, int H, int W, const CellParams& p,
                          const float* gD, size_t gstride, float* Wm, float* gC, int nb,
                          hipStream_t s) {
    k_cells_adj_w<<<dim3((FW + 127) / 128, FH, nb), 128, 0, s>>>(C, FH, FW, H, W, p, gD, gstride,
                                                                 Wm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_cells_adj_g<<<dim3((FW + 127) / 128, FH, kNumOri * nb), 128, 0, s>>>(Wm, FH, FW, p, gC);
    return hipGetLastError();
}

hipError_t diff_grad_z(const float* gC, float2* Z, size_t n, hipStream_t s) {
    k_grad_z<<<ew(n), 256, 0, s>>>(gC, Z, n);
    return hipGetLastError();
}

hipError_t diff_accum_conjq(const float2* Q, const float2* G, float2* out, size_t n, int nb,
                            float invN, hipStream_t s) {
    k_accum_conjq<<<ew(n * nb), 256, 0, s>>>(Q, G, out, n, nb, invN);
    return hipGetLastError();
}

hipError_t diff_mul_conjd(float2* G, const float2* D, size_t n, int nb, float invN,
                          hipStream_t s) {
    k_mul_conjd<<<ew(n * nb), 256, 0, s>>>(G, D, n, nb, invN);
    return hipGetLastError();
}

int diff_param_blocks(int kW) { return (kNumOri * kW * kW + 255) / 256; }

hipError_t diff_gather_tap_rows(const float2* Z, int kW, int kX, int FH, int FW, float2* Yc,
                                int nb, hipStream_t s) {
    // map index z = image * 8 + orientation in both Z and Yc
    k_gather_tap_rows<<<dim3((FW + 127) / 128, kW, kNumOri * nb), 128, 0, s>>>(Z, kW, kX, FH, FW,
                                                                               Yc);
    return hipGetLastError();
}

hipError_t diff_param_partials(const float2* Yc, int kW, int kX, int FW, const float* dK,
                               float* partials, size_t pstride, int nb, hipStream_t s) {
    k_param_partials<<<dim3(diff_param_blocks(kW), nb), 256, 0, s>>>(Yc, kW, kX, FW, dK, partials,
                                                                     pstride);
    return hipGetLastError();
}

hipError_t diff_fill(float* p, size_t n, hipStream_t s) {
    k_fill<<<ew(n), 256, 0, s>>>(p, n);
    return hipGetLastError();
}

}  // namespace bimpk
