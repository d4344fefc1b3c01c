// diff.cpp -- host side of the differentiation form (C ABI bimp_diff_*, include/bimp_hip.h).
//
// A diff context holds, per scale, the bank spectra Qhat_o of the current Gabor parameters
// (sigma ratio, gamma), their real quadrature form G_o and the tap derivatives dK/dtheta; per
// frame it recomputes the complex responses instead of storing them (forward and backward
// each run the filter).  The forward takes the detection path's transforms (two image rows
// per complex row FFT, half spectra, inverse columns with the Hermitian partner, G_o for
// o = 0..4 with 5..7 mirrored); the backward needs Z_o, FFT(Pd) and C_o on the whole periodic
// domain and recomputes them with full complex 2D transforms.  Images are FP32, scaled by
// 255 * input_scale before filtering, so an image u8/255 gives the maps of the uint8 path up
// to the pyramid's rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "../../include/bimp_hip.h"
#include "bank.hpp"
#include "kernels.hpp"

namespace bimpk {
hipError_t diff_prepare(int n);
hipError_t diff_rows(float2* data, int rows, int n, bool inv, const float2* tw, hipStream_t s);
hipError_t diff_rows_maps(float2* data, int maps, int per_map, size_t map_stride, int n,
                          bool inv, const float2* tw, hipStream_t s);
hipError_t diff_gather_tap_rows(const float2* Z, int kW, int kX, int FH, int FW, float2* Yc,
                                int nb, hipStream_t s);
hipError_t diff_cols(float2* data, int n, int w, int batch, bool inv, const float2* tw,
                     hipStream_t s);
hipError_t diff_pyr_down(const float* src, int sw, int sh, float* dst, int dw, int dh, int nb,
                         hipStream_t s);
hipError_t diff_pyr_down_adj(const float* gdst, int dw, int dh, float* gsrc, int sw, int sh,
                             int nb, hipStream_t s);
hipError_t diff_pad(const float* L, int pitch, size_t lstride, int H, int W, int FH, int FW,
                    int k, float sc, float2* out, int nb, hipStream_t s);
hipError_t diff_pad_adj(const float2* g, int H, int W, int FH, int FW, int k, float sc,
                        float* gL, int pitch, size_t lstride, int nb, hipStream_t s);
hipError_t diff_q_to_g(const float2* Q, int FH, int FW, float invN, float* G, hipStream_t s);
hipError_t diff_bank_pad(const float* ke, const float* ko, int kW, int kX, int FH, int FW,
                         float2* out, hipStream_t s);
hipError_t diff_spec_mul(const float2* D, const float2* Q, float2* Z, size_t n, int nb,
                         float invN, hipStream_t s);
hipError_t diff_mag(const float2* Z, float* C, size_t n, hipStream_t s);
hipError_t diff_cells_adj(const float* C, int FH, int FW, int H, int W, const CellParams& p,
                          const float* gD, size_t gstride, float* Wm, float* gC, int nb,
                          hipStream_t s);
hipError_t diff_grad_z(const float* gC, float2* Z, size_t n, hipStream_t s);
hipError_t diff_accum_conjq(const float2* Q, const float2* G, float2* out, size_t n, int nb,
                            float invN, hipStream_t s);
hipError_t diff_mul_conjd(float2* G, const float2* D, size_t n, int nb, float invN,
                          hipStream_t s);
int diff_param_blocks(int kW);
hipError_t diff_param_partials(const float2* Yc, int kW, int kX, int FW, const float* dK,
                               float* partials, size_t pstride, int nb, hipStream_t s);
hipError_t diff_fill(float* p, size_t n, hipStream_t s);
}  // namespace bimpk

namespace {

using bimpk::kNumOri;

struct DScale {
    bimph::ScaleGeom g;
    bimpk::CellParams cp;
    float2* Q = nullptr;   // 8 x FH x FW bank spectra
    float* G = nullptr;    // 5 x FW x FH quadrature bank Re(Q_o) / (FH*FW), kx-major
    int Hc = 0, Wc = 0, ldc = 0;  // complex-cell region of the forward (as in detection)
    float2* DT = nullptr;  // forward: (FW/2+1) x FH half spectrum, column-major, per image
    float* Cf = nullptr;   // forward: 8 x Hc x ldc complex-cell maps, per image
    int nbf = 1;           // forward: images per launch of its group (DT, Cf copies)
    float* dK = nullptr;   // 4 x 8 x kW x kW tap derivatives (dKe/dr, dKo/dr, dKe/dg, dKo/dg)
    size_t respOff = 0;    // offset of this scale's D map in a response vector
};

int round_half_away(double v) { return (int)std::round(v); }

}  // namespace

struct bimp_diff {
    int width = 0, height = 0;
    bimp_options opt{};
    float ratio = 0.f, gamma = 0.f;
    std::vector<DScale> sc;
    size_t respSize = 0;
    int maxLevel = 0;
    std::vector<int> lw, lh;
    std::vector<float*> lev, glev;
    std::map<int, float2*> tw;
    size_t Pmax = 0;
    float2 *Dhat = nullptr, *Z = nullptr, *acc = nullptr;
    float2* Yc = nullptr;  // backward: 8 x kW tap rows of Y_o
    size_t ycSize = 0;
    // forward: scales of one level and transform size share each launch (first, count)
    std::vector<std::pair<int, int>> groups;
    float2 *R = nullptr, *T = nullptr;  // forward: row spectra and inverse-column output
    float* Sf = nullptr;                // forward: single-stopped maps (scratch)
    size_t rSize = 0, tSize = 0;
    float *C = nullptr, *Wm = nullptr, *gC = nullptr, *S = nullptr, *taps = nullptr;
    unsigned* maxslot = nullptr;
    float* partials = nullptr;
    int maxBlocks = 0;
    int nbMax = 1;  // images per backward launch (the per-image buffers hold that many)
    std::string err;
};

namespace {

int dfail(bimp_diff* d, int code, const std::string& m) {
    if (d) d->err = m;
    return code;
}

#define DCHK(d, expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return dfail(d, BIMP_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// In-place 2D complex transform of `nb` FH x FW maps (scratch of the same size).
int fft2(bimp_diff* d, float2* a, int nb, int FH, int FW, bool inv, hipStream_t s) {
    DCHK(d, bimpk::diff_rows(a, nb * FH, FW, inv, d->tw[FW], s));
    DCHK(d, bimpk::diff_cols(a, FH, FW, nb, inv, d->tw[FH], s));
    return BIMP_OK;
}

int set_params(bimp_diff* d, float ratio, float gamma, hipStream_t s) {
    for (auto& x : d->sc) {
        const auto& g = x.g;
        const int kk = g.kW * g.kW;
        std::vector<float> ke((size_t)kNumOri * kk), ko((size_t)kNumOri * kk);
        std::vector<float> dk((size_t)4 * kNumOri * kk);
        std::vector<double> a(kk), b(kk), c(kk), e(kk);
        for (int o = 0; o < kNumOri; ++o) {
            bimph::gabor_taps(g, o, gamma, ratio, ke.data() + (size_t)o * kk, ko.data() + (size_t)o * kk);
            bimph::gabor_taps_grad(g, o, gamma, ratio, a.data(), b.data(), c.data(), e.data());
            const size_t n = (size_t)kNumOri * kk;
            for (int m = 0; m < kk; ++m) {
                dk[0 * n + (size_t)o * kk + m] = (float)a[m];
                dk[1 * n + (size_t)o * kk + m] = (float)b[m];
                dk[2 * n + (size_t)o * kk + m] = (float)c[m];
                dk[3 * n + (size_t)o * kk + m] = (float)e[m];
            }
        }
        DCHK(d, hipMemcpy(d->taps, ke.data(), ke.size() * 4, hipMemcpyHostToDevice));
        DCHK(d, hipMemcpy(d->taps + ke.size(), ko.data(), ko.size() * 4, hipMemcpyHostToDevice));
        DCHK(d, hipMemcpy(x.dK, dk.data(), dk.size() * 4, hipMemcpyHostToDevice));
        DCHK(d, bimpk::diff_bank_pad(d->taps, d->taps + ke.size(), g.kW, g.kX, g.FH, g.FW, x.Q, s));
        const int st = fft2(d, x.Q, kNumOri, g.FH, g.FW, false, s);
        if (st) return st;
        DCHK(d, bimpk::diff_q_to_g(x.Q, g.FH, g.FW, 1.0f / ((float)g.FH * (float)g.FW), x.G, s));
        DCHK(d, hipStreamSynchronize(s));
    }
    d->ratio = ratio;
    d->gamma = gamma;
    return BIMP_OK;
}

// Float pyramid of one image into d->lev.
// Float pyramids of nb consecutive images into d->lev (level l of image b at
// lev[l] + b * lw[l] * lh[l]).
int pyramid(bimp_diff* d, const float* img, int nb, hipStream_t s) {
    DCHK(d, hipMemcpyAsync(d->lev[0], img, (size_t)nb * d->width * d->height * 4,
                           hipMemcpyDeviceToDevice, s));
    for (int l = 1; l <= d->maxLevel; ++l)
        DCHK(d, bimpk::diff_pyr_down(d->lev[l - 1], d->lw[l - 1], d->lh[l - 1], d->lev[l], d->lw[l],
                                     d->lh[l], nb, s));
    return BIMP_OK;
}

// Forward of one scale for the backward pass, for the nb images whose pyramids are in
// d->lev: Dhat, Z, C per image on the whole FH x FW domain (Z and C valid in rows [0, Hc)),
// and the D maps into scratch.
int scale_forward(bimp_diff* d, DScale& x, int nb, hipStream_t s) {
    const auto& g = x.g;
    const size_t P = (size_t)g.FH * g.FW;
    const float sc = 255.0f * d->opt.input_scale;
    DCHK(d, bimpk::diff_pad(d->lev[g.level], d->lw[g.level],
                            (size_t)d->lw[g.level] * d->lh[g.level], g.H, g.W, g.FH, g.FW, g.kX,
                            sc, d->Dhat, nb, s));
    int st = fft2(d, d->Dhat, nb, g.FH, g.FW, false, s);
    if (st) return st;
    DCHK(d, bimpk::diff_spec_mul(d->Dhat, x.Q, d->Z, P, nb, 1.0f / (float)P, s));
    // inverse: columns first, then only the rows [0, Hc) the cells stencil and its adjoint
    // read (the rest of Z and C is left column-transformed, finite, and never read)
    DCHK(d, bimpk::diff_cols(d->Z, g.FH, g.FW, kNumOri * nb, true, d->tw[g.FH], s));
    DCHK(d, bimpk::diff_rows_maps(d->Z, kNumOri * nb, x.Hc, P, g.FW, true, d->tw[g.FW], s));
    DCHK(d, bimpk::diff_mag(d->Z, d->C, (size_t)kNumOri * P * nb, s));
    bimpk::CellScale cs[bimpk::kCellMaxScales];
    for (int b = 0; b < nb; ++b)
        cs[b] = bimpk::CellScale{d->C + (size_t)b * kNumOri * P, g.FH, g.FW, g.FH, g.FW, g.H,
                                 g.W, x.cp, d->S + (size_t)b * 2 * P + P,
                                 d->S + (size_t)b * 2 * P, d->maxslot};
    DCHK(d, bimpk::launch_cells(cs, nb, s));
    return BIMP_OK;
}

// The D maps of scales first..first+n-1 (one pyramid level, one transform size) into the
// response vector `resp` through the detection path's kernels, one launch per stage: real
// row FFTs of the float level (two rows per complex FFT), column FFTs of the half spectra,
// inverse columns with the quadrature bank G, inverse rows with the complex-cell magnitude
// over the Hc x Wc region, then the cells stencil.
int group_forward(bimp_diff* d, int first, int n, int img0, int nb, float* resp, hipStream_t s) {
    // entry k = b * n + j of the launches: image img0 + b of the pyramids in d->lev, scale
    // first + j; its D map goes to resp + b * respSize
    const DScale& x0 = d->sc[first];
    const auto& g = x0.g;
    const size_t dtSize = (size_t)(g.FW / 2 + 1) * g.FH;
    bimpk::ScaleSet ss;
    ss.n = n * nb;
    bimpk::CellScale cs[bimpk::kCellMaxScales];
    for (int b = 0; b < nb; ++b)
        for (int j = 0; j < n; ++j) {
            const DScale& x = d->sc[first + j];
            const int k = b * n + j;
            ss.kX[k] = x.g.kX;
            ss.G[k] = x.G;
            ss.DT[k] = x.DT + (size_t)b * dtSize;
            ss.C[k] = x.Cf + (size_t)b * kNumOri * x.Hc * x.ldc;
            cs[k] = bimpk::CellScale{ss.C[k], x.Hc, x.ldc, g.FH, g.FW, g.H, g.W, x.cp,
                                     resp + (size_t)b * d->respSize + x.respOff,
                                     d->Sf + (size_t)b * d->respSize + x.respOff, d->maxslot};
        }
    const int ncols = g.FW / 2 + 1;
    const size_t rstride = (size_t)g.H * ncols;
    const size_t tstride = (size_t)kNumOri * x0.Hc * g.FW;
    const size_t lsz = (size_t)d->lw[g.level] * d->lh[g.level];
    for (int b = 0; b < nb; ++b) {
        // the row pass reads one image: its n entries, written to their slots of R
        bimpk::ScaleSet sb;
        sb.n = n;
        for (int j = 0; j < n; ++j) {
            sb.kX[j] = ss.kX[b * n + j];
            sb.G[j] = ss.G[b * n + j];
            sb.DT[j] = ss.DT[b * n + j];
            sb.C[j] = ss.C[b * n + j];
        }
        DCHK(d, bimpk::launch_fwd_rows_f32(g.FW, d->lev[g.level] + (img0 + b) * lsz,
                                           (size_t)d->lw[g.level],
                                           g.H, g.W, sb, 255.0f * d->opt.input_scale,
                                           d->R + (size_t)b * n * rstride, ncols, rstride,
                                           d->tw[g.FW], s));
    }
    DCHK(d, bimpk::launch_fwd_cols(g.FH, d->R, ncols, rstride, g.H, ss, ncols, d->tw[g.FH], s));
    DCHK(d, bimpk::launch_inv_cols(g.FH, ss, g.FW, x0.Hc, d->T, tstride, d->tw[g.FH], s));
    DCHK(d, bimpk::launch_inv_rows(g.FW, g.FH, ss, d->T, tstride, x0.Hc, x0.Wc, x0.ldc,
                                   d->tw[g.FW], s));
    DCHK(d, bimpk::launch_cells(cs, n * nb, s));
    return BIMP_OK;
}

}  // namespace

extern "C" {

int bimp_diff_create(int width, int height, const float* lambdas, int n, const bimp_options* opt,
                     bimp_diff** out) {
    if (!out) return BIMP_E_ARG;
    *out = nullptr;
    if (width < 8 || height < 8 || !lambdas || n <= 0 || n > 64) return BIMP_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BIMP_E_DEVICE;
    bimp_diff* d = new bimp_diff;
    d->width = width;
    d->height = height;
    if (opt)
        d->opt = *opt;
    else
        bimp_default_options(&d->opt);
    auto bail = [&](int code) {
        bimp_diff_destroy(d);
        return code;
    };
    std::vector<float> lam(lambdas, lambdas + n);
    std::sort(lam.begin(), lam.end());
    std::vector<int> sizes;
    for (float l : lam) {
        DScale x;
        if (!bimph::derive_scale(l, width, height, d->opt.pyr_threshold, d->opt.support, &x.g) ||
            x.g.lambda < 1.f || x.g.H < 4 || x.g.W < 4)
            return bail(BIMP_E_ARG);
        if (!bimpk::fft_size_supported(x.g.FH) || !bimpk::fft_size_supported(x.g.FW) ||
            x.g.kW > x.g.FH || x.g.kW > x.g.FW)
            return bail(BIMP_E_UNSUPPORTED);
        for (int o = 0; o < kNumOri; ++o) {
            const double theta = (double)(float)(o * 3.1415926535897932384626433832795 / kNumOri);
            x.cp.ea[o] = round_half_away(0.6 * x.g.lambda * std::cos(theta));
            x.cp.eb[o] = round_half_away(0.6 * x.g.lambda * std::sin(theta));
            x.cp.ia[o] = round_half_away(0.6 * x.g.lambda * std::sin(theta));
            x.cp.ib[o] = round_half_away(0.6 * x.g.lambda * std::cos(theta));
            x.cp.ra[o] = round_half_away(0.5 * x.cp.ia[o]);
            x.cp.rb[o] = round_half_away(0.5 * x.cp.ib[o]);
        }
        x.respOff = d->respSize;
        d->respSize += (size_t)x.g.H * x.g.W;
        // the complex-cell region the cells stencil reads (the detection context's rule)
        x.Wc = std::min(x.g.FW, x.g.W + 1);
        x.Hc = std::min(x.g.FH, x.g.H + (x.g.W == x.g.FW ? 2 : 1));
        x.ldc = (x.Wc + 31) / 32 * 32;

        d->maxLevel = std::max(d->maxLevel, x.g.level);
        d->Pmax = std::max(d->Pmax, (size_t)x.g.FH * x.g.FW);
        sizes.push_back(x.g.FH);
        sizes.push_back(x.g.FW);
        d->sc.push_back(x);
    }
    // launch groups: consecutive scales (sorted by wavelength) of one level and transform size
    for (int i = 0; i < (int)d->sc.size();) {
        const auto& g = d->sc[i].g;
        int n = 1;
        while (i + n < (int)d->sc.size() && n < bimpk::kMaxBatch && n < bimpk::kCellMaxScales &&
               d->sc[i + n].g.level == g.level && d->sc[i + n].g.FH == g.FH &&
               d->sc[i + n].g.FW == g.FW)
            ++n;
        d->groups.emplace_back(i, n);
        // the forward takes nbf images of the group per launch: n * nbf entries, at most 8
        const int nbf = std::max(1, bimpk::kMaxBatch / n);
        for (int j = 0; j < n; ++j) d->sc[i + j].nbf = nbf;
        d->rSize = std::max(d->rSize, (size_t)n * nbf * g.H * (g.FW / 2 + 1));
        d->tSize = std::max(d->tSize, (size_t)n * nbf * kNumOri * d->sc[i].Hc * g.FW);
        i += n;
    }
    d->lw.assign(d->maxLevel + 1, 0);
    d->lh.assign(d->maxLevel + 1, 0);
    d->lw[0] = width;
    d->lh[0] = height;
    for (int l = 1; l <= d->maxLevel; ++l) {
        d->lw[l] = (d->lw[l - 1] + 1) / 2;
        d->lh[l] = (d->lh[l - 1] + 1) / 2;
    }
    // images per backward launch: up to the cells launch's 8 jobs, while the per-image
    // buffers (8 complex and 44 real maps of the largest transform) stay within 4 GB
    {
        size_t pm = 0;
        for (const auto& x : d->sc) pm = std::max(pm, (size_t)x.g.FH * x.g.FW);
        const size_t per = pm * (2 * sizeof(float2) + kNumOri * sizeof(float2) +
                                 (2 + 4) * kNumOri * sizeof(float) + 2 * sizeof(float));
        d->nbMax = (int)std::max<size_t>(1, std::min<size_t>(bimpk::kCellMaxScales,
                                                             (size_t(4) << 30) / per));
    }
    const size_t NB = (size_t)d->nbMax;
    d->lev.assign(d->maxLevel + 1, nullptr);
    d->glev.assign(d->maxLevel + 1, nullptr);
    for (int l = 0; l <= d->maxLevel; ++l)
        if (hipMalloc(&d->lev[l], NB * d->lw[l] * d->lh[l] * 4) != hipSuccess ||
            hipMalloc(&d->glev[l], NB * d->lw[l] * d->lh[l] * 4) != hipSuccess)
            return bail(BIMP_E_ALLOC);
    const size_t P = d->Pmax;
    size_t maxTaps = 0;
    for (auto& x : d->sc) {
        const size_t kk = (size_t)x.g.kW * x.g.kW;
        maxTaps = std::max(maxTaps, kk);
        d->ycSize = std::max(d->ycSize, (size_t)kNumOri * x.g.kW * x.g.FW);
        d->maxBlocks = std::max(d->maxBlocks, bimpk::diff_param_blocks(x.g.kW));
        if (hipMalloc(&x.Q, (size_t)kNumOri * x.g.FH * x.g.FW * sizeof(float2)) != hipSuccess ||
            hipMalloc(&x.G, (size_t)bimpk::kUniqueOri * x.g.FH * x.g.FW * sizeof(float)) != hipSuccess ||
            hipMalloc(&x.DT, (size_t)x.nbf * (x.g.FW / 2 + 1) * x.g.FH * sizeof(float2)) != hipSuccess ||
            hipMalloc(&x.Cf, (size_t)x.nbf * kNumOri * x.Hc * x.ldc * sizeof(float)) != hipSuccess ||
            hipMalloc(&x.dK, 4 * kNumOri * kk * sizeof(float)) != hipSuccess)
            return bail(BIMP_E_ALLOC);
    }
    if (hipMalloc(&d->Dhat, NB * P * sizeof(float2)) != hipSuccess ||
        hipMalloc(&d->acc, NB * P * sizeof(float2)) != hipSuccess ||
        hipMalloc(&d->Z, NB * kNumOri * P * sizeof(float2)) != hipSuccess ||
        hipMalloc(&d->C, NB * kNumOri * P * sizeof(float)) != hipSuccess ||
        hipMalloc(&d->gC, NB * kNumOri * P * sizeof(float)) != hipSuccess ||
        hipMalloc(&d->Wm, NB * 4 * kNumOri * P * sizeof(float)) != hipSuccess ||
        hipMalloc(&d->S, NB * 2 * P * sizeof(float)) != hipSuccess ||
        hipMalloc(&d->taps, 2 * kNumOri * maxTaps * sizeof(float)) != hipSuccess ||
        hipMalloc(&d->maxslot, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&d->R, d->rSize * sizeof(float2)) != hipSuccess ||
        hipMalloc(&d->T, d->tSize * sizeof(float2)) != hipSuccess ||
        hipMalloc(&d->Sf, (size_t)bimpk::kMaxBatch * d->respSize * sizeof(float)) != hipSuccess ||
        hipMalloc(&d->Yc, NB * d->ycSize * sizeof(float2)) != hipSuccess ||
        hipMalloc(&d->partials, NB * 2 * (size_t)d->maxBlocks * d->sc.size() * sizeof(float)) !=
                hipSuccess)
        return bail(BIMP_E_ALLOC);
    std::sort(sizes.begin(), sizes.end());
    sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
    for (int n2 : sizes) {
        if (bimpk::diff_prepare(n2) != hipSuccess || bimpk::prepare_size(n2) != hipSuccess)
            return bail(BIMP_E_DEVICE);
        std::vector<float2> t;
        if (!bimpk::twiddle_table(n2, t)) return bail(BIMP_E_UNSUPPORTED);
        float2* p = nullptr;
        if (hipMalloc(&p, t.size() * sizeof(float2) + 256) != hipSuccess) return bail(BIMP_E_ALLOC);
        d->tw[n2] = p;
        if (hipMemcpy(p, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess)
            return bail(BIMP_E_DEVICE);
    }
    if (bimpk::prepare_common() != hipSuccess) return bail(BIMP_E_DEVICE);
    const int st = set_params(d, d->opt.sigma_ratio, d->opt.gamma, nullptr);
    if (st) return bail(st);
    *out = d;
    return BIMP_OK;
}

int bimp_diff_destroy(bimp_diff* d) {
    if (!d) return BIMP_E_ARG;
    (void)hipDeviceSynchronize();
    for (auto& x : d->sc) {
        (void)hipFree(x.Q);
        (void)hipFree(x.G);
        (void)hipFree(x.DT);
        (void)hipFree(x.Cf);
        (void)hipFree(x.dK);
    }
    for (auto* p : d->lev) (void)hipFree(p);
    for (auto* p : d->glev) (void)hipFree(p);
    for (auto& kv : d->tw) (void)hipFree(kv.second);
    for (void* p : {(void*)d->Dhat, (void*)d->acc, (void*)d->Z, (void*)d->C,
                    (void*)d->gC, (void*)d->Wm, (void*)d->S, (void*)d->taps, (void*)d->maxslot,
                    (void*)d->partials, (void*)d->R, (void*)d->T, (void*)d->Sf, (void*)d->Yc})
        (void)hipFree(p);
    delete d;
    return BIMP_OK;
}

long long bimp_diff_response_size(const bimp_diff* d) { return d ? (long long)d->respSize : BIMP_E_ARG; }

int bimp_diff_set_params(bimp_diff* d, float sigma_ratio, float gamma) {
    if (!d || !(sigma_ratio > 0.f) || !(gamma > 0.f)) return BIMP_E_ARG;
    if (sigma_ratio == d->ratio && gamma == d->gamma) return BIMP_OK;
    return set_params(d, sigma_ratio, gamma, nullptr);
}

int bimp_diff_forward(bimp_diff* d, const float* d_img, int nimg, float* d_resp,
                      bimp_stream_t stream) {
    if (!d || !d_img || !d_resp || nimg <= 0) return BIMP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    const size_t isz = (size_t)d->width * d->height;
    // chunks of images whose pyramids fit d->lev; each group takes nbf of them per launch
    for (int b0 = 0; b0 < nimg; b0 += d->nbMax) {
        const int nb = std::min(d->nbMax, nimg - b0);
        int st = pyramid(d, d_img + b0 * isz, nb, s);
        if (st) return st;
        for (const auto& gr : d->groups) {
            const int nbf = d->sc[gr.first].nbf;
            for (int b = 0; b < nb; b += nbf)
                if ((st = group_forward(d, gr.first, gr.second, b, std::min(nbf, nb - b),
                                        d_resp + (size_t)(b0 + b) * d->respSize, s)))
                    return st;
        }
    }
    return BIMP_OK;
}

int bimp_diff_backward(bimp_diff* d, const float* d_img, int nimg, const float* d_grad_resp,
                       float* d_grad_img, double* grad_params, bimp_stream_t stream) {
    if (!d || !d_img || !d_grad_resp || !d_grad_img || nimg <= 0) return BIMP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    const size_t isz = (size_t)d->width * d->height;
    const float sc = 255.0f * d->opt.input_scale;
    double gr = 0, gg = 0;
    std::vector<float> part;
    const size_t pstride = 2 * (size_t)d->maxBlocks * d->sc.size();  // partials per image
    // chunks of up to nbMax images: every launch below covers the chunk's images at once
    for (int b0 = 0; b0 < nimg; b0 += d->nbMax) {
        const int nb = std::min(d->nbMax, nimg - b0);
        int st = pyramid(d, d_img + b0 * isz, nb, s);
        if (st) return st;
        for (int l = 0; l <= d->maxLevel; ++l)
            DCHK(d, bimpk::diff_fill(d->glev[l], (size_t)nb * d->lw[l] * d->lh[l], s));
        for (size_t si = 0; si < d->sc.size(); ++si) {
            DScale& x = d->sc[si];
            const auto& g = x.g;
            const size_t P = (size_t)g.FH * g.FW;
            const float invN = 1.0f / (float)P;
            // recompute the responses (the D maps go to scratch)
            if ((st = scale_forward(d, x, nb, s))) return st;
            DCHK(d, bimpk::diff_cells_adj(d->C, g.FH, g.FW, g.H, g.W, x.cp,
                                          d_grad_resp + b0 * d->respSize + x.respOff, d->respSize,
                                          d->Wm, d->gC, nb, s));
            DCHK(d, bimpk::diff_grad_z(d->gC, d->Z, (size_t)kNumOri * P * nb, s));
            // gZ is zero outside rows [0, Hc): their row transforms are zero, not computed
            DCHK(d, bimpk::diff_rows_maps(d->Z, kNumOri * nb, x.Hc, P, g.FW, false, d->tw[g.FW], s));
            DCHK(d, bimpk::diff_cols(d->Z, g.FH, g.FW, kNumOri * nb, false, d->tw[g.FH], s));
            DCHK(d, bimpk::diff_accum_conjq(x.Q, d->Z, d->acc, P, nb, invN, s));
            if ((st = fft2(d, d->acc, nb, g.FH, g.FW, true, s))) return st;
            DCHK(d, bimpk::diff_pad_adj(d->acc, g.H, g.W, g.FH, g.FW, g.kX, sc, d->glev[g.level],
                                        d->lw[g.level], (size_t)d->lw[g.level] * d->lh[g.level],
                                        nb, s));
            DCHK(d, bimpk::diff_mul_conjd(d->Z, d->Dhat, P, nb, invN, s));
            // Y_o = IFFT2(...) is read at the kW x kW taps only: columns, then the kW tap rows
            DCHK(d, bimpk::diff_cols(d->Z, g.FH, g.FW, kNumOri * nb, true, d->tw[g.FH], s));
            DCHK(d, bimpk::diff_gather_tap_rows(d->Z, g.kW, g.kX, g.FH, g.FW, d->Yc, nb, s));
            DCHK(d, bimpk::diff_rows(d->Yc, kNumOri * g.kW * nb, g.FW, true, d->tw[g.FW], s));
            DCHK(d, bimpk::diff_param_partials(d->Yc, g.kW, g.kX, g.FW, x.dK,
                                               d->partials + 2 * (size_t)d->maxBlocks * si,
                                               pstride, nb, s));
        }
        for (int l = d->maxLevel; l >= 1; --l)
            DCHK(d, bimpk::diff_pyr_down_adj(d->glev[l], d->lw[l], d->lh[l], d->glev[l - 1],
                                             d->lw[l - 1], d->lh[l - 1], nb, s));
        DCHK(d, hipMemcpyAsync(d_grad_img + b0 * isz, d->glev[0], (size_t)nb * isz * 4,
                               hipMemcpyDeviceToDevice, s));
        part.resize(pstride * nb);
        DCHK(d, hipMemcpyAsync(part.data(), d->partials, part.size() * 4, hipMemcpyDeviceToHost, s));
        DCHK(d, hipStreamSynchronize(s));
        // image by image, scale by scale, block by block: the order of one image at a time
        for (int b = 0; b < nb; ++b)
            for (size_t si = 0; si < d->sc.size(); ++si) {
                const int nbk = bimpk::diff_param_blocks(d->sc[si].g.kW);
                const float* pp = part.data() + pstride * b + 2 * (size_t)d->maxBlocks * si;
                for (int k = 0; k < nbk; ++k) {
                    gr += pp[2 * k];
                    gg += pp[2 * k + 1];
                }
            }
    }
    if (grad_params) {
        grad_params[0] = gr;
        grad_params[1] = gg;
    }
    return BIMP_OK;
}

const char* bimp_diff_last_error(const bimp_diff* d) { return d ? d->err.c_str() : ""; }

}  // extern "C"
