// bank.hpp -- host-side (FP64) Gabor filter-bank spectra.
//
// The reference generates the even/odd kernels in host FP32 and transforms them with cuFFT
// R2C (bimp/src/cuda_bimp.cu:304-349).  Here the kernels follow the same formula (same
// float/double mix, so the FP32 taps are the reference's), and their spectra are evaluated in
// FP64 before a single rounding to FP32.
#pragma once
#include <complex>
#include <cstdint>
#include <vector>

namespace bimph {

struct ScaleGeom {
    float lambda2, lambda, pyrstep, fs;
    int kW, kX, off, level, H, W, FH, FW;
};

// snapTransformSize (cuda_bimp.cu:200-219)
int snap_transform_size(int n);
// CudaStuff parameter derivation (cuda_bimp.cu:231-256) with run-time constants
bool derive_scale(float lambda2, int width, int height, float pyr_threshold, float support,
                  ScaleGeom* g);
// Even/odd taps of orientation o (cuda_bimp.cu:304-331), kW*kW each, row-major.
// dc_out, when given, receives the subtracted DC offset evensum / fs^2.
void gabor_taps(const ScaleGeom& g, int o, float gamma, float sigma_ratio, float* ke, float* ko,
                float* dc_out = nullptr);
// Derivatives of the taps of orientation o with respect to sigma_ratio (sigma = ratio *
// lambda) and gamma, evaluated in FP64 on the same grid: d(Ke)/d., d(Ko)/d. (the DC term of Ke
// differentiated too).  Arrays of kW*kW, row-major.
void gabor_taps_grad(const ScaleGeom& g, int o, float gamma, float sigma_ratio, double* dke_dr,
                     double* dko_dr, double* dke_dg, double* dko_dg);
// Quadrature spectra G_o(ky, kx) = Re[DFT2(Ke + i*Ko)] / (FH*FW) for o = 0..4, stored
// column-major: out[(o*FW + kx)*FH + ky].  DFT(Ke) is real and DFT(Ko) imaginary (Ke even,
// Ko odd), so the real part carries the whole quadrature pair.
void quadrature_bank(const ScaleGeom& g, float gamma, float sigma_ratio, int threads,
                     std::vector<float>& out);

// Precision form: the 16 filters of one scale (even and odd kernel of each orientation,
// o = 0..7, filter index 2*o and 2*o+1) flipped for convolution, scaled by 1/fscale and
// rounded once to BF16 or FP16.  nsig <= 0 (or an extent that reaches kW) keeps the whole
// kW x kW kernel Ke; nsig > 0 truncates to the odd extent ceil(nsig*sigma) and the even filter
// becomes the carrier Ke + dc inside the window.  The caller subtracts dc[f] times the
// kW x kW box sum of the image from filter f's response,
//   dc[f] = (fscale * sum(rounded taps of f) - sum(exact kW x kW kernel of f)) / kW^2,
// so that a constant image gets the exact full kernel's response.
//
// Layout: the B operand of v_mfma_f32_16x16x32 for every tap block of 4 rows x 8 columns,
//   frag[((i*nvb + vb)*64 + lane)*8 + t] = tap (u0 + lane/16, 8*vb + t) of filter lane%16,
// with u0 = i - (rows - 1) for i < 4*nrho + rows - 1, nvb = ceil(ext/8),
// nrho = ceil((rows - 1 + ext)/4), and zero outside 0 <= u, v < ext.  `rows` is the number of
// output rows one wavefront accumulates (it sets the range of u0).
struct DirectBank {
    int ext = 0, nvb = 0, nrho = 0;
    float fscale = 1.f;
    float dc[16] = {0};
    std::vector<uint16_t> frag;
};
void direct_bank(const ScaleGeom& g, float gamma, float sigma_ratio, float nsig, bool fp16,
                 int rows, DirectBank& out);
// Extent direct_bank would choose (without building the taps).
int direct_extent(const ScaleGeom& g, float sigma_ratio, float nsig);
uint16_t float_to_bf16(float f);  // round to nearest even
uint16_t float_to_f16(float f);   // round to nearest even
double bf16_to_double(uint16_t h);
double f16_to_double(uint16_t h);

}  // namespace bimph
