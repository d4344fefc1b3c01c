// bank.hpp -- host-side (FP64) Gabor filter-bank spectra.
//
// The reference generates the even/odd kernels in host FP32 and transforms them with cuFFT
// R2C (bimp/src/cuda_bimp.cu:304-349).  Here the kernels follow the same formula (same
// float/double mix, so the FP32 taps are the reference's), and their spectra are evaluated in
// FP64 before a single rounding to FP32.
#pragma once
#include <complex>
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
void gabor_taps(const ScaleGeom& g, int o, float gamma, float sigma_ratio, float* ke, float* ko);
// Quadrature spectra G_o(ky, kx) = Re[DFT2(Ke + i*Ko)] / (FH*FW) for o = 0..4, stored
// column-major: out[(o*FW + kx)*FH + ky].  DFT(Ke) is real and DFT(Ko) imaginary (Ke even,
// Ko odd), so the real part carries the whole quadrature pair.
void quadrature_bank(const ScaleGeom& g, float gamma, float sigma_ratio, int threads,
                     std::vector<float>& out);

}  // namespace bimph
