/*
 * bimp_oracle.h -- CPU restatement of the reference BIMP keypoint path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke check and the
 * CPU-baseline leg of bench.py load it to compare against the HIP library.  The product
 * (cuda-bimp_amd/) never links, loads or calls it.
 *
 * Every function restates the reference semantics it cites (daniel-koudouna/cuda-bimp;
 * file:line with basenames: cuda_bimp.cu = bimp/src/cuda_bimp.cu, cuda_bimp.cuh =
 * bimp/include/detection/cuda_bimp.cuh, convolutionFFT2D.cuh =
 * bimp/include/detection/convolutionFFT2D.cuh).  The reference's FFT (cuFFT), pyramid
 * (OpenCV-CUDA pyrDown) and max reduction (thrust) are third-party and cannot run here;
 * the oracle restates their published mathematics instead:
 *   - FFT convolution: textbook DFT, own mixed-radix Stockham FFT (FP64 or FP32);
 *   - pyrDown: 5-tap binomial [1 4 6 4 1]^2/256, reflect-101 border, round-half-even
 *     (OpenCV-CUDA 3.3.1 semantics; parity unpinned: no OpenCV in this image);
 *   - max: exact float maximum.
 */
#ifndef BIMP_ORACLE_H
#define BIMP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cv::KeyPoint field order (pt.x, pt.y, size, angle, response, octave, class_id). */
typedef struct {
    float x, y, size, angle, response;
    int octave, class_id;
} bo_keypoint;

/* Per-scale parameters derived exactly as in CudaStuff::CudaStuff (cuda_bimp.cu:222-256). */
typedef struct {
    float lambda2;  /* wavelength requested at full resolution */
    float lambda;   /* wavelength at the pyramid level */
    float pyrstep;  /* 1, 2, 4, ... */
    float fs;       /* float filter size (7*lambda, +1 if its integer part is even) */
    int kW;         /* kernel width = height (odd) */
    int kX;         /* kernel centre = floor(fs/2) */
    int off;        /* NMS border = (int)lambda */
    int level;      /* log2(pyrstep) */
    int dataH, dataW;
    int fftH, fftW;
} bo_scale;

int bo_snap_transform_size(int n);
int bo_make_lambdas_log(float start, float end, int spo, float* out, int cap);
int bo_scale_params(float lambda2, int width, int height, bo_scale* s);
void bo_gabor(const bo_scale* s, int o, float* ke, float* ko);
int bo_pyr_down(const uint8_t* src, int w, int h, size_t pitch, uint8_t* dst);
/* OpenCV's host cv::pyrDown rounding, (S + 128) >> 8: the reference's resize_cpu option */
int bo_pyr_down_host(const uint8_t* src, int w, int h, size_t pitch, uint8_t* dst);
void bo_pad_clamp(const uint8_t* img, size_t pitch, const bo_scale* s, float* out);
int bo_end_stop_offsets(float lambda, int o, int* a, int* b);
int bo_inhibition_offsets(float lambda, int o, int* a, int* b);

/* In-place 1D complex FFT in FP64 on interleaved (re, im) pairs. sign -1 forward, +1 inverse
 * (unnormalised).  Returns 0, or -1 for an unsupported length. */
int bo_fft_c2c(double* data, int n, int sign);

/* Brute-force circular convolution over the FH x FW torus of the padded image with the
 * padded kernel (the mathematics the FFT path must equal; O(FH*FW*kW^2), tiny inputs only).
 * e and o are FP64 maps of FH*FW. */
void bo_brute_filter(const uint8_t* level_img, size_t pitch, const bo_scale* s, int o,
                     double* e, double* od);

/* ---- context: bank precomputed at creation, as CudaStuff does ---- */
typedef struct bo_ctx bo_ctx;
bo_ctx* bo_create(int width, int height, const float* lambdas, int n, int fp64);
/* The context's pyramid (bo_set_image) rounded as bo_pyr_down_host when on != 0 */
int bo_set_resize_cpu(bo_ctx* c, int on);
void bo_destroy(bo_ctx* c);
int bo_num_scales(const bo_ctx* c);
int bo_get_scale(const bo_ctx* c, int i, bo_scale* out);
int bo_set_threads(int n);

/* Stage 0: pyramid levels and per-scale crops of the uint8 frame. */
int bo_set_image(bo_ctx* c, const uint8_t* img, size_t pitch);
/* Copy the uint8 data image of scale i (dataH x dataW) after bo_set_image. */
int bo_get_level_image(const bo_ctx* c, int i, uint8_t* dst);
/* Stage 1: complex-cell maps of scale i, 8 maps of fftH x fftW (row stride fftW). */
int bo_filter_scale(bo_ctx* c, int i, float* maps);
/* Stage 2: end-stopping, inhibition and inhibit from 8 fftH x fftW maps -> D, S (dataH x dataW). */
int bo_cells_scale(const bo_ctx* c, int i, const float* maps, float* D, float* S);
/* Stage 3: NMS of D for scale i; appends at out[*count..] (writes only below cap), *count
 * grows by the number of keypoints found. */
int bo_nms_scale(const bo_ctx* c, int i, const float* D, bo_keypoint* out, int cap, int* count);
/* The reference's CPU keypoint mode for scale i (cuda_bimp.cu:745-808) on the D and S maps;
 * appends the D keypoints (class_id -1) like bo_nms_scale. */
int bo_nms_cpu_mode(const bo_ctx* c, int i, const float* D, const float* S, bo_keypoint* out,
                    int cap, int* count);
/* Whole frame: stages 0-3 over every scale in ascending lambda order.  D_out / S_out, if
 * non-null, receive the per-scale maps concatenated (sum of dataH*dataW). */
int bo_detect(bo_ctx* c, const uint8_t* img, size_t pitch, bo_keypoint* out, int cap,
              int* count, float* D_out, float* S_out);

/* Precision form of the library (no counterpart in the reference, which filters in FP32 via
 * cuFFT).  Taps: the even kernel Ke (nsig <= 0: the whole kW x kW kernel) or, truncated to
 * the centred odd extent ext = ceil(nsig * 0.56 * lambda) < kW, the carrier Ke + dc inside
 * the window; the odd kernel Ko.  fscale = max |tap| over the 16 filters (as float); each
 * tap is stored as round((float)(tap / fscale)) to BF16 (fp16 = 0) or FP16, nearest-even.
 * taps receives the rounded values as 16 kW x kW kernels (filter 2o even, 2o+1 odd; zero
 * outside the extent).  dc receives 16 box coefficients, one per filter f,
 *   dc_f = (fscale * sum(rounded taps of f) - sum(exact kW x kW kernel of f)) / kW^2
 * which make the response to a constant image that of the exact full kernel. */
int bo_direct_taps(const bo_scale* s, float nsig, int fp16, int* ext, double* fscale,
                   double* dc, double* taps);
/* Complex-cell maps of scale i in the precision form, computed in FP64:
 *   e_o = fscale * sum(tap * u8 sample) - dc_2o * (kW x kW box sum),
 *   o_o = fscale * sum(tap * u8 sample) - dc_2o+1 * (kW x kW box sum),
 *   C_o = (255/65536) * sqrt(e_o^2 + o_o^2)
 * over the periodic padded image, 8 maps of fftH x fftW as bo_filter_scale. */
int bo_direct_filter_scale(bo_ctx* c, int i, float nsig, int fp16, float* maps);

/* Colour histogram of keypoints, addColorFeatures_kernel (cuda_bimp.cuh:8-71): keypoint i of
 * the 7-row SURF_CUDA matrix kp (leading dimension ld) gets 24 floats at desc + i*desc_ld:
 * 8 bins of B, then G, then R over the pixels (ix, iy) with ix in [(int)x - p, (int)x + p),
 * iy likewise, p = (int)(log2(size) / 2), inside the image.  A channel byte is read as a
 * signed char v and counts in bin k when 31.875*k < v < 31.875*(k+1) (so bytes 0 and >= 128
 * count nowhere); bins are then scaled by 0.5 / (3 * pixels inside), in double.  The
 * reference also reads column `width` and row `height` (its bounds are inclusive), outside
 * the image; they are skipped here. */
int bo_color_features(const float* kp, int ld, int n, const uint8_t* bgr, size_t pitch,
                      int width, int height, float* desc, int desc_ld);
/* getGPUImage's minimum-size upscale (cuda_bimp.cu:156-168): cv::resize INTER_LINEAR of an
 * 8-bit frame (cn interleaved channels) by factors fx, fy, restated from its published
 * definition (OpenCV 3.3.1; parity with OpenCV itself unpinned): per axis the source
 * coordinate f = (float)((d + 0.5) / factor - 0.5), s = floor(f), fraction f - s (0 and s
 * clamped at either end), weights round((1 - frac) * 2048) and round(frac * 2048) (each to
 * nearest even), value = (wy0 * (wx0 * p00 + wx1 * p01) + wy1 * (wx0 * p10 + wx1 * p11)
 * + 2^21) >> 22. */
int bo_resize_linear(const uint8_t* src, size_t spitch, int sw, int sh, int cn, uint8_t* dst,
                     size_t dpitch, int dw, int dh, double fx, double fy);
/* Integral image (exclusive: (H+1) x (W+1), first row and column 0) of an 8-bit frame. */
int bo_integral_image(const uint8_t* img, size_t pitch, int width, int height, long* ii);
/* SURF descriptors (64 floats, unit length) of keypoints of a 7-row matrix, following the
 * published method (Bay et al. 2008) as the library states it (surf.hip): orientation from
 * Haar responses in a disc of radius 6s and 72 windows of pi/3 (upright = 0), 4 x 4
 * subregions of 5 x 5 samples.  angle, if not null, receives each keypoint's angle. */
int bo_surf_descriptors(const uint8_t* img, size_t pitch, int width, int height, const float* kp,
                        int ld, int n, int upright, float* desc, int desc_ld, double* angle);
/* Brute-force 2-nearest-neighbour matching by L2 distance (cv::cuda BFMatcher knnMatch with
 * k = 2, bimp.cpp:188-205): for each query row, the train rows with the smallest and second
 * smallest distance sqrt(sum (q - t)^2), ties to the lower train index; sums in FP64.
 * idx2 / dist2 receive 2 entries per query (index -1 and +inf when there are fewer rows). */
int bo_knn2(const float* query, int nq, int q_ld, const float* train, int nt, int t_ld, int dim,
            int* idx2, float* dist2);

#ifdef __cplusplus
}
#endif
#endif
