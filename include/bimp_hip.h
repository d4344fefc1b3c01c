/*
 * bimp_hip.h -- C ABI of the MI355X-native BIMP keypoint extractor (libbimp_hip.so).
 *
 * The drop-in boundary for the reference library daniel-koudouna/cuda-bimp.  Each entry
 * point names the reference interface it replaces (file:line, basenames: cuda_bimp.hpp =
 * bimp/include/detection/cuda_bimp.hpp, cuda_bimp.cu = bimp/src/cuda_bimp.cu, bimp.hpp /
 * bimp.cpp = the Context API, utils.cpp = keypoint layout converters).  Plain pointers and
 * sizes only; every call returns a status and never exits the process.  The C++ Context
 * facade over these calls is include/bimp/bimp.hpp.
 *
 * Threading: one context per thread/stream.  The library holds no global mutable state;
 * calls on one context must not overlap.
 */
#ifndef BIMP_HIP_H
#define BIMP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t without pulling in the HIP headers. */
typedef struct ihipStream_t* bimp_stream_t;

#define BIMP_NUM_ORI 8              /* NUM_ORI,          cuda_bimp.hpp:10 */
#define BIMP_KP_VECTOR_LENGTH 7     /* KP_VECTOR_LENGTH, cuda_bimp.hpp:13 */
#define BIMP_KP_MAX_POINTS 8000     /* KP_MAX_POINTS,    cuda_bimp.hpp:14 */

/* Status codes (the reference prints and exit()s inside helper_cuda.h macros). */
enum {
    BIMP_OK = 0,
    BIMP_W_OVERFLOW = 1,      /* more keypoints than the cap: count is still the total
                                 (cuda_bimp.cu:132-138 warns and clamps) */
    BIMP_E_ARG = -1,          /* invalid argument */
    BIMP_E_ALLOC = -2,        /* device or host allocation failed */
    BIMP_E_DEVICE = -3,       /* HIP runtime error (no device, launch failure, ...) */
    BIMP_E_UNSUPPORTED = -4   /* transform size outside the supported set */
};

/* Map selectors for bimp_get_map / bimp_set_map. */
enum {
    BIMP_MAP_DOUBLE = 0,  /* double-stopped response D after inhibition (d_ResultDouble) */
    BIMP_MAP_SINGLE = 1,  /* single-stopped response S after inhibition (d_ResultSingle) */
    BIMP_MAP_COMPLEX = 2  /* complex-cell magnitude C_o of one orientation (d_ResultComplex[o]) */
};

/* cv::KeyPoint field order: pt.x, pt.y, size, angle, response, octave, class_id
 * (utils.cpp:130-150 fills cv::KeyPoint from the 7-row SURF_CUDA matrix). */
typedef struct bimp_keypoint {
    float x, y;
    float size;
    float angle;
    float response;
    int octave;
    int class_id;
} bimp_keypoint;

/* Run-time options; the reference hard-codes these (cuda_bimp.cu:236-247, :481, :721). */
typedef struct bimp_options {
    float gamma;          /* 0.5: envelope aspect (not squared) */
    float sigma_ratio;    /* 0.56: sigma = sigma_ratio * lambda */
    float support;        /* 7: filter size = support * lambda (+1 if even) */
    float pyr_threshold;  /* 7.5: halve the frame while lambda/pyrstep >= this */
    double nms_threshold; /* 0.104: keypoint threshold relative to the scale's max (the
                             reference multiplies in double, cuda_bimp.cu:721) */
    float input_scale;    /* 255/65536: uint8 -> float conversion factor */
    int flags;            /* BIMP_FLAG_* */
    int precision;        /* BIMP_PRECISION_*: arithmetic of the Gabor bank (the reference
                             runs FP32 cuFFT, cuda_bimp.cu:486-507) */
    float direct_max_lambda; /* low-precision forms: scales whose requested wavelength is
                                at most this (full-resolution pixels) run the direct form */
    float direct_sigmas;  /* low-precision forms: 0 convolves with the full kW x kW kernel;
                             s > 0 truncates the taps to an odd extent of ceil(s*sigma)
                             (default 8).  Either way an exact integer box sum restores every
                             filter's response to a constant image */
} bimp_options;

#define BIMP_FLAG_PROFILE 1  /* time every kernel with HIP events (see bimp_get_profile) */
/* Every launch on the caller's stream.  Without it, a context whose scales span pyramid level
 * 0 and higher levels runs part of a frame on a second stream of its own: the pyramid, the
 * higher levels' work (FFT groups, or the precision form's direct launch for those levels,
 * with their cells), and level-0 cells stages beside the remaining level-0 transforms.  That
 * stream is forked from the caller's stream once the frame is on the device and joined back
 * before the keypoint pass (and not used while the caller's stream is being captured into a
 * graph).  Read at bimp_create (no second stream
 * is made, and the precision form keeps one direct launch for all levels) and at every frame.
 * The batching form's contexts set it. */
#define BIMP_FLAG_SINGLE_STREAM 2
/* Replay each frame of bimp_detect_u8, bimp_detect_u8_matrix and bimp_time_detect(_frames)
 * as one hipGraph instead of issuing its ~30 launches (no counterpart in the reference).  The
 * graph is built on the first such frame and again whenever the frame or output pointers, the
 * capacity or the flags change; host frames are uploaded to the context's own level-0 buffer,
 * so successive host frames reuse one graph.  With a second stream the graph has two
 * independent branches (the work above pyramid level 0 and the level-0 work) between the
 * per-frame reset and the keypoint pass.  Ignored while BIMP_FLAG_PROFILE is set; not used
 * by bimp_detect_u8_noalloc. */
#define BIMP_FLAG_GRAPH 4
/* Pyramid levels rounded as OpenCV's host cv::pyrDown rounds, (S + 128) >> 8 (half up), in
 * place of cv::cuda::pyrDown's round half to even: the reference's resize_cpu option
 * (setupCUDAImages, cuda_bimp.cu:649-687, selected by getKeypoints(..., cpu_mode, resize_cpu),
 * cuda_bimp.hpp:109-111, and Context::resize_cpu, bimp.hpp:24).  Computed on the GPU.  The
 * reference honours it only in its CPU keypoint mode; the C++ facade sets it only there. */
#define BIMP_FLAG_RESIZE_CPU 8

/* Arithmetic of the Gabor bank.  FP32: FFT convolution in FP32 (the reference's form).
 * BF16 / FP16: direct spatial convolution on the matrix cores, uint8 samples times taps
 * rounded once to BF16 / FP16, accumulated in FP32; used for every scale with
 * lambda <= direct_max_lambda whose kernel width kW is at most BIMP_DIRECT_MAX_KERNEL, the
 * others keep the FP32 FFT form. */
#define BIMP_PRECISION_FP32 0
#define BIMP_PRECISION_BF16 1
#define BIMP_PRECISION_FP16 2
#define BIMP_DIRECT_MAX_KERNEL 55

typedef struct bimp_scale_info {
    float lambda2;   /* requested wavelength (full resolution) */
    float lambda;    /* wavelength at the pyramid level */
    float pyrstep;   /* 1, 2, 4, ... */
    int kernel_size; /* kW (odd) */
    int kernel_centre;
    int border;      /* NMS border = (int)lambda */
    int level;       /* pyramid level = log2(pyrstep) */
    int data_h, data_w;
    int fft_h, fft_w;
    int map_h, map_w, map_pitch; /* stored complex-cell region and its row pitch (floats) */
    int precision;               /* BIMP_PRECISION_* this scale runs in */
    int tap_extent;              /* direct forms: convolved tap extent (odd), else kW */
} bimp_scale_info;

typedef struct bimp_ctx bimp_ctx;

/* Defaults equal to the reference constants. */
void bimp_default_options(bimp_options* opt);

/* snapTransformSize (cuda_bimp.hpp:140, cuda_bimp.cu:200-219). */
int bimp_snap_transform_size(int n);

/* makeLambdasLog (cuda_bimp.hpp:146, cuda_bimp.cu:927-937): writes up to cap wavelengths,
 * returns how many the reference loop produces. */
int bimp_make_lambdas_log(float start, float end, int scales_per_octave, float* out, int cap);

/* CudaStuff::CudaStuff(width, height, lambdas) (cuda_bimp.hpp:97, cuda_bimp.cu:222-397):
 * sorts the wavelengths, derives per-scale parameters, builds the Gabor bank spectra
 * (FP64 on the host, stored FP32 on the device) and allocates every buffer a frame needs.
 * opt may be NULL (reference defaults). */
int bimp_create(int width, int height, const float* lambdas, int n, const bimp_options* opt,
                bimp_ctx** out);

/* CudaStuff::~CudaStuff (cuda_bimp.cu:399-460). */
int bimp_destroy(bimp_ctx* ctx);

int bimp_num_scales(const bimp_ctx* ctx);
int bimp_get_scale_info(const bimp_ctx* ctx, int scale, bimp_scale_info* out);

/* bimp::cuda::getKeypoints (cuda_bimp.hpp:107-111, cuda_bimp.cu:109-140) followed by
 * utils::downloadKeypoints (utils.cpp:130-150): runs the whole frame (pyramid, Gabor bank,
 * complex cells, end-stopping, inhibition, NMS) on `stream` and writes min(total, cap)
 * keypoints ordered by (scale, row, column).  img is width x height uint8 with row pitch
 * `pitch` bytes, on the device when img_on_device != 0.  out is device memory when
 * out_on_device != 0 (then *count is written after a stream synchronisation).
 * Returns BIMP_W_OVERFLOW when total > cap; *count is the total either way. */
int bimp_detect_u8(bimp_ctx* ctx, const uint8_t* img, size_t pitch, int img_on_device,
                   bimp_keypoint* out, int cap, int out_on_device, int* count,
                   bimp_stream_t stream);

/* As bimp_detect_u8 with the keypoints delivered to host memory, without any device
 * allocation: no counterpart in the reference.  For callers that cannot allocate (serving
 * loops, code running between captured graphs, harnesses that forbid hipMalloc): the keypoint
 * records are staged in a static device array of 65536 records (1.75 MiB, one per device, held
 * until the process ends) instead of the context's lazily allocated buffer.  cap above 65536
 * returns BIMP_E_ARG; a pitched host frame (which needs pinned staging the context allocates
 * on first use) and a context with BIMP_FLAG_PROFILE return BIMP_E_UNSUPPORTED.  Calls on one
 * device share the staging: calls that overlap on one device race. */
int bimp_detect_u8_noalloc(bimp_ctx* ctx, const uint8_t* img, size_t pitch, int img_on_device,
                           bimp_keypoint* out, int cap, int* count, bimp_stream_t stream);
/* Same frame with the 7-row SURF_CUDA keypoint matrix the reference returns from
 * getKeypoints (cuda_bimp.cu:129-140): rows X, Y, LAPLACIAN (int bits), OCTAVE (int bits),
 * SIZE, ANGLE, HESSIAN, each of leading dimension ld >= cap floats. */
int bimp_detect_u8_matrix(bimp_ctx* ctx, const uint8_t* img, size_t pitch, int img_on_device,
                          float* mat, int ld, int cap, int mat_on_device, int* count,
                          bimp_stream_t stream);

/* Stage entry points of DoFullGPUFiltering (cuda_bimp.cu:462-559), for callers that want
 * the response maps (Context::getResponses, bimp.cpp:94-104) or that inject maps:
 *   bimp_filter_u8  -- pyramid + filter bank + complex cells (cuda_bimp.cu:466-524)
 *   bimp_cells      -- end-stopping + inhibition from the current complex maps (:526-552)
 *   bimp_keypoints  -- NMS + compaction from the current D maps (:554, :714-743) */
int bimp_filter_u8(bimp_ctx* ctx, const uint8_t* img, size_t pitch, int img_on_device,
                   bimp_stream_t stream);
int bimp_cells(bimp_ctx* ctx, bimp_stream_t stream);
int bimp_keypoints(bimp_ctx* ctx, bimp_keypoint* out, int cap, int out_on_device, int* count,
                   bimp_stream_t stream);

/* Read or replace one response map of one scale (d_ResultDouble / d_ResultSingle /
 * d_ResultComplex[o], cuda_bimp.hpp:50-53).  DOUBLE/SINGLE maps are data_h x data_w;
 * COMPLEX maps are map_h x map_w.  pitch is in floats.  Synchronous on `stream`. */
int bimp_get_map(bimp_ctx* ctx, int scale, int which, int orientation, float* dst,
                 size_t pitch, int dst_on_device, bimp_stream_t stream);
/* Device view of the same map (Context::getResponses hands out the context's own device
 * pointers, bimp.cpp:94-104): pointer, pitch in floats, rows and columns.  The view stays valid
 * until the context is destroyed; its contents change with each frame. */
int bimp_map_view(const bimp_ctx* ctx, int scale, int which, int orientation, const float** d_map,
                  size_t* pitch, int* rows, int* cols);
int bimp_set_map(bimp_ctx* ctx, int scale, int which, int orientation, const float* src,
                 size_t pitch, int src_on_device, bimp_stream_t stream);

/* utils::uploadKeypoints / downloadKeypoints layout converters (utils.cpp:109-150),
 * host memory: keypoints <-> 7 x ld SURF_CUDA matrix. */
int bimp_keypoints_to_matrix(const bimp_keypoint* kps, int n, float* mat, int ld);
int bimp_matrix_to_keypoints(const float* mat, int ld, int n, bimp_keypoint* kps);

/* Message of the last failure on ctx; with ctx == NULL, of the last failed bimp_create on the
 * calling thread. */
const char* bimp_last_error(const bimp_ctx* ctx);
const char* bimp_status_string(int status);

/* ---- profiling (replaces utils::log / showTimes, utils.cpp:11-30, :54-85) ---- */
typedef struct bimp_profile_entry {
    char name[32];      /* kernel family */
    int launches;       /* launches accumulated */
    double total_ms;    /* summed HIP-event time */
    double bytes;       /* summed algorithmic bytes of those launches */
} bimp_profile_entry;
/* With BIMP_FLAG_PROFILE set, every launch is bracketed by HIP events on the stream it runs
 * on; this returns the accumulated per-family totals (waits for every bracketed launch).
 * Launches on the context's two streams overlap, so set BIMP_FLAG_SINGLE_STREAM as well for
 * durations that are each launch's own. */
int bimp_get_profile(bimp_ctx* ctx, bimp_profile_entry* out, int cap, int* n);
int bimp_reset_profile(bimp_ctx* ctx);
int bimp_set_flags(bimp_ctx* ctx, int flags);

/* ---- device plumbing for hosts without a GPU array library (no counterpart in the
 * reference, which relies on cv::cuda::GpuMat) ---- */
int bimp_device_count(int* n);
int bimp_device_malloc(void** p, size_t bytes);
int bimp_device_free(void* p);
int bimp_memcpy_h2d(void* dst, const void* src, size_t bytes);
/* rows x width_bytes from pitched host memory to pitched device memory (synchronous). */
int bimp_memcpy_h2d_2d(void* dst, size_t dpitch, const void* src, size_t spitch,
                       size_t width_bytes, size_t rows);
int bimp_memcpy_d2h(void* dst, const void* src, size_t bytes);
int bimp_stream_create(bimp_stream_t* s);
int bimp_stream_destroy(bimp_stream_t s);
int bimp_stream_sync(bimp_stream_t s);
/* Time `iters` back-to-back detects (device-resident frame and keypoint buffer) with HIP
 * events on `stream`; writes the elapsed milliseconds. */
int bimp_time_detect(bimp_ctx* ctx, const uint8_t* d_img, size_t pitch, bimp_keypoint* d_out,
                     int cap, int iters, bimp_stream_t stream, float* elapsed_ms);
/* As bimp_time_detect, with one event pair per frame: frame_ms[k] is frame k's elapsed
 * milliseconds (iters entries; frames run back to back, so their sum is the whole run). */
int bimp_time_detect_frames(bimp_ctx* ctx, const uint8_t* d_img, size_t pitch,
                            bimp_keypoint* d_out, int cap, int iters, bimp_stream_t stream,
                            float* frame_ms);
/* Pyramid level `level` (1 .. deepest level of the bank) of the last filtered frame, copied to
 * host memory (uint8, (w+1)/2 x (h+1)/2 per level, row pitch `pitch`); synchronises the stream.
 * The reference computes it with cv::cuda::pyrDown (cuda_bimp.cu:473-479) and does not expose
 * it. */
int bimp_get_level(bimp_ctx* ctx, int level, uint8_t* dst, size_t pitch, bimp_stream_t stream);
/* A plain device copy kernel of known size, used to calibrate HBM byte counters. */
int bimp_copy_kernel(void* dst, const void* src, size_t bytes, bimp_stream_t stream);

/* Seeded synthetic frame generator for tests and benchmarks (no counterpart in the
 * reference, which ships no images): background U[0,255], nshapes rectangles/ellipses,
 * N(0,5^2) noise, uint8.  Host memory. */
int bimp_synth_frame(int width, int height, int nshapes, uint64_t seed, uint8_t* out,
                     size_t pitch);

/* The reference's CPU keypoint mode (getKeypoints with cpu_mode, cuda_bimp.cu:117-124;
 * cuda_keypoints cuda_bimp.cu:745-808): after bimp_filter_u8 + bimp_cells (or a detect), D
 * and S are copied to the host and keypoints are extracted there with that mode's rules
 * (dilation-based maxima, threshold 0.104*max(D) for both maps, inclusive border off,
 * S-map offsets added to the D-map position when both peak, class_id -1), in (scale, row,
 * column) order.  Synchronises the stream. */
int bimp_keypoints_cpu_mode(bimp_ctx* ctx, bimp_keypoint* out, int cap, int* count,
                            bimp_stream_t stream);

/* getGPUImage's colour step (cuda_bimp.cu:156-179, cv::cuda::cvtColor COLOR_BGR2GRAY): 8-bit
 * interleaved BGR (host or device) -> 8-bit gray on the device, with OpenCV's Q14 weights.
 * The MIN_SIZE upscale of getGPUImage is not provided. */
int bimp_bgr_to_gray(const uint8_t* bgr, size_t bgr_pitch, int width, int height, int bgr_on_device,
                     uint8_t* d_gray, size_t gray_pitch, bimp_stream_t stream);

/* ---- Frame preparation (getGPUImage, cuda_bimp.cu:156-177) ------------------------------
 * The reference upscales a frame with rows < 350 <= cols by 350/rows, else one with
 * cols < 350 by 350/cols (a float factor), with cv::resize's bilinear filter, then converts
 * BGR to gray on the device.  A Context keeps its CudaStuff sized from the frame before the
 * upscale (bimp.cpp:8-10), so detection reads the top-left width x height of the upscaled
 * frame: pass the prepared frame and its pitch to bimp_detect_u8 with the original size. */
/* Size of the prepared frame (the input size when no upscale applies) and the factor (1 when
 * none). */
int bimp_min_size_shape(int width, int height, int* out_w, int* out_h, float* scale);
/* Bilinear resize of a device 8-bit frame of 1 or 3 interleaved channels (cv::resize
 * INTER_LINEAR: source coordinate (x + 0.5) / fx - 0.5 clamped to the frame, 11-bit fixed-point
 * weights, one rounding of the two-axis sum).  fx, fy > 0 give the factors (as cv::resize's
 * fx, fy); otherwise the factors are dw/sw and dh/sh.  No synchronisation. */
int bimp_resize_linear_u8(const uint8_t* d_src, size_t src_pitch, int sw, int sh, int channels,
                          uint8_t* d_dst, size_t dst_pitch, int dw, int dh, double fx, double fy,
                          bimp_stream_t stream);
/* getGPUImage on a device frame: the minimum-size upscale, then BGR -> gray when channels is
 * 3 and grayscale is set (otherwise the channels are kept).  d_dst holds the prepared frame
 * (bimp_min_size_shape gives its size).  No synchronisation. */
int bimp_get_gpu_image(const uint8_t* d_src, size_t src_pitch, int width, int height,
                       int channels, int grayscale, uint8_t* d_dst, size_t dst_pitch,
                       bimp_stream_t stream);

/* ---- Descriptors and matching (the keypoints' consumers) --------------------------------
 * Context::getDescriptors (bimp.cpp:32-59) appends BIMP_COLOR_DIMENSIONS colour-histogram
 * floats to each keypoint's descriptor (columns SURF_DIMENSIONS..DIMENSIONS of a
 * DIMENSIONS-wide row, cuda_bimp.hpp:17-19). */
#define BIMP_SURF_DIMENSIONS 128
#define BIMP_COLOR_DIMENSIONS 24
#define BIMP_DIMENSIONS (BIMP_SURF_DIMENSIONS + BIMP_COLOR_DIMENSIONS)

/* cv::DMatch field order (queryIdx, trainIdx, imgIdx, distance). */
typedef struct {
    int query_idx, train_idx, img_idx;
    float distance;
} bimp_match;

/* addColorFeatures (cuda_bimp.cu:952-958, kernel cuda_bimp.cuh:8-71): for the n keypoints of
 * a device 7-row SURF_CUDA matrix (leading dimension ld), 24 floats per keypoint at
 * d_desc + i*desc_ld: 8 bins of B, G, R over the pixels of a (2p) x (2p) window at
 * ((int)x - p, (int)y - p), p = (int)(log2(size)/2), of a device BGR frame (3 bytes per pixel,
 * row pitch bgr_pitch).  The reference's arithmetic: a byte is read as a signed char v and
 * counts in bin k when 31.875k < v < 31.875(k+1) (bytes 0 and >= 128 count nowhere), and bins
 * are scaled by 0.5 / (3 x pixels inside the image).  Window pixels outside the frame are
 * skipped (the reference's inclusive bounds read one column and one row past it).  No
 * synchronisation. */
int bimp_color_features(const float* d_kp, int ld, int n, const uint8_t* d_bgr, size_t bgr_pitch,
                        int width, int height, float* d_desc, int desc_ld, bimp_stream_t stream);
/* Integral image of a device 8-bit frame: (height+1) x (width+1) uint32 sums, first row and
 * column 0 (SURF's box filters; width*height*255 must stay below 2^31).  No synchronisation. */
int bimp_integral_image(const uint8_t* d_img, size_t pitch, int width, int height,
                        unsigned* d_integral, bimp_stream_t stream);
/* SURF descriptors of given keypoints (the reference runs cv::cuda::SURF_CUDA with
 * useProvidedKeypoints on the BIMP keypoints, bimp.cpp:32-59, :188-199), restated from the
 * published method (Bay et al. 2008; OpenCV's code is not followed): scale 1.2*size/9, the
 * orientation from Haar responses of a disc of radius 6s (72 windows of pi/3; upright != 0 keeps
 * angle 0), 4 x 4 subregions of 5 x 5 rotated samples summing dx, dy, |dx|, |dy|, unit length.
 * 64 floats per keypoint at d_desc + i*desc_ld; d_angle (may be NULL) gets the angles in
 * radians.  Keypoints are a device 7-row SURF_CUDA matrix (leading dimension ld).  No
 * synchronisation. */
int bimp_surf_descriptors(const unsigned* d_integral, int width, int height, const float* d_kp,
                          int ld, int n, int upright, float* d_desc, int desc_ld, float* d_angle,
                          bimp_stream_t stream);
/* Brute-force 2-nearest-neighbour matching by L2 distance (cv::cuda::DescriptorMatcher
 * createBFMatcher + knnMatchAsync with k = 2, bimp.cpp:202-204): for each of nq device query
 * rows (q_ld floats apart) the two train rows (t_ld apart) with the smallest distances,
 * sqrt of the FP32 sum of squared differences over dim dimensions in order; ties go to the
 * lower train index.  d_idx2 / d_dist2 get 2 entries per query (index -1 and +inf when there
 * is no second row).  No synchronisation. */
int bimp_match_knn2(const float* d_query, int nq, int q_ld, const float* d_train, int nt,
                    int t_ld, int dim, int* d_idx2, float* d_dist2, bimp_stream_t stream);
/* downloadMatches (bimp.cpp:208-240), host memory: query i keeps (i, idx2[2i], 0, dist2[2i])
 * when dist2[2i] < ratio * dist2[2i+1] (the reference's ratio is 0.6, compared in double).
 * Returns BIMP_W_OVERFLOW with *count the total when more than cap matches pass. */
int bimp_ratio_matches(const int* idx2, const float* dist2, int nq, double ratio,
                       bimp_match* out, int cap, int* count);

/* ---- Batching form --------------------------------------------------------------------
 * Many frames of uneven sizes in one launch.  The reference has no batch entry: its caller
 * runs getKeypoints frame by frame (example_bimp.cpp:48-54, cuda_bimp.cu:109-140); a batch
 * gives every frame the same result as bimp_detect_u8 on a context of that frame's size.
 * The batch owns one device input slot per frame (fill it, then run), per-frame keypoint
 * outputs (cap records each, (scale, row, column) order) and per-frame counts.  The whole
 * batch is one hipGraph: one branch per context, frames of one size spread over that size's
 * contexts.  One context per frame size; tune != 0 (a plan search over contexts per size)
 * is switched off and returns BIMP_E_UNSUPPORTED.  With BIMP_TRACE set in the environment,
 * creation prints its steps to stderr. */
typedef struct bimp_batch bimp_batch;

int bimp_batch_create(const int* widths, const int* heights, int nframes, const float* lambdas,
                      int n, const bimp_options* opt, int cap, int tune, bimp_batch** out);
int bimp_batch_destroy(bimp_batch* batch);
/* Device input slot of frame i (uint8, row pitch *pitch bytes). */
int bimp_batch_frame(bimp_batch* batch, int i, uint8_t** d_frame, size_t* pitch);
/* Launch the batch graph on `stream` (no host synchronisation). */
int bimp_batch_run(bimp_batch* batch, bimp_stream_t stream);
/* After the stream has finished: frame i's keypoints (min(count, cap) records) and count. */
int bimp_batch_results(bimp_batch* batch, int i, bimp_keypoint* out, int cap, int* count);
/* Device outputs of frame i (cap records, then the count as one int). */
int bimp_batch_outputs(bimp_batch* batch, int i, bimp_keypoint** d_keypoints, int** d_count);
/* Time `iters` batch launches (elapsed milliseconds per batch) and, when frame_ms is not NULL,
 * each frame alone as its own graph launch (milliseconds, nframes entries). */
int bimp_batch_time(bimp_batch* batch, int iters, bimp_stream_t stream, float* batch_ms,
                    float* frame_ms);
/* Human-readable plan (contexts and sub-batches per frame size), NUL-terminated. */
int bimp_batch_plan(const bimp_batch* batch, char* buf, int cap);
const char* bimp_batch_last_error(const bimp_batch* batch);

/* ---- Differentiation form -------------------------------------------------------------
 * The filter bank, complex cells, end-stopping and inhibition (cuda_bimp.cu:462-552; NMS
 * excluded) as a differentiable map from float images to the per-scale response maps D, with
 * a hand-written adjoint.  No counterpart in the reference.  Images are nimg x height x width
 * FP32 (device, contiguous), scaled by 255 * input_scale before filtering (an image u8/255
 * gives the uint8 path's maps up to the pyramid's rounding; levels above 0 use the float
 * pyramid).  Responses are nimg x bimp_diff_response_size() floats: the D maps of the scales in
 * ascending wavelength order, each data_h x data_w.  The Gabor parameters (sigma ratio, gamma)
 * start at opt's values.  The backward pass recomputes the complex responses, writes
 * d(loss)/d(image) for an upstream gradient of the responses, and the gradient with respect to
 * (sigma ratio, gamma) summed over the images into grad_params (host, 2 doubles).  It
 * synchronises the stream. */
typedef struct bimp_diff bimp_diff;
int bimp_diff_create(int width, int height, const float* lambdas, int n, const bimp_options* opt,
                     bimp_diff** out);
int bimp_diff_destroy(bimp_diff* d);
long long bimp_diff_response_size(const bimp_diff* d);
int bimp_diff_set_params(bimp_diff* d, float sigma_ratio, float gamma);
int bimp_diff_forward(bimp_diff* d, const float* d_img, int nimg, float* d_resp,
                      bimp_stream_t stream);
int bimp_diff_backward(bimp_diff* d, const float* d_img, int nimg, const float* d_grad_resp,
                       float* d_grad_img, double* grad_params, bimp_stream_t stream);
const char* bimp_diff_last_error(const bimp_diff* d);

#ifdef __cplusplus
}
#endif
#endif
