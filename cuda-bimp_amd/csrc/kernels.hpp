// kernels.hpp -- launch interface between the host context and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace bimpk {

constexpr int kNumOri = 8;
constexpr int kUniqueOri = 5;  // orientations 5..7 mirror 3..1 in the spectrum

// Scales of one resolution group (same pyramid level and transform size) processed by one
// launch of each FFT stage.
constexpr int kMaxBatch = 8;
struct ScaleSet {
    int n;
    int kX[kMaxBatch];       // kernel centre = clamp-padding width
    const float* G[kMaxBatch];
    float2* DT[kMaxBatch];
    float* C[kMaxBatch];
};

// Per-scale constants of the end-stop / inhibition stencil (host-derived).
struct CellParams {
    int ea[kNumOri], eb[kNumOri];  // end-stop offsets (rows, cols)
    int ia[kNumOri], ib[kNumOri];  // inhibition offsets (rows, cols)
    int ra[kNumOri], rb[kNumOri];  // radial half offsets
};

// Global-memory pointer qualifier: loads and stores through such pointers are global_*
// instructions, counted by vmcnt only (through a generic pointer they are flat_* and also
// hold up every LDS wait).
#define BIMP_GLOBAL __attribute__((address_space(1)))

// Per-scale descriptor for the NMS kernels (device table).
struct NmsScale {
    const float* D;     // dataH x dataW response map
    int H, W, off;
    int block0;         // first NMS block of this scale
    int nblocks;
    float lambda, pyrstep;
    int slot;           // index of the scale's max slot
};

// Dynamic LDS a workgroup may request (160 KB per CU less the runtime's reserve).
constexpr size_t kMaxLds = 160 * 1024 - 1024;

// One scale of the precision form (direct MFMA convolution, see direct.hip).
// Workgroup tile: kDirTY output rows (passes of kDirRows rows per wavefront) x kDirGroups x 16
// columns.
constexpr int kDirRows = 4, kDirGroups = 8;
constexpr int kDirTY = 32, kDirTX = kDirGroups * 16;
// widest staged image row (elements): tile columns plus the kernel's 55-tap support, rounded
constexpr int kDirMaxLwp = 256;
struct DirectScale {
    int H, W, FH, FW;        // data size, transform size (the periodic domain)
    int kpad, kW;            // kernel centre kX (clamp padding, box half-width) and width
    int Hc, Wc, ldc;         // complex-cell map region and pitch
    int off, ext, nvb, nrho; // tap window offset kX - R, extent, 8-column blocks, 4-row blocks
    int LH, lwp, vsp;        // staged rows, LDS row pitch (elements), box-sum row pitch (floats)
    int cstride;             // element stride between the two shifted copies
    float fscale, iscale;    // filter scale, input scale (255/65536)
    float dc[2 * kNumOri];   // box coefficients, filter 2o even, 2o+1 odd
};

// One scale of a direct launch: its geometry, filter fragments, output maps, tile range and
// pyramid level.
struct DirectJob {
    DirectScale ds;
    const void* frag;
    float* C;
    int tiles_x, tile0;
    int level;
};
// The pyramid levels a direct launch reads (level 0 is the caller's frame).
constexpr int kDirMaxLevels = 8;
struct DirectLevels {
    const uint8_t* img[kDirMaxLevels];
    size_t pitch[kDirMaxLevels];
};

// Fills off/nvb/nrho/LH/lwp/vsp/cstride from kpad, kW, ext; returns the dynamic LDS bytes.
size_t direct_geometry(DirectScale& ds);
// Job table of n scales (any pyramid levels), heaviest first; returns the launch's LDS bytes
// (0 if a scale does not fit) and the total tile count.
size_t direct_jobs(const DirectScale* ds, const void* const* frag, float* const* C,
                   const int* level, int n, DirectJob* jobs, int* total_tiles);
hipError_t prepare_direct();
hipError_t launch_direct(const DirectLevels& lv, const DirectJob* d_jobs, int njobs,
                         int total_tiles, size_t lds, bool fp16, hipStream_t s);

constexpr int kNmsThreads = 256;
constexpr int kNmsPerThread = 16;
constexpr int kNmsBlockPixels = kNmsThreads * kNmsPerThread;

bool fft_size_supported(int n);
// Packed per-stage twiddle table of the length-n transform (FP64-evaluated, FP32-stored).
// Transform lengths with a plan (fft_device.hpp): every kernel family is instantiated for each.
#define BIMP_FOR_SIZES(X) \
    X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(1536) X(2048) X(2560) X(3072) X(3584) X(4096)
bool twiddle_table(int n, std::vector<float2>& out);
// Register the dynamic-LDS limits of every kernel instance for size n (once per process).
hipError_t prepare_size(int n);
hipError_t prepare_common();

hipError_t launch_init(unsigned* maxslots, int nslots, int* counter, hipStream_t s);
// half_up: round (S + 128) >> 8 as OpenCV's host pyrDown; else half to even as its CUDA one
hipError_t launch_pyr_down(const uint8_t* src, int sw, int sh, size_t spitch, uint8_t* dst,
                           int dw, int dh, size_t dpitch, bool half_up, hipStream_t s);
hipError_t launch_fwd_rows(int FW, const uint8_t* img, size_t pitch, int H, int W,
                           const ScaleSet& ss, float scale, float2* R, int ldr, size_t rstride,
                           const float2* tw, hipStream_t s);
// The same pass over a float image (pitch in floats), e.g. a differentiation-form level.
hipError_t launch_fwd_rows_f32(int FW, const float* img, size_t pitch, int H, int W,
                               const ScaleSet& ss, float scale, float2* R, int ldr,
                               size_t rstride, const float2* tw, hipStream_t s);
hipError_t launch_fwd_cols(int FH, const float2* R, int ldr, size_t rstride, int H,
                           const ScaleSet& ss, int ncols, const float2* tw, hipStream_t s);
hipError_t launch_inv_cols(int FH, const ScaleSet& ss, int FW, int Hc, float2* T,
                           size_t tstride, const float2* tw, hipStream_t s);
hipError_t launch_inv_rows(int FW, int FH, const ScaleSet& ss, const float2* T, size_t tstride,
                           int Hc, int Wc, int ldc, const float2* tw, hipStream_t s);
hipError_t launch_cells(const float* C, int Hc, int ldc, int FH, int FW, int H, int W,
                        const CellParams& p, float* Dm, float* Sm, unsigned* maxslot,
                        hipStream_t s);
// End-stop/inhibition of up to kCellMaxScales scales in one launch (the arguments of
// launch_cells per scale).
constexpr int kCellMaxScales = 8;
struct CellScale {
    const float* C;
    int Hc, ldc, FH, FW, H, W;
    CellParams p;
    float* Dm;
    float* Sm;
    unsigned* maxslot;
};
hipError_t launch_cells(const CellScale* scales, int n, hipStream_t s);
hipError_t launch_max(const float* D, int n, unsigned* maxslot, hipStream_t s);
// k_nms_count stores each block's keypoint count and each thread's hit mask (total_blocks x
// kNmsThreads words) for k_nms_write, which ranks the keypoints from the counts, stores the
// total in counter, reads the response maps only around the hits and clears the max slots
// (one per scale, nscales of them).
hipError_t launch_nms_count(const NmsScale* scales, int nscales, int total_blocks,
                            const unsigned* maxslots, double thresh, int* block_counts,
                            unsigned* masks, hipStream_t s);
hipError_t launch_nms_write(const NmsScale* scales, int nscales, int total_blocks,
                            const int* block_counts, int* counter, unsigned* maxslots,
                            const unsigned* masks, void* out, int cap, int matrix_ld,
                            hipStream_t s);
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);
hipError_t launch_bgr_to_gray(const uint8_t* bgr, size_t bpitch, int w, int h, uint8_t* gray,
                              size_t gpitch, hipStream_t s);
// dst[0] = counter[0] on the stream (a kernel, so that it can sit in a captured graph).
hipError_t launch_store_count(const int* counter, int* dst, hipStream_t s);

// column-kernel batch (columns per workgroup) for transform length n
int col_batch(int n);
int row_vectors_per_block(int n);

}  // namespace bimpk
