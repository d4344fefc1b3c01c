// context.hpp -- the per-frame-size context behind the C ABI (the reference's CudaStuff,
// bimp/src/cuda_bimp.cu:222-460) and the per-frame launch sequence shared by the single-frame
// entry points and the batching form.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/bimp_hip.h"
#include "bank.hpp"
#include "kernels.hpp"

namespace bimpi {

enum Family {
    F_INIT, F_COPY_IN, F_PYR, F_FWD_ROWS, F_FWD_COLS, F_INV_COLS, F_INV_ROWS, F_CELLS,
    F_MAX, F_NMS_COUNT, F_NMS_WRITE, F_DIRECT, F_COUNT
};
// Profile names follow the kernel symbols (size-templated kernels carry their length).
inline const char* const kFamilyNames[F_COUNT] = {"k_init",      "copy_in",    "k_pyr_down",  "k_fwd_rows",
                                     "k_fwd_cols",  "k_inv_cols", "k_inv_rows",  "k_cells",
                                     "k_max",       "k_nms_count", "k_nms_write",
                                     "k_direct"};

struct Scale {
    bimph::ScaleGeom g;
    int Hc = 0, Wc = 0, ldc = 0;
    bimpk::CellParams cp;
    float* G = nullptr;    // kUniqueOri x FW x FH
    float2* DT = nullptr;  // (FW/2+1) x FH
    float* C = nullptr;    // 8 x Hc x ldc
    float* D = nullptr;    // H x W
    float* S = nullptr;    // H x W
    // direct (low-precision) form
    int precision = BIMP_PRECISION_FP32;
    bimpk::DirectScale ds{};
    void* frag = nullptr;  // B-operand fragments of every tap block (16-bit taps)
};

// One step of a frame: an FFT resolution group or one direct-form scale.
struct Step {
    bool direct;
    int index;  // FFT: group index; direct: first scale
    int n = 1;  // direct: consecutive scales of one level in one launch
    bimpk::DirectJob* jobs = nullptr;  // direct: device job table
    int tiles = 0;
    size_t lds = 0;
};

// Scales sharing pyramid level and transform size; each FFT stage runs once per group.
struct Group {
    int first, n;
};

struct ProfEv {
    hipEvent_t a, b;
    int fam, size;
    double bytes;
};

struct ProfAcc {
    int launches = 0;
    double ms = 0, bytes = 0;
};

}  // namespace bimpi

struct bimp_ctx {
    int width = 0, height = 0;
    bimp_options opt;
    std::vector<bimpi::Scale> sc;
    std::vector<bimpi::Group> groups;
    std::vector<bimpi::Step> steps;
    int maxLevel = 0;
    std::vector<uint8_t*> level;  // level[0] is the staging copy of a host frame
    std::vector<int> levelW, levelH;
    float2* R = nullptr;
    float2* T = nullptr;
    size_t tCap = 0;  // T capacity in complex elements
    std::map<int, float2*> tw;
    unsigned* maxslots = nullptr;
    int* counter = nullptr;
    int* blockCounts = nullptr;
    unsigned* blockMasks = nullptr;  // NMS hit masks, totalBlocks x kNmsThreads
    int totalBlocks = 0;
    bimpk::NmsScale* nmsTable = nullptr;
    void* kpBuf = nullptr;
    size_t kpBufBytes = 0;
    // pinned staging of pitched host frames (a 2D copy from pageable memory is slow), and the
    // event after its last upload, waited on before the staging is rewritten
    uint8_t* hostIn = nullptr;
    hipEvent_t hostInDone = nullptr;
    bool maxValid = false;
    // the per-scale max slots are zero at this point of the stream (set at creation and by
    // the keypoint pass, which clears them once it has read them): a frame then needs no reset
    bool slotsClean = true;
    std::string err;
    // profiling
    std::vector<bimpi::ProfEv> ev;
    size_t evUsed = 0;
    std::map<std::string, bimpi::ProfAcc> prof;
    hipStream_t lastStream = nullptr;
    int invBatch = 0;  // inverse-pass sub-batch size; 0 = by the last-level-cache budget
    // FFT groups above pyramid level 0 run on a side stream (forked after the pyramid, joined
    // before the keypoint pass), with their own forward and inverse intermediates
    hipStream_t side = nullptr;
    hipEvent_t evFork = nullptr, evJoin = nullptr, evCells = nullptr;
    float2* R2 = nullptr;
    float2* T2 = nullptr;
    size_t tCap2 = 0;
    // this device's copy of the static keypoint staging of bimp_detect_u8_noalloc
    bimp_keypoint* noallocStaging = nullptr;
    // BIMP_FLAG_GRAPH: the frame as one executable graph, rebuilt when its frame or output
    // pointers change (captured on capStream, a stream of the context's own)
    hipStream_t capStream = nullptr;
    hipGraphExec_t frameExec = nullptr;
    const void* gSrc = nullptr;
    size_t gPitch = 0;
    const void* gOut = nullptr;
    int gCap = -1, gLd = -1, gFlags = -1;
};

namespace bimpi {
using bimpk::kNumOri;
using bimpk::kUniqueOri;

// One frame on `s` with no host synchronisation: pyramid, bank, cells, NMS and compaction
// into `out` (device; keypoint records, or the 7-row matrix when matrix_ld > 0), the total
// count left in c->counter.  Returns a BIMP_* status.
int run_frame(bimp_ctx* c, const uint8_t* d_img, size_t pitch, void* out, int cap,
              int matrix_ld, hipStream_t s);
}  // namespace bimpi
