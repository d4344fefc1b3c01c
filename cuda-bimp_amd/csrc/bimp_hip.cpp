// bimp_hip.cpp -- host context and C ABI of libbimp_hip.so (declared in include/bimp_hip.h).
//
// One context = the reference's CudaStuff (bimp/src/cuda_bimp.cu:222-460): per-scale
// parameters, the quadrature filter bank resident in HBM, and every buffer one frame needs,
// allocated once.  A frame is a fixed sequence of kernel launches on the caller's stream with
// no host synchronisation until the keypoint count is read back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bimp_hip.h"
#include "bank.hpp"
#include "context.hpp"
#include "kernels.hpp"

using namespace bimpi;


namespace {

// Message of the last failed bimp_create on this thread (the context does not exist then).
thread_local std::string g_create_error;

int fail(bimp_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(bimp_ctx* c, hipError_t e, const char* what) {
    return fail(c, BIMP_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, expr)                                    \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
    } while (0)

template <class T>
hipError_t dalloc(T** p, size_t count) {
    return hipMalloc((void**)p, count * sizeof(T) + 256);
}

int round_half_away(double v) { return (int)std::round(v); }

// Profiling brackets around one launch.
struct LaunchScope {
    bimp_ctx* c;
    hipStream_t s;
    ProfEv* e = nullptr;
    LaunchScope(bimp_ctx* c_, hipStream_t s_, int fam, double bytes, int size = 0) : c(c_), s(s_) {
        if (!(c->opt.flags & BIMP_FLAG_PROFILE)) return;
        if (c->evUsed == c->ev.size()) {
            ProfEv n;
            (void)hipEventCreate(&n.a);
            (void)hipEventCreate(&n.b);
            c->ev.push_back(n);
        }
        e = &c->ev[c->evUsed++];
        e->fam = fam;
        e->size = size;
        e->bytes = bytes;
        (void)hipEventRecord(e->a, s);
    }
    ~LaunchScope() {
        if (e) (void)hipEventRecord(e->b, s);
    }
};

void collect_profile(bimp_ctx* c) {
    if (!c->evUsed) return;
    // launches on two streams complete out of record order: wait for every one
    for (size_t i = 0; i < c->evUsed; ++i) (void)hipEventSynchronize(c->ev[i].b);
    for (size_t i = 0; i < c->evUsed; ++i) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, c->ev[i].a, c->ev[i].b) == hipSuccess) {
            std::string name = kFamilyNames[c->ev[i].fam];
            if (c->ev[i].size) name += "<" + std::to_string(c->ev[i].size) + ">";
            ProfAcc& a = c->prof[name];
            a.ms += ms;
            a.bytes += c->ev[i].bytes;
            a.launches += 1;
        }
    }
    c->evUsed = 0;
}

void free_all(bimp_ctx* c) {
    for (auto& s : c->sc) {
        (void)hipFree(s.G);
        (void)hipFree(s.DT);
        (void)hipFree(s.C);
        (void)hipFree(s.D);
        (void)hipFree(s.S);
        (void)hipFree(s.frag);
    }
    for (auto* p : c->level) (void)hipFree(p);
    for (auto& st : c->steps) (void)hipFree(st.jobs);
    for (auto& kv : c->tw) (void)hipFree(kv.second);
    (void)hipFree(c->R);
    (void)hipFree(c->T);
    (void)hipFree(c->R2);
    (void)hipFree(c->T2);
    if (c->frameExec) (void)hipGraphExecDestroy(c->frameExec);
    if (c->capStream) (void)hipStreamDestroy(c->capStream);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->evFork) (void)hipEventDestroy(c->evFork);
    if (c->evJoin) (void)hipEventDestroy(c->evJoin);
    if (c->evCells) (void)hipEventDestroy(c->evCells);
    (void)hipFree(c->maxslots);
    (void)hipFree(c->counter);
    (void)hipFree(c->blockCounts);
    (void)hipFree(c->blockMasks);
    (void)hipFree(c->nmsTable);
    (void)hipFree(c->kpBuf);
    if (c->hostIn) (void)hipHostFree(c->hostIn);
    if (c->hostInDone) (void)hipEventDestroy(c->hostInDone);
    for (auto& e : c->ev) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
}

// End-stop/inhibition of scales first .. first + n - 1, which are ready together: one launch
// per (up to) bimpk::kCellMaxScales of them.
int launch_cells_group(bimp_ctx* c, size_t first, int n, hipStream_t s) {
    for (int k0 = 0; k0 < n; k0 += bimpk::kCellMaxScales) {
        const int m = std::min(n - k0, bimpk::kCellMaxScales);
        bimpk::CellScale cs[bimpk::kCellMaxScales];
        double bytes = 0;
        for (int k = 0; k < m; ++k) {
            const size_t i = first + (size_t)(k0 + k);
            auto& sc = c->sc[i];
            const auto& g = sc.g;
            cs[k] = bimpk::CellScale{sc.C, sc.Hc, sc.ldc, g.FH, g.FW, g.H, g.W, sc.cp, sc.D, sc.S,
                                     c->maxslots + i};
            bytes += 4.0 * kNumOri * (g.H + 1.0) * (g.W + 1.0) + 8.0 * g.H * g.W;
        }
        LaunchScope ls(c, s, F_CELLS, bytes);
        HIPCHK(c, bimpk::launch_cells(cs, m, s));
    }
    return BIMP_OK;
}

// Budget for the inverse-pass intermediate T: scales of one group share an inverse launch
// while their T stays within it, so that T is consumed from the 256 MB last-level cache.
constexpr size_t kTBudgetBytes = 160u << 20;

// Direct-form scales per launch: consecutive direct scales of every level share one launch,
// so the lighter ones fill the tail of the heavier ones.
constexpr int kDirectLaunchScales = 8;

// Pyramid levels a context can hold (a frame side would have to exceed 2^31 to need more).
constexpr int kMaxLevels = 32;

int inverse_batch(size_t tstride_elems) {
    const size_t per = tstride_elems * sizeof(float2);
    return (int)std::max<size_t>(1, std::min<size_t>(bimpk::kMaxBatch, kTBudgetBytes / per));
}

// Pyramid, forward transforms (one launch per stage and resolution group), then per
// sub-batch of scales the inverse column and row passes and, with `cells`, the
// end-stop/inhibition stencil.  Direct-form scales replace the transforms by one MFMA
// convolution launch.  Running the consumers right after their producers keeps the
// intermediate T and the complex-cell maps resident in the 256 MB last-level cache.
// A host frame into the context's level-0 buffer on s: contiguous frames with one copy, pitched
// ones gathered into pinned staging first (a 2D copy from pageable memory is slow).
int upload_frame(bimp_ctx* c, const uint8_t* img, size_t pitch, hipStream_t s) {
    const int W0 = c->width, H0 = c->height;
    const size_t bytes = (size_t)W0 * H0;
    if (pitch == (size_t)W0) {
        HIPCHK(c, hipMemcpyAsync(c->level[0], img, bytes, hipMemcpyHostToDevice, s));
        return BIMP_OK;
    }
    if (!c->hostIn) {
        HIPCHK(c, hipHostMalloc((void**)&c->hostIn, bytes, hipHostMallocDefault));
        HIPCHK(c, hipEventCreateWithFlags(&c->hostInDone, hipEventDisableTiming));
    } else {
        HIPCHK(c, hipEventSynchronize(c->hostInDone));
    }
    for (int y = 0; y < H0; ++y)
        std::memcpy(c->hostIn + (size_t)y * W0, img + (size_t)y * pitch, (size_t)W0);
    HIPCHK(c, hipMemcpyAsync(c->level[0], c->hostIn, bytes, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipEventRecord(c->hostInDone, s));
    return BIMP_OK;
}

// Parts of a frame for graph capture (run_filter's `part`): the whole frame as issued eagerly,
// or, captured one per graph: the per-frame reset, the work on levels above 0 (pyramid
// included), the level-0 work, or both of those in order on one stream.
enum FramePart { kPartAll = 0, kPartHead, kPartUpper, kPartLevel0, kPartBody };

int run_filter(bimp_ctx* c, const uint8_t* img, size_t pitch, int on_device, bool cells,
               hipStream_t s, int part = kPartAll) {
    const int W0 = c->width, H0 = c->height;
    if (part == kPartAll || part == kPartHead) {
        // the cells stage takes maxima into the per-scale slots, which must start at zero:
        // the keypoint pass clears them after reading them, so only a frame after work that
        // left them holding maxima needs the reset (a graph's head always takes it)
        if (part == kPartHead || !c->slotsClean) {
            LaunchScope ls(c, s, F_INIT, 0);
            HIPCHK(c, bimpk::launch_init(c->maxslots, (int)c->sc.size(), c->counter, s));
        }
        // the cells stages below write the slots: dirty from their first launch on, so that a
        // call failing part way leaves the next frame to reset them
        c->slotsClean = !cells;
        c->maxValid = false;
        if (part == kPartHead) return BIMP_OK;
    }
    const uint8_t* src0 = img;
    size_t pitch0 = pitch;
    if (!on_device && part == kPartAll) {
        LaunchScope ls(c, s, F_COPY_IN, 2.0 * W0 * H0);
        const int st = upload_frame(c, img, pitch, s);
        if (st) return st;
        src0 = c->level[0];
        pitch0 = W0;
    }
    // Work above pyramid level 0 (the pyramid itself, FFT groups and direct launches) goes to
    // the side stream, forked once the frame is on the device, except while s is being
    // captured into a graph (there it stays on s).  Level-0 work reads the frame only.
    hipStreamCaptureStatus capturing = hipStreamCaptureStatusActive;
    const bool useSide = part == kPartAll && c->side &&
                         !(c->opt.flags & BIMP_FLAG_SINGLE_STREAM) &&
                         hipStreamIsCapturing(s, &capturing) == hipSuccess &&
                         capturing == hipStreamCaptureStatusNone;
    if (useSide) {
        HIPCHK(c, hipEventRecord(c->evFork, s));
        HIPCHK(c, hipStreamWaitEvent(c->side, c->evFork, 0));
    }
    const hipStream_t pq = useSide ? c->side : s;
    const uint8_t* lv[kMaxLevels] = {};
    size_t lp[kMaxLevels] = {};
    lv[0] = src0;
    lp[0] = pitch0;
    for (int l = 1; l <= c->maxLevel; ++l) {
        if (part == kPartLevel0) {  // the pyramid belongs to the upper-level part
            lv[l] = c->level[l];
            lp[l] = (size_t)c->levelW[l];
            continue;
        }
        LaunchScope ls(c, pq, F_PYR,
                       (double)c->levelW[l - 1] * c->levelH[l - 1] + (double)c->levelW[l] * c->levelH[l]);
        HIPCHK(c, bimpk::launch_pyr_down(lv[l - 1], c->levelW[l - 1], c->levelH[l - 1], lp[l - 1],
                                         c->level[l], c->levelW[l], c->levelH[l],
                                         (size_t)c->levelW[l],
                                         (c->opt.flags & BIMP_FLAG_RESIZE_CPU) != 0, pq));
        lv[l] = c->level[l];
        lp[l] = (size_t)c->levelW[l];
    }
    // the last step at pyramid level 0: the frame's final transforms on s
    const Step* lastL0 = nullptr;
    for (const Step& st : c->steps)
        if ((st.direct ? c->sc[st.index].g.level : c->sc[c->groups[st.index].first].g.level) == 0)
            lastL0 = &st;
    auto run_step = [&](const Step& st) -> int {
        if (st.direct) {
            const Scale& s0 = c->sc[st.index];
            const auto& g = s0.g;
            // with a side stream a direct step holds one level class (bimp_create)
            const hipStream_t q = useSide && g.level > 0 ? c->side : s;
            {
                double bytes = (double)g.H * g.W;
                int maxext = 0;
                for (int k = 0; k < st.n; ++k) {
                    const Scale& sk = c->sc[st.index + k];
                    bytes += 4.0 * kNumOri * sk.Hc * sk.Wc;
                    maxext = std::max(maxext, sk.ds.ext);
                }
                LaunchScope ls(c, q, F_DIRECT, bytes, maxext);
                bimpk::DirectLevels dl{};
                for (int l = 0; l <= c->maxLevel && l < bimpk::kDirMaxLevels; ++l) {
                    dl.img[l] = lv[l];
                    dl.pitch[l] = lp[l];
                }
                HIPCHK(c, bimpk::launch_direct(dl, st.jobs, st.n, st.tiles,
                                               st.lds, s0.precision == BIMP_PRECISION_FP16, q));
            }
            if (cells) {
                // the launch's cells stage follows it on the same stream: handing half of the
                // level-0 scales to the second stream adds an event on the way to the keypoint
                // pass and measured slower (BF16 8 sigma 2014 against 2039 frames/s)
                const int r = launch_cells_group(c, (size_t)st.index, st.n, q);
                if (r) return r;
            }
            return BIMP_OK;
        }
        const Group& gr = c->groups[st.index];
        const auto& g = c->sc[gr.first].g;
        // groups above level 0 own R2/T2 whenever the context has a side stream, on s as well
        const bool upper = c->side && g.level > 0;
        const hipStream_t q = upper && useSide ? c->side : s;
        float2* const R = upper ? c->R2 : c->R;
        float2* const T = upper ? c->T2 : c->T;
        const size_t tCap = upper ? c->tCap2 : c->tCap;
        const int ncols = g.FW / 2 + 1;
        const float2* twW = c->tw[g.FW];
        const float2* twH = c->tw[g.FH];
        bimpk::ScaleSet ss;
        ss.n = gr.n;
        for (int k = 0; k < gr.n; ++k) {
            const Scale& sk = c->sc[gr.first + k];
            ss.kX[k] = sk.g.kX;
            ss.G[k] = sk.G;
            ss.DT[k] = sk.DT;
            ss.C[k] = sk.C;
        }
        const size_t rstride = (size_t)g.H * ncols;
        const double n = gr.n;
        {
            LaunchScope ls(c, q, F_FWD_ROWS, n * ((double)g.H * g.W + 8.0 * g.H * ncols), g.FW);
            HIPCHK(c, bimpk::launch_fwd_rows(g.FW, lv[g.level], lp[g.level], g.H, g.W, ss,
                                             c->opt.input_scale, R, ncols, rstride, twW, q));
        }
        {
            LaunchScope ls(c, q, F_FWD_COLS, n * (8.0 * g.H * ncols + 8.0 * ncols * g.FH), g.FH);
            HIPCHK(c, bimpk::launch_fwd_cols(g.FH, R, ncols, rstride, g.H, ss, ncols, twH, q));
        }
        // inverse passes over sub-batches of scales whose intermediate T fits the budget
        const Scale& s0 = c->sc[gr.first];
        const size_t tstride = (size_t)kNumOri * s0.Hc * g.FW;
        const int nb = std::max(1, std::min(c->invBatch > 0 ? c->invBatch : inverse_batch(tstride),
                                            (int)(tCap / tstride)));
        for (int k0 = 0; k0 < gr.n; k0 += nb) {
            bimpk::ScaleSet sub;
            sub.n = std::min(nb, gr.n - k0);
            for (int k = 0; k < sub.n; ++k) {
                const Scale& sk = c->sc[gr.first + k0 + k];
                sub.kX[k] = sk.g.kX;
                sub.G[k] = sk.G;
                sub.DT[k] = sk.DT;
                sub.C[k] = sk.C;
            }
            const double m = sub.n;
            {
                LaunchScope ls(c, q, F_INV_COLS,
                               m * (8.0 * ncols * g.FH + 4.0 * kUniqueOri * g.FW * g.FH +
                                    8.0 * kNumOri * s0.Hc * g.FW),
                               g.FH);
                HIPCHK(c, bimpk::launch_inv_cols(g.FH, sub, g.FW, s0.Hc, T, tstride, twH, q));
            }
            {
                LaunchScope ls(c, q, F_INV_ROWS,
                               m * (8.0 * kNumOri * s0.Hc * g.FW + 4.0 * kNumOri * s0.Hc * s0.Wc),
                               g.FW);
                HIPCHK(c, bimpk::launch_inv_rows(g.FW, g.FH, sub, T, tstride, s0.Hc, s0.Wc, s0.ldc,
                                                 twW, q));
            }
            if (cells) {
                // level 0 with a second stream: the cells stage of these scales runs there,
                // beside the next transforms on s; the frame's last level-0 sub-batch keeps
                // it on s, where the keypoint pass follows: a hand-off there would only add
                // the event's gap to the frame's critical path (1080p 1208 -> 1229 frames/s)
                hipStream_t cq = q;
                const bool last = &st == lastL0 && k0 + nb >= gr.n;
                if (useSide && !upper && !last) {
                    HIPCHK(c, hipEventRecord(c->evCells, q));
                    HIPCHK(c, hipStreamWaitEvent(c->side, c->evCells, 0));
                    cq = c->side;
                }
                const int st = launch_cells_group(c, (size_t)(gr.first + k0), sub.n, cq);
                if (st) return st;
            }
        }
        return BIMP_OK;
    };
    // Second-stream steps are issued first, so that the level-0 cells handed to that stream
    // below queue behind its own transforms rather than ahead of them.
    auto on_side = [&](const Step& st) {
        return (st.direct ? c->sc[st.index].g.level : c->sc[c->groups[st.index].first].g.level) > 0;
    };
    for (int pass = useSide ? 0 : 1; pass < 2; ++pass) {
        for (const Step& st : c->steps) {
            if (useSide && on_side(st) != (pass == 0)) continue;
            if ((part == kPartUpper && !on_side(st)) || (part == kPartLevel0 && on_side(st)))
                continue;
            const int r = run_step(st);
            if (r) return r;
        }
    }
    if (useSide) {
        HIPCHK(c, hipEventRecord(c->evJoin, c->side));
        HIPCHK(c, hipStreamWaitEvent(s, c->evJoin, 0));
    }
    if (cells) {
        c->maxValid = true;
        c->slotsClean = false;
    }
    return BIMP_OK;
}

int run_cells(bimp_ctx* c, hipStream_t s) {
    if (!c->slotsClean) {
        LaunchScope ls(c, s, F_INIT, 0);
        HIPCHK(c, bimpk::launch_init(c->maxslots, (int)c->sc.size(), c->counter, s));
    }
    c->slotsClean = false;
    const int st = launch_cells_group(c, 0, (int)c->sc.size(), s);
    if (st) return st;
    c->maxValid = true;
    return BIMP_OK;
}

int run_keypoints(bimp_ctx* c, void* out, int cap, int matrix_ld, hipStream_t s) {
    if (!c->maxValid) {
        {
            LaunchScope ls(c, s, F_INIT, 0);
            HIPCHK(c, bimpk::launch_init(c->maxslots, (int)c->sc.size(), c->counter, s));
        }
        c->slotsClean = false;
        for (size_t i = 0; i < c->sc.size(); ++i) {
            auto& sc = c->sc[i];
            LaunchScope ls(c, s, F_MAX, 4.0 * sc.g.H * sc.g.W);
            HIPCHK(c, bimpk::launch_max(sc.D, sc.g.H * sc.g.W, c->maxslots + i, s));
        }
        c->maxValid = true;
    }
    double dbytes = 0;
    for (auto& sc : c->sc) dbytes += 4.0 * sc.g.H * sc.g.W;
    {
        LaunchScope ls(c, s, F_NMS_COUNT, dbytes);
        HIPCHK(c, bimpk::launch_nms_count(c->nmsTable, (int)c->sc.size(), c->totalBlocks,
                                          c->maxslots, c->opt.nms_threshold, c->blockCounts,
                                          c->blockMasks, s));
    }
    {
        LaunchScope ls(c, s, F_NMS_WRITE, 4.0 * bimpk::kNmsThreads * c->totalBlocks);
        HIPCHK(c, bimpk::launch_nms_write(c->nmsTable, (int)c->sc.size(), c->totalBlocks,
                                          c->blockCounts, c->counter, c->maxslots, c->blockMasks,
                                          out, cap, matrix_ld, s));
    }
    // the write pass cleared the slots: a later keypoint pass recomputes the maxima
    c->maxValid = false;
    c->slotsClean = true;
    return BIMP_OK;
}

int ensure_kp_buf(bimp_ctx* c, size_t bytes) {
    if (c->kpBufBytes >= bytes) return BIMP_OK;
    (void)hipFree(c->kpBuf);
    c->kpBuf = nullptr;
    c->kpBufBytes = 0;
    HIPCHK(c, hipMalloc(&c->kpBuf, bytes));
    c->kpBufBytes = bytes;
    return BIMP_OK;
}

// Runs NMS into `out` (host or device) and reads the count back.
// With `frame`, the whole frame (filtering and NMS, eagerly or as the frame's graph) is issued
// into the keypoint target instead of the NMS alone.
struct FrameIn {
    const uint8_t* img;
    size_t pitch;
    int on_device;
};
int run_frame_any(bimp_ctx* c, const uint8_t* img, size_t pitch, int on_device, void* out,
                  int cap, int ld, hipStream_t s);

int keypoints_out(bimp_ctx* c, void* out, int cap, int matrix_ld, int out_on_device, int* count,
                  hipStream_t s, void* staging = nullptr, const FrameIn* frame = nullptr) {
    const size_t elem = matrix_ld > 0 ? 0 : sizeof(bimp_keypoint);
    void* target = out;
    if (!out_on_device && cap > 0 && staging) {
        target = staging;
    } else if (!out_on_device && cap > 0) {
        const size_t bytes = matrix_ld > 0 ? (size_t)BIMP_KP_VECTOR_LENGTH * matrix_ld * 4
                                           : (size_t)cap * elem;
        int st = ensure_kp_buf(c, bytes);
        if (st) return st;
        target = c->kpBuf;
    }
    int st = frame ? run_frame_any(c, frame->img, frame->pitch, frame->on_device, target, cap,
                                   matrix_ld, s)
                   : run_keypoints(c, target, cap, matrix_ld, s);
    if (st) return st;
    int total = 0;
    HIPCHK(c, hipMemcpyAsync(&total, c->counter, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    const int n = std::min(total, cap);
    if (!out_on_device && n > 0) {
        if (matrix_ld > 0) {
            HIPCHK(c, hipMemcpy2D(out, (size_t)matrix_ld * 4, target, (size_t)matrix_ld * 4,
                                  (size_t)n * 4, BIMP_KP_VECTOR_LENGTH, hipMemcpyDeviceToHost));
        } else {
            HIPCHK(c, hipMemcpy(out, target, (size_t)n * elem, hipMemcpyDeviceToHost));
        }
    }
    if (count) *count = total;
    if (c->opt.flags & BIMP_FLAG_PROFILE) collect_profile(c);
    if (total > cap) {
        c->err = "keypoints (" + std::to_string(total) + ") exceed the output capacity (" +
                 std::to_string(cap) + ")";
        return BIMP_W_OVERFLOW;
    }
    return BIMP_OK;
}

// ---- per-frame graph (BIMP_FLAG_GRAPH)
// The frame is captured as separate single-stream graphs, composed into one graph whose
// dependencies say what the second stream's events say in the eager form: the per-frame reset
// first, then the work above level 0 (pyramid included) and the level-0 work as two
// independent branches, then the keypoint pass.  Host frames are uploaded to the context's
// level-0 buffer before the graph is launched; the graph reads that buffer.
int capture_part(bimp_ctx* c, hipGraph_t* g, int (*body)(bimp_ctx*, hipStream_t, void*),
                 void* arg) {
    HIPCHK(c, hipStreamBeginCapture(c->capStream, hipStreamCaptureModeThreadLocal));
    const int st = body(c, c->capStream, arg);
    hipGraph_t out = nullptr;
    const hipError_t e = hipStreamEndCapture(c->capStream, &out);
    if (st != BIMP_OK) {
        if (out) (void)hipGraphDestroy(out);
        return st;
    }
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    *g = out;
    return BIMP_OK;
}

struct GraphArgs {
    const uint8_t* src;
    size_t pitch;
    void* out;
    int cap, ld, part;
};

int build_frame_graph(bimp_ctx* c, const uint8_t* src, size_t pitch, void* out, int cap, int ld) {
    if (c->frameExec) {
        (void)hipGraphExecDestroy(c->frameExec);
        c->frameExec = nullptr;
    }
    if (!c->capStream) HIPCHK(c, hipStreamCreateWithFlags(&c->capStream, hipStreamNonBlocking));
    GraphArgs ga{src, pitch, out, cap, ld, kPartHead};
    auto filt = [](bimp_ctx* x, hipStream_t q, void* a) {
        const GraphArgs* g = static_cast<const GraphArgs*>(a);
        return run_filter(x, g->src, g->pitch, 1, true, q, g->part);
    };
    auto kp = [](bimp_ctx* x, hipStream_t q, void* a) {
        const GraphArgs* g = static_cast<const GraphArgs*>(a);
        return run_keypoints(x, g->out, g->cap, g->ld, q);
    };
    const bool split = c->side && !(c->opt.flags & BIMP_FLAG_SINGLE_STREAM);
    hipGraph_t parts[4] = {nullptr, nullptr, nullptr, nullptr};  // head, upper, level 0, tail
    auto release = [&]() {
        for (auto& g : parts)
            if (g) (void)hipGraphDestroy(g);
    };
    int st = capture_part(c, &parts[0], filt, &ga);
    if (!st && split) {
        ga.part = kPartUpper;
        st = capture_part(c, &parts[1], filt, &ga);
        ga.part = kPartLevel0;
        if (!st) st = capture_part(c, &parts[2], filt, &ga);
    } else if (!st) {
        ga.part = kPartBody;
        st = capture_part(c, &parts[2], filt, &ga);
    }
    if (!st) st = capture_part(c, &parts[3], kp, &ga);
    if (st) {
        release();
        return st;
    }
    hipGraph_t g = nullptr;
    hipGraphNode_t head = nullptr, upper = nullptr, l0 = nullptr, tail = nullptr;
    hipError_t e = hipGraphCreate(&g, 0);
    if (e == hipSuccess) e = hipGraphAddChildGraphNode(&head, g, nullptr, 0, parts[0]);
    if (e == hipSuccess && parts[1]) e = hipGraphAddChildGraphNode(&upper, g, &head, 1, parts[1]);
    if (e == hipSuccess) e = hipGraphAddChildGraphNode(&l0, g, &head, 1, parts[2]);
    if (e == hipSuccess) {
        hipGraphNode_t deps[2] = {l0, upper};
        e = hipGraphAddChildGraphNode(&tail, g, deps, upper ? 2 : 1, parts[3]);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&c->frameExec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    release();
    if (e != hipSuccess) {
        c->frameExec = nullptr;
        return hip_fail(c, e, "frame graph");
    }
    c->gSrc = src;
    c->gPitch = pitch;
    c->gOut = out;
    c->gCap = cap;
    c->gLd = ld;
    c->gFlags = c->opt.flags;
    return BIMP_OK;
}

// One frame into `out` (device) on `s`: eager launches, or with BIMP_FLAG_GRAPH the frame's
// graph (built on first use and whenever the frame or output pointers change).
int run_frame_any(bimp_ctx* c, const uint8_t* img, size_t pitch, int on_device, void* out,
                  int cap, int ld, hipStream_t s) {
    if (!(c->opt.flags & BIMP_FLAG_GRAPH) || (c->opt.flags & BIMP_FLAG_PROFILE)) {
        int st = run_filter(c, img, pitch, on_device, true, s);
        return st ? st : run_keypoints(c, out, cap, ld, s);
    }
    const uint8_t* src = img;
    size_t p = pitch;
    if (!on_device) {
        // the graph reads the level-0 buffer; the upload is ordered before it on s
        const int st = upload_frame(c, img, pitch, s);
        if (st) return st;
        src = c->level[0];
        p = (size_t)c->width;
    }
    if (!c->frameExec || c->gSrc != src || c->gPitch != p || c->gOut != out || c->gCap != cap ||
        c->gLd != ld || c->gFlags != c->opt.flags) {
        // a capture runs nothing: the slots keep the state they had before it
        const bool clean = c->slotsClean, valid = c->maxValid;
        const int st = build_frame_graph(c, src, p, out, cap, ld);
        c->slotsClean = clean;
        c->maxValid = valid;
        if (st) return st;
    }
    HIPCHK(c, hipGraphLaunch(c->frameExec, s));
    // the graph ends with the keypoint pass, which clears the max slots
    c->maxValid = false;
    c->slotsClean = true;
    return BIMP_OK;
}

}  // namespace

// ---- keypoint staging of bimp_detect_u8_noalloc: no device allocation per call
namespace {
constexpr int kNoAllocMaxKeypoints = 65536;  // 65536 records x 28 B = 1.75 MiB per device
}  // namespace

// One copy per device, held until the process ends.  Every bimp_detect_u8_noalloc call on a
// device shares it: calls that overlap on one device race.  Nothing in it needs clearing: a
// call reads back only the records its own keypoint pass wrote.
static __device__ __align__(256) bimp_keypoint scratch[kNoAllocMaxKeypoints];

namespace {
std::mutex g_scratch_mu;
std::vector<bimp_keypoint*> g_scratch;  // per device: the array's device address, or null

// Device address of the current device's copy of `scratch`, looked up once per device with
// that device current.  In host code the array's name is not a device address.
hipError_t noalloc_staging(bimp_keypoint** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch.size() <= dev) g_scratch.resize((size_t)dev + 1, nullptr);
    if (!g_scratch[(size_t)dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[(size_t)dev] = static_cast<bimp_keypoint*>(p);
    }
    *out = g_scratch[(size_t)dev];
    return hipSuccess;
}
}  // namespace

int bimpi::run_frame(bimp_ctx* c, const uint8_t* d_img, size_t pitch, void* out, int cap,
                     int matrix_ld, hipStream_t s) {
    const int st = run_filter(c, d_img, pitch, 1, true, s);
    if (st) return st;
    return run_keypoints(c, out, cap, matrix_ld, s);
}

// ======================================================================= C ABI
extern "C" {

void bimp_default_options(bimp_options* o) {
    if (!o) return;
    o->gamma = 0.5f;
    o->sigma_ratio = 0.56f;
    o->support = 7.0f;
    o->pyr_threshold = 7.5f;
    o->nms_threshold = 0.104;
    o->input_scale = (float)(255.0 / 65536.0);
    o->flags = 0;
    o->precision = BIMP_PRECISION_FP32;
    o->direct_max_lambda = 64.0f;  // every scale of the reference's banks (4 .. 64)
    o->direct_sigmas = 8.f;  // 8 sigma: inside the 1e-2 map / 99% keypoint bars, 6 sigma is not
}

int bimp_snap_transform_size(int n) { return n > 0 ? bimph::snap_transform_size(n) : BIMP_E_ARG; }

int bimp_make_lambdas_log(float start, float end, int spo, float* out, int cap) {
    if (spo <= 0 || !(start > 0.f) || cap < 0) return BIMP_E_ARG;
    const float multiplier = (float)std::pow(2., 1. / spo);
    int n = 0;
    for (float lambda = start; (double)lambda <= (double)end * 1.05; lambda *= multiplier) {
        if (out && n < cap) out[n] = lambda;
        if (++n > 1000000) break;
    }
    return n;
}

int bimp_create(int width, int height, const float* lambdas, int n, const bimp_options* opt,
                bimp_ctx** out) {
    if (!out) return BIMP_E_ARG;
    *out = nullptr;
    if (width < 4 || height < 4 || !lambdas || n <= 0 || n > 64) {
        g_create_error = "invalid frame size or wavelength list";
        return BIMP_E_ARG;
    }
    bimp_ctx* c = new bimp_ctx;
    c->width = width;
    c->height = height;
    if (opt)
        c->opt = *opt;
    else
        bimp_default_options(&c->opt);
    if (c->opt.precision < BIMP_PRECISION_FP32 || c->opt.precision > BIMP_PRECISION_FP16 ||
        !(c->opt.direct_sigmas >= 0.f)) {
        delete c;
        g_create_error = "invalid precision options";
        return BIMP_E_ARG;
    }
    std::vector<float> lam(lambdas, lambdas + n);
    std::sort(lam.begin(), lam.end());  // cuda_bimp.cu:224
    for (float l : lam) {
        Scale s;
        if (!bimph::derive_scale(l, width, height, c->opt.pyr_threshold, c->opt.support, &s.g) ||
            s.g.lambda < 1.0f || s.g.H < 4 || s.g.W < 4 || s.g.kW < 1) {
            delete c;
            g_create_error = "wavelength " + std::to_string(l) + " gives an invalid scale";
            return BIMP_E_ARG;
        }
        if (c->opt.precision != BIMP_PRECISION_FP32 && l <= c->opt.direct_max_lambda &&
            s.g.kW <= BIMP_DIRECT_MAX_KERNEL) {
            bimpk::DirectScale d{};
            d.kpad = s.g.kX;
            d.kW = s.g.kW;
            d.ext = bimph::direct_extent(s.g, c->opt.sigma_ratio, c->opt.direct_sigmas);
            if (bimpk::direct_geometry(d) <= bimpk::kMaxLds && d.lwp <= bimpk::kDirMaxLwp)
                s.precision = c->opt.precision;
        }
        if (s.precision == BIMP_PRECISION_FP32 &&
            (!bimpk::fft_size_supported(s.g.FH) || !bimpk::fft_size_supported(s.g.FW) ||
             s.g.FW % bimpk::col_batch(s.g.FH) != 0)) {
            delete c;
            g_create_error = "transform size " + std::to_string(s.g.FH) + " x " +
                             std::to_string(s.g.FW) + " is not supported";
            return BIMP_E_UNSUPPORTED;
        }
        const auto& g = s.g;
        s.Wc = std::min(g.FW, g.W + 1);
        s.Hc = std::min(g.FH, g.H + (g.W == g.FW ? 2 : 1));
        s.ldc = (s.Wc + 31) / 32 * 32;
        for (int o = 0; o < kNumOri; ++o) {
            const double theta = (double)(float)(o * 3.1415926535897932384626433832795 / kNumOri);
            s.cp.ea[o] = round_half_away(0.6 * g.lambda * std::cos(theta));
            s.cp.eb[o] = round_half_away(0.6 * g.lambda * std::sin(theta));
            s.cp.ia[o] = round_half_away(0.6 * g.lambda * std::sin(theta));
            s.cp.ib[o] = round_half_away(0.6 * g.lambda * std::cos(theta));
            s.cp.ra[o] = round_half_away(0.5 * s.cp.ia[o]);
            s.cp.rb[o] = round_half_away(0.5 * s.cp.ib[o]);
        }
        c->sc.push_back(s);
    }
    int st = BIMP_OK;
    auto bail = [&](int code, const std::string& msg) {
        free_all(c);
        delete c;
        g_create_error = msg;
        return code;
    };
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return bail(BIMP_E_DEVICE, "no HIP device");
    // pyramid levels
    for (auto& s : c->sc) c->maxLevel = std::max(c->maxLevel, s.g.level);
    if (c->maxLevel >= kMaxLevels) return bail(BIMP_E_ARG, "too many pyramid levels");
    c->levelW.resize(c->maxLevel + 1);
    c->levelH.resize(c->maxLevel + 1);
    c->level.assign(c->maxLevel + 1, nullptr);
    c->levelW[0] = width;
    c->levelH[0] = height;
    for (int l = 1; l <= c->maxLevel; ++l) {
        c->levelW[l] = (c->levelW[l - 1] + 1) / 2;
        c->levelH[l] = (c->levelH[l - 1] + 1) / 2;
    }
    for (int l = 0; l <= c->maxLevel; ++l)
        if (dalloc(&c->level[l], (size_t)c->levelW[l] * c->levelH[l]) != hipSuccess)
            return bail(BIMP_E_ALLOC, "level");
    // resolution groups: consecutive FFT-form scales (ascending lambda) with one level and
    // size; a direct-form scale is a step of its own
    // A second stream when the frame's work spans pyramid level 0 and higher levels (see
    // run_filter); then direct launches do not mix levels either.
    bool spans[2] = {false, false};
    for (const auto& s : c->sc) spans[s.g.level > 0 ? 1 : 0] = true;
    const bool sideStream = spans[0] && spans[1] && !(c->opt.flags & BIMP_FLAG_SINGLE_STREAM);
    bool prevFft = false;
    for (int i = 0; i < (int)c->sc.size(); ++i) {
        const auto& g = c->sc[i].g;
        if (c->sc[i].precision != BIMP_PRECISION_FP32) {
            // consecutive direct scales share a launch
            if (!c->steps.empty() && c->steps.back().direct && !prevFft) {
                Step& last = c->steps.back();
                const bool sameSide = (c->sc[last.index].g.level > 0) == (g.level > 0);
                if (last.n < kDirectLaunchScales && (!sideStream || sameSide)) {
                    ++last.n;
                    continue;
                }
            }
            Step st;
            st.direct = true;
            st.index = i;
            c->steps.push_back(st);
            prevFft = false;
            continue;
        }
        if (prevFft) {
            Group& last = c->groups.back();
            const auto& h0 = c->sc[last.first].g;
            if (h0.level == g.level && h0.H == g.H && h0.W == g.W && h0.FH == g.FH &&
                h0.FW == g.FW && last.n < bimpk::kMaxBatch) {
                ++last.n;
                continue;
            }
        }
        c->groups.push_back({i, 1});
        c->steps.push_back({false, (int)c->groups.size() - 1});
        prevFft = true;
    }
    size_t rMax[2] = {0, 0}, tMax[2] = {0, 0};
    std::vector<int> sizes;
    const unsigned hw = std::thread::hardware_concurrency();
    const int threads = (int)std::max(1u, std::min(16u, hw ? hw : 1u));
    for (const auto& gr : c->groups) {
        const Scale& s0 = c->sc[gr.first];
        const int k = s0.g.level > 0 ? 1 : 0;
        rMax[k] = std::max(rMax[k], (size_t)gr.n * s0.g.H * (s0.g.FW / 2 + 1));
        const size_t ts = (size_t)kNumOri * s0.Hc * s0.g.FW;
        tMax[k] = std::max(tMax[k], ts * std::min(gr.n, inverse_batch(ts)));
    }
    // FFT groups above level 0 own R2/T2 only with a side stream
    if (!sideStream) {
        rMax[0] = std::max(rMax[0], rMax[1]);
        tMax[0] = std::max(tMax[0], tMax[1]);
        rMax[1] = tMax[1] = 0;
    }
    bool anyDirect = false;
    for (auto& s : c->sc) {
        const auto& g = s.g;
        if (dalloc(&s.C, (size_t)kNumOri * s.Hc * s.ldc) != hipSuccess ||
            dalloc(&s.D, (size_t)g.H * g.W) != hipSuccess ||
            dalloc(&s.S, (size_t)g.H * g.W) != hipSuccess)
            return bail(BIMP_E_ALLOC, "scale buffers");
        if (s.precision != BIMP_PRECISION_FP32) {
            anyDirect = true;
            bimph::DirectBank db;
            bimph::direct_bank(g, c->opt.gamma, c->opt.sigma_ratio, c->opt.direct_sigmas,
                               s.precision == BIMP_PRECISION_FP16, bimpk::kDirRows, db);
            auto& d = s.ds;
            d.H = g.H;
            d.W = g.W;
            d.FH = g.FH;
            d.FW = g.FW;
            d.kpad = g.kX;
            d.kW = g.kW;
            d.Hc = s.Hc;
            d.Wc = s.Wc;
            d.ldc = s.ldc;
            d.ext = db.ext;
            d.fscale = db.fscale;
            d.iscale = c->opt.input_scale;
            for (int f = 0; f < 2 * kNumOri; ++f) d.dc[f] = db.dc[f];
            if (bimpk::direct_geometry(d) > bimpk::kMaxLds || d.nvb != db.nvb || d.nrho != db.nrho)
                return bail(BIMP_E_UNSUPPORTED, "direct-form tile exceeds LDS");
            if (dalloc((uint16_t**)&s.frag, db.frag.size()) != hipSuccess)
                return bail(BIMP_E_ALLOC, "direct taps");
            if (hipMemcpy(s.frag, db.frag.data(), db.frag.size() * sizeof(uint16_t),
                          hipMemcpyHostToDevice) != hipSuccess)
                return bail(BIMP_E_DEVICE, "direct taps upload");
        } else {
            const int ncols = g.FW / 2 + 1;
            sizes.push_back(g.FH);
            sizes.push_back(g.FW);
            if (dalloc(&s.G, (size_t)kUniqueOri * g.FW * g.FH) != hipSuccess ||
                dalloc(&s.DT, (size_t)ncols * g.FH) != hipSuccess)
                return bail(BIMP_E_ALLOC, "scale buffers");
            std::vector<float> bank;
            bimph::quadrature_bank(g, c->opt.gamma, c->opt.sigma_ratio, threads, bank);
            if (hipMemcpy(s.G, bank.data(), bank.size() * sizeof(float), hipMemcpyHostToDevice) !=
                hipSuccess)
                return bail(BIMP_E_DEVICE, "bank upload");
        }
        // complex-cell maps start at zero so the unused padding columns are defined
        if (hipMemset(s.C, 0, (size_t)kNumOri * s.Hc * s.ldc * sizeof(float)) != hipSuccess)
            return bail(BIMP_E_DEVICE, "memset");
    }
    c->tCap = tMax[0];
    c->tCap2 = tMax[1];
    if ((rMax[0] && dalloc(&c->R, rMax[0]) != hipSuccess) ||
        (tMax[0] && dalloc(&c->T, tMax[0]) != hipSuccess) ||
        (rMax[1] && dalloc(&c->R2, rMax[1]) != hipSuccess) ||
        (tMax[1] && dalloc(&c->T2, tMax[1]) != hipSuccess))
        return bail(BIMP_E_ALLOC, "scratch");
    // The fork, hand-off and join events order kernels of this device only, so they skip the
    // system-scope fence (the host sees results through the caller's stream, whose kernels
    // keep theirs): less of a gap on the main stream at each record (1080p 1237 -> 1244
    // frames/s, 640x480 5528 -> 5637; two-stream frames equal the one-stream ones bit for bit).
    constexpr unsigned kEvFlags = hipEventDisableTiming | hipEventDisableSystemFence;
    if (sideStream &&
        (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
         hipEventCreateWithFlags(&c->evFork, kEvFlags) != hipSuccess ||
         hipEventCreateWithFlags(&c->evJoin, kEvFlags) != hipSuccess ||
         hipEventCreateWithFlags(&c->evCells, kEvFlags) != hipSuccess))
        return bail(BIMP_E_DEVICE, "side stream");
    if (anyDirect) {
        if (hipError_t e = bimpk::prepare_direct(); e != hipSuccess)
            return bail(BIMP_E_DEVICE, std::string("kernel attributes: ") + hipGetErrorString(e));
        for (Step& st : c->steps) {
            if (!st.direct) continue;
            std::vector<bimpk::DirectScale> ds;
            std::vector<const void*> fr;
            std::vector<float*> cm;
            std::vector<int> lvl;
            for (int k = 0; k < st.n; ++k) {
                const Scale& sk = c->sc[st.index + k];
                ds.push_back(sk.ds);
                fr.push_back(sk.frag);
                cm.push_back(sk.C);
                lvl.push_back(sk.g.level);
            }
            std::vector<bimpk::DirectJob> jobs(st.n);
            st.lds = bimpk::direct_jobs(ds.data(), fr.data(), cm.data(), lvl.data(), st.n,
                                        jobs.data(), &st.tiles);
            if (!st.lds) return bail(BIMP_E_UNSUPPORTED, "direct-form job table");
            if (dalloc(&st.jobs, jobs.size()) != hipSuccess) return bail(BIMP_E_ALLOC, "direct jobs");
            if (hipMemcpy(st.jobs, jobs.data(), jobs.size() * sizeof(bimpk::DirectJob),
                          hipMemcpyHostToDevice) != hipSuccess)
                return bail(BIMP_E_DEVICE, "direct jobs upload");
        }
    }
    if (hipError_t e = bimpk::prepare_common(); e != hipSuccess)
        return bail(BIMP_E_DEVICE, std::string("kernel attributes: ") + hipGetErrorString(e));
    std::sort(sizes.begin(), sizes.end());
    sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
    for (int n2 : sizes) {
        if (hipError_t e = bimpk::prepare_size(n2); e != hipSuccess)
            return bail(BIMP_E_DEVICE, std::string("kernel attributes: ") + hipGetErrorString(e));
        std::vector<float2> t;
        if (!bimpk::twiddle_table(n2, t)) return bail(BIMP_E_UNSUPPORTED, "twiddle table");
        float2* d = nullptr;
        if (dalloc(&d, t.size()) != hipSuccess) return bail(BIMP_E_ALLOC, "twiddles");
        c->tw[n2] = d;
        if (hipMemcpy(d, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess)
            return bail(BIMP_E_DEVICE, "twiddle upload");
    }
    // NMS descriptors
    std::vector<bimpk::NmsScale> tab;
    int blocks = 0;
    for (size_t i = 0; i < c->sc.size(); ++i) {
        const auto& s = c->sc[i];
        bimpk::NmsScale d;
        d.D = s.D;
        d.H = s.g.H;
        d.W = s.g.W;
        d.off = s.g.off;
        d.block0 = blocks;
        d.nblocks = (s.g.H * s.g.W + bimpk::kNmsBlockPixels - 1) / bimpk::kNmsBlockPixels;
        d.lambda = s.g.lambda;
        d.pyrstep = s.g.pyrstep;
        d.slot = (int)i;
        blocks += d.nblocks;
        tab.push_back(d);
    }
    c->totalBlocks = blocks;
    if (dalloc(&c->nmsTable, tab.size()) != hipSuccess || dalloc(&c->blockCounts, blocks + 1) != hipSuccess ||
        dalloc(&c->blockMasks, (size_t)blocks * bimpk::kNmsThreads) != hipSuccess ||
        dalloc(&c->maxslots, c->sc.size()) != hipSuccess || dalloc(&c->counter, 1) != hipSuccess)
        return bail(BIMP_E_ALLOC, "nms buffers");
    if (hipMemset(c->maxslots, 0, c->sc.size() * sizeof(unsigned)) != hipSuccess ||
        hipMemset(c->counter, 0, sizeof(int)) != hipSuccess)
        return bail(BIMP_E_DEVICE, "nms buffers");
    if (hipMemcpy(c->nmsTable, tab.data(), tab.size() * sizeof(bimpk::NmsScale),
                  hipMemcpyHostToDevice) != hipSuccess)
        return bail(BIMP_E_DEVICE, "nms table upload");
    if (noalloc_staging(&c->noallocStaging) != hipSuccess)
        return bail(BIMP_E_DEVICE, "no-allocation staging lookup");
    (void)st;
    *out = c;
    return BIMP_OK;
}

int bimp_destroy(bimp_ctx* c) {
    if (!c) return BIMP_E_ARG;
    (void)hipDeviceSynchronize();
    free_all(c);
    delete c;
    return BIMP_OK;
}

int bimp_num_scales(const bimp_ctx* c) { return c ? (int)c->sc.size() : BIMP_E_ARG; }

int bimp_get_scale_info(const bimp_ctx* c, int i, bimp_scale_info* o) {
    if (!c || !o || i < 0 || i >= (int)c->sc.size()) return BIMP_E_ARG;
    const auto& s = c->sc[i];
    o->lambda2 = s.g.lambda2;
    o->lambda = s.g.lambda;
    o->pyrstep = s.g.pyrstep;
    o->kernel_size = s.g.kW;
    o->kernel_centre = s.g.kX;
    o->border = s.g.off;
    o->level = s.g.level;
    o->data_h = s.g.H;
    o->data_w = s.g.W;
    o->fft_h = s.g.FH;
    o->fft_w = s.g.FW;
    o->map_h = s.Hc;
    o->map_w = s.Wc;
    o->map_pitch = s.ldc;
    o->precision = s.precision;
    o->tap_extent = s.precision != BIMP_PRECISION_FP32 ? s.ds.ext : s.g.kW;
    return BIMP_OK;
}

int bimp_filter_u8(bimp_ctx* c, const uint8_t* img, size_t pitch, int on_device, bimp_stream_t stream) {
    if (!c || !img || pitch < (size_t)c->width) return fail(c, BIMP_E_ARG, "bad image arguments");
    hipStream_t s = (hipStream_t)stream;
    c->lastStream = s;
    return run_filter(c, img, pitch, on_device, false, s);
}

int bimp_cells(bimp_ctx* c, bimp_stream_t stream) {
    if (!c) return BIMP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    c->lastStream = s;
    return run_cells(c, s);
}

int bimp_keypoints(bimp_ctx* c, bimp_keypoint* out, int cap, int out_on_device, int* count,
                   bimp_stream_t stream) {
    if (!c || cap < 0 || (cap > 0 && !out)) return fail(c, BIMP_E_ARG, "bad output arguments");
    hipStream_t s = (hipStream_t)stream;
    c->lastStream = s;
    return keypoints_out(c, out, cap, 0, out_on_device, count, s);
}

int bimp_detect_u8(bimp_ctx* c, const uint8_t* img, size_t pitch, int img_on_device,
                   bimp_keypoint* out, int cap, int out_on_device, int* count, bimp_stream_t stream) {
    if (!c || !img || pitch < (size_t)c->width || cap < 0 || (cap > 0 && !out))
        return fail(c, BIMP_E_ARG, "bad arguments");
    hipStream_t s = (hipStream_t)stream;
    c->lastStream = s;
    const FrameIn fr{img, pitch, img_on_device};
    return keypoints_out(c, out, cap, 0, out_on_device, count, s, nullptr, &fr);
}

int bimp_detect_u8_noalloc(bimp_ctx* c, const uint8_t* img, size_t pitch, int img_on_device,
                           bimp_keypoint* out, int cap, int* count, bimp_stream_t stream) {
    if (!c || !img || pitch < (size_t)c->width || cap < 0 || (cap > 0 && !out) || !count)
        return fail(c, BIMP_E_ARG, "bad arguments");
    if (cap > kNoAllocMaxKeypoints)
        return fail(c, BIMP_E_ARG, "cap " + std::to_string(cap) +
                                       " exceeds the no-allocation staging of " +
                                       std::to_string(kNoAllocMaxKeypoints) + " keypoints");
    if (c->opt.flags & BIMP_FLAG_PROFILE)
        return fail(c, BIMP_E_UNSUPPORTED, "profiling creates events: clear BIMP_FLAG_PROFILE");
    if (!img_on_device && pitch != (size_t)c->width && !c->hostIn)
        return fail(c, BIMP_E_UNSUPPORTED,
                    "a pitched host frame needs pinned staging: pass a device frame or a "
                    "contiguous host frame");
    if (!c->noallocStaging) return fail(c, BIMP_E_DEVICE, "no staging on this device");
    hipStream_t s = (hipStream_t)stream;
    c->lastStream = s;
    int st = run_filter(c, img, pitch, img_on_device, true, s);
    if (st) return st;
    return keypoints_out(c, out, cap, 0, 0, count, s, c->noallocStaging);
}

int bimp_detect_u8_matrix(bimp_ctx* c, const uint8_t* img, size_t pitch, int img_on_device,
                          float* mat, int ld, int cap, int mat_on_device, int* count,
                          bimp_stream_t stream) {
    if (!c || !img || pitch < (size_t)c->width || cap < 0 || ld < cap || ld <= 0 || !mat)
        return fail(c, BIMP_E_ARG, "bad arguments");
    hipStream_t s = (hipStream_t)stream;
    c->lastStream = s;
    const FrameIn fr{img, pitch, img_on_device};
    return keypoints_out(c, mat, cap, ld, mat_on_device, count, s, nullptr, &fr);
}

int bimp_get_map(bimp_ctx* c, int i, int which, int o, float* dst, size_t pitch, int dst_on_device,
                 bimp_stream_t stream) {
    if (!c || !dst || i < 0 || i >= (int)c->sc.size()) return BIMP_E_ARG;
    const auto& s = c->sc[i];
    hipStream_t st = (hipStream_t)stream;
    const float* src;
    int rows, cols;
    size_t spitch;
    if (which == BIMP_MAP_DOUBLE || which == BIMP_MAP_SINGLE) {
        src = which == BIMP_MAP_DOUBLE ? s.D : s.S;
        rows = s.g.H;
        cols = s.g.W;
        spitch = s.g.W;
    } else if (which == BIMP_MAP_COMPLEX && o >= 0 && o < kNumOri) {
        src = s.C + (size_t)o * s.Hc * s.ldc;
        rows = s.Hc;
        cols = s.Wc;
        spitch = s.ldc;
    } else {
        return fail(c, BIMP_E_ARG, "bad map selector");
    }
    if (pitch < (size_t)cols) return fail(c, BIMP_E_ARG, "pitch too small");
    HIPCHK(c, hipMemcpy2DAsync(dst, pitch * 4, src, spitch * 4, (size_t)cols * 4, rows,
                               dst_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    return BIMP_OK;
}

int bimp_map_view(const bimp_ctx* c, int i, int which, int o, const float** d_map, size_t* pitch,
                  int* rows, int* cols) {
    if (!c || !d_map || !pitch || !rows || !cols || i < 0 || i >= (int)c->sc.size())
        return BIMP_E_ARG;
    const auto& s = c->sc[i];
    if (which == BIMP_MAP_DOUBLE || which == BIMP_MAP_SINGLE) {
        *d_map = which == BIMP_MAP_DOUBLE ? s.D : s.S;
        *rows = s.g.H;
        *cols = s.g.W;
        *pitch = (size_t)s.g.W;
    } else if (which == BIMP_MAP_COMPLEX && o >= 0 && o < kNumOri) {
        *d_map = s.C + (size_t)o * s.Hc * s.ldc;
        *rows = s.Hc;
        *cols = s.Wc;
        *pitch = (size_t)s.ldc;
    } else {
        return BIMP_E_ARG;
    }
    return BIMP_OK;
}

int bimp_set_map(bimp_ctx* c, int i, int which, int o, const float* src, size_t pitch,
                 int src_on_device, bimp_stream_t stream) {
    if (!c || !src || i < 0 || i >= (int)c->sc.size()) return BIMP_E_ARG;
    auto& s = c->sc[i];
    hipStream_t st = (hipStream_t)stream;
    float* dst;
    int rows, cols;
    size_t dpitch;
    if (which == BIMP_MAP_DOUBLE || which == BIMP_MAP_SINGLE) {
        dst = which == BIMP_MAP_DOUBLE ? s.D : s.S;
        rows = s.g.H;
        cols = s.g.W;
        dpitch = s.g.W;
        if (which == BIMP_MAP_DOUBLE) c->maxValid = false;
    } else if (which == BIMP_MAP_COMPLEX && o >= 0 && o < kNumOri) {
        dst = s.C + (size_t)o * s.Hc * s.ldc;
        rows = s.Hc;
        cols = s.Wc;
        dpitch = s.ldc;
    } else {
        return fail(c, BIMP_E_ARG, "bad map selector");
    }
    if (pitch < (size_t)cols) return fail(c, BIMP_E_ARG, "pitch too small");
    HIPCHK(c, hipMemcpy2DAsync(dst, dpitch * 4, src, pitch * 4, (size_t)cols * 4, rows,
                               src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    HIPCHK(c, hipStreamSynchronize(st));
    return BIMP_OK;
}

int bimp_keypoints_to_matrix(const bimp_keypoint* k, int n, float* mat, int ld) {
    if (n < 0 || ld < n || (n > 0 && (!k || !mat))) return BIMP_E_ARG;
    for (int j = 0; j < n; ++j) {
        mat[0 * (size_t)ld + j] = k[j].x;
        mat[1 * (size_t)ld + j] = k[j].y;
        const int one = 1, oct = k[j].octave;
        std::memcpy(&mat[2 * (size_t)ld + j], &one, 4);
        std::memcpy(&mat[3 * (size_t)ld + j], &oct, 4);
        mat[4 * (size_t)ld + j] = k[j].size;
        mat[5 * (size_t)ld + j] = k[j].angle;
        mat[6 * (size_t)ld + j] = k[j].response;
    }
    return BIMP_OK;
}

int bimp_matrix_to_keypoints(const float* mat, int ld, int n, bimp_keypoint* k) {
    if (n < 0 || ld < n || (n > 0 && (!k || !mat))) return BIMP_E_ARG;
    for (int j = 0; j < n; ++j) {
        k[j].x = mat[0 * (size_t)ld + j];
        k[j].y = mat[1 * (size_t)ld + j];
        std::memcpy(&k[j].class_id, &mat[2 * (size_t)ld + j], 4);
        std::memcpy(&k[j].octave, &mat[3 * (size_t)ld + j], 4);
        k[j].size = mat[4 * (size_t)ld + j];
        k[j].angle = mat[5 * (size_t)ld + j];
        k[j].response = mat[6 * (size_t)ld + j];
    }
    return BIMP_OK;
}

const char* bimp_last_error(const bimp_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

const char* bimp_status_string(int s) {
    switch (s) {
        case BIMP_OK: return "ok";
        case BIMP_W_OVERFLOW: return "keypoint overflow";
        case BIMP_E_ARG: return "invalid argument";
        case BIMP_E_ALLOC: return "allocation failed";
        case BIMP_E_DEVICE: return "device error";
        case BIMP_E_UNSUPPORTED: return "unsupported transform size";
        default: return "unknown status";
    }
}

int bimp_get_profile(bimp_ctx* c, bimp_profile_entry* out, int cap, int* n) {
    if (!c || !n) return BIMP_E_ARG;
    collect_profile(c);
    int k = 0;
    for (const auto& kv : c->prof) {
        if (out && k < cap) {
            std::snprintf(out[k].name, sizeof(out[k].name), "%s", kv.first.c_str());
            out[k].launches = kv.second.launches;
            out[k].total_ms = kv.second.ms;
            out[k].bytes = kv.second.bytes;
        }
        ++k;
    }
    *n = k;
    return BIMP_OK;
}

int bimp_reset_profile(bimp_ctx* c) {
    if (!c) return BIMP_E_ARG;
    collect_profile(c);
    c->prof.clear();
    return BIMP_OK;
}

int bimp_set_flags(bimp_ctx* c, int flags) {
    if (!c) return BIMP_E_ARG;
    collect_profile(c);
    c->opt.flags = flags;
    return BIMP_OK;
}

int bimp_device_count(int* n) {
    if (!n) return BIMP_E_ARG;
    int k = 0;
    if (hipGetDeviceCount(&k) != hipSuccess) k = 0;
    *n = k;
    return BIMP_OK;
}

int bimp_device_malloc(void** p, size_t bytes) {
    if (!p) return BIMP_E_ARG;
    return hipMalloc(p, bytes ? bytes : 1) == hipSuccess ? BIMP_OK : BIMP_E_ALLOC;
}
int bimp_device_free(void* p) { return hipFree(p) == hipSuccess ? BIMP_OK : BIMP_E_DEVICE; }
int bimp_memcpy_h2d(void* d, const void* s, size_t b) {
    return hipMemcpy(d, s, b, hipMemcpyHostToDevice) == hipSuccess ? BIMP_OK : BIMP_E_DEVICE;
}
int bimp_memcpy_h2d_2d(void* dst, size_t dpitch, const void* src, size_t spitch,
                       size_t width_bytes, size_t rows) {
    if (!dst || !src || dpitch < width_bytes || spitch < width_bytes) return BIMP_E_ARG;
    if (!rows || !width_bytes) return BIMP_OK;
    // one contiguous upload: HIP's pitched copy from pageable memory is many times slower
    std::vector<uint8_t> packed;
    const uint8_t* s8 = static_cast<const uint8_t*>(src);
    if (spitch != width_bytes) {
        packed.resize(width_bytes * rows);
        for (size_t y = 0; y < rows; ++y)
            std::memcpy(packed.data() + y * width_bytes, s8 + y * spitch, width_bytes);
        s8 = packed.data();
    }
    if (dpitch == width_bytes)
        return hipMemcpy(dst, s8, width_bytes * rows, hipMemcpyHostToDevice) == hipSuccess
                   ? BIMP_OK
                   : BIMP_E_DEVICE;
    for (size_t y = 0; y < rows; ++y)
        if (hipMemcpy(static_cast<uint8_t*>(dst) + y * dpitch, s8 + y * width_bytes, width_bytes,
                      hipMemcpyHostToDevice) != hipSuccess)
            return BIMP_E_DEVICE;
    return BIMP_OK;
}
int bimp_memcpy_d2h(void* d, const void* s, size_t b) {
    return hipMemcpy(d, s, b, hipMemcpyDeviceToHost) == hipSuccess ? BIMP_OK : BIMP_E_DEVICE;
}
int bimp_stream_create(bimp_stream_t* s) {
    if (!s) return BIMP_E_ARG;
    hipStream_t h;
    if (hipStreamCreate(&h) != hipSuccess) return BIMP_E_DEVICE;
    *s = (bimp_stream_t)h;
    return BIMP_OK;
}
int bimp_stream_destroy(bimp_stream_t s) {
    return hipStreamDestroy((hipStream_t)s) == hipSuccess ? BIMP_OK : BIMP_E_DEVICE;
}
int bimp_stream_sync(bimp_stream_t s) {
    return hipStreamSynchronize((hipStream_t)s) == hipSuccess ? BIMP_OK : BIMP_E_DEVICE;
}

int bimp_time_detect(bimp_ctx* c, const uint8_t* d_img, size_t pitch, bimp_keypoint* d_out,
                     int cap, int iters, bimp_stream_t stream, float* ms) {
    if (!c || !d_img || !d_out || iters <= 0 || !ms) return BIMP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    HIPCHK(c, hipEventRecord(a, s));
    for (int it = 0; it < iters; ++it) {
        int st = run_frame_any(c, d_img, pitch, 1, d_out, cap, 0, s);
        if (st) {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            return st;
        }
    }
    HIPCHK(c, hipEventRecord(b, s));
    HIPCHK(c, hipEventSynchronize(b));
    HIPCHK(c, hipEventElapsedTime(ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (c->opt.flags & BIMP_FLAG_PROFILE) collect_profile(c);
    return BIMP_OK;
}

int bimp_time_detect_frames(bimp_ctx* c, const uint8_t* d_img, size_t pitch, bimp_keypoint* d_out,
                            int cap, int iters, bimp_stream_t stream, float* frame_ms) {
    if (!c || !d_img || !d_out || iters <= 0 || !frame_ms) return BIMP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    std::vector<hipEvent_t> ev((size_t)iters + 1, nullptr);
    auto release = [&]() {
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    for (auto& e : ev) {
        const hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) {
            release();
            return hip_fail(c, r, "hipEventCreate");
        }
    }
    hipError_t r = hipEventRecord(ev[0], s);
    for (int it = 0; it < iters && r == hipSuccess; ++it) {
        int st = run_frame_any(c, d_img, pitch, 1, d_out, cap, 0, s);
        if (st) {
            (void)hipStreamSynchronize(s);
            release();
            return st;
        }
        r = hipEventRecord(ev[(size_t)it + 1], s);
    }
    if (r == hipSuccess) r = hipEventSynchronize(ev[(size_t)iters]);
    for (int it = 0; it < iters && r == hipSuccess; ++it)
        r = hipEventElapsedTime(&frame_ms[it], ev[(size_t)it], ev[(size_t)it + 1]);
    release();
    if (r != hipSuccess) return hip_fail(c, r, "per-frame events");
    if (c->opt.flags & BIMP_FLAG_PROFILE) collect_profile(c);
    return BIMP_OK;
}

int bimp_get_level(bimp_ctx* c, int level, uint8_t* dst, size_t pitch, bimp_stream_t stream) {
    if (!c || !dst || level < 1 || level > c->maxLevel) return BIMP_E_ARG;
    const int w = c->levelW[level], h = c->levelH[level];
    if (pitch < (size_t)w) return BIMP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(c, hipMemcpy2DAsync(dst, pitch, c->level[level], (size_t)w, (size_t)w, (size_t)h,
                               hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return BIMP_OK;
}

// CPU keypoint mode, cuda_keypoints (cuda_bimp.cu:745-808): the reference's getKeypoints with
// cpu_mode (cuda_bimp.cu:117-124) downloads D and S and extracts keypoints on the host with
// its own rules: 8-neighbour dilation ignoring outside neighbours, threshold 0.104*max(D) for
// both maps, rows and columns in [off, n - off), and a keypoint position that adds the S-map
// parabola offsets to the D-map ones when both maps peak at a pixel.
namespace {
float parabola_peak_cpu(float v1, float v2, float v3) {
    // parabolaPeak with positions (-1, 0, 1) (cuda_bimp.cu:689-711)
    const float b = (v3 - v1) * 0.5f;
    const float a = v1 + b - v2;
    return -b / (a * 2);
}
float neighbour_max(const std::vector<float>& m, int W, int H, int x, int y) {
    float r = -INFINITY;
    for (int dy = -1; dy <= 1; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= H) continue;
        for (int dx = -1; dx <= 1; ++dx) {
            const int xx = x + dx;
            if ((!dx && !dy) || xx < 0 || xx >= W) continue;
            r = std::max(r, m[(size_t)yy * W + xx]);
        }
    }
    return r;
}
}  // namespace

int bimp_keypoints_cpu_mode(bimp_ctx* c, bimp_keypoint* out, int cap, int* count,
                            bimp_stream_t stream) {
    if (!c || !count || (cap > 0 && !out)) return BIMP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    int total = 0;
    for (const Scale& sk : c->sc) {
        const auto& g = sk.g;
        const int H = g.H, W = g.W;
        std::vector<float> Dm((size_t)H * W), Sm((size_t)H * W);
        HIPCHK(c, hipMemcpyAsync(Dm.data(), sk.D, Dm.size() * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipMemcpyAsync(Sm.data(), sk.S, Sm.size() * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        float maxd = Dm[0];
        for (float v : Dm) maxd = std::max(maxd, v);
        const double threshd = 0.104 * (double)maxd;
        const int off = (int)g.lambda;
        for (int row = off; row < H - off; ++row) {
            for (int col = off; col < W - off; ++col) {
                const size_t q = (size_t)row * W + col;
                float xpos = (float)col, ypos = (float)row;
                const float sv = Sm[q], dv = Dm[q];
                if (sv - neighbour_max(Sm, W, H, col, row) > 0 && (double)sv > threshd) {
                    xpos += parabola_peak_cpu(Sm[q - 1], sv, Sm[q + 1]);
                    ypos += parabola_peak_cpu(Sm[q - W], sv, Sm[q + W]);
                }
                if (!(dv - neighbour_max(Dm, W, H, col, row) > 0 && (double)dv > threshd)) continue;
                xpos += parabola_peak_cpu(Dm[q - 1], dv, Dm[q + 1]);
                ypos += parabola_peak_cpu(Dm[q - W], dv, Dm[q + W]);
                if (total < cap) {
                    bimp_keypoint& kp = out[total];
                    kp.x = xpos * g.pyrstep;
                    kp.y = ypos * g.pyrstep;
                    kp.size = g.lambda * g.pyrstep;
                    kp.angle = -1;
                    kp.response = dv;
                    kp.octave = (int)(g.lambda * g.pyrstep);
                    kp.class_id = -1;
                }
                ++total;
            }
        }
    }
    *count = total;
    if (total > cap) {
        c->err = "keypoints (" + std::to_string(total) + ") exceed the output capacity (" +
                 std::to_string(cap) + ")";
        return BIMP_W_OVERFLOW;
    }
    return BIMP_OK;
}

int bimp_bgr_to_gray(const uint8_t* bgr, size_t bgr_pitch, int width, int height, int bgr_on_device,
                     uint8_t* d_gray, size_t gray_pitch, bimp_stream_t stream) {
    if (!bgr || !d_gray || width <= 0 || height <= 0 || bgr_pitch < (size_t)3 * width ||
        gray_pitch < (size_t)width)
        return BIMP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    const uint8_t* src = bgr;
    uint8_t* staged = nullptr;
    if (!bgr_on_device) {
        if (hipMalloc(&staged, bgr_pitch * height) != hipSuccess) return BIMP_E_ALLOC;
        if (hipMemcpyAsync(staged, bgr, bgr_pitch * height, hipMemcpyHostToDevice, s) != hipSuccess) {
            (void)hipFree(staged);
            return BIMP_E_DEVICE;
        }
        src = staged;
    }
    hipError_t e = bimpk::launch_bgr_to_gray(src, bgr_pitch, width, height, d_gray, gray_pitch, s);
    if (staged) {
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(staged);
    }
    return e == hipSuccess ? BIMP_OK : BIMP_E_DEVICE;
}

int bimp_copy_kernel(void* dst, const void* src, size_t bytes, bimp_stream_t stream) {
    if (!dst || !src) return BIMP_E_ARG;
    if (bimpk::launch_copy(dst, src, bytes, (hipStream_t)stream) != hipSuccess) return BIMP_E_DEVICE;
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? BIMP_OK : BIMP_E_DEVICE;
}

}  // extern "C"
