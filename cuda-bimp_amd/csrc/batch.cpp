// batch.cpp -- batching form: many frames of uneven sizes as one hipGraph launch.
//
// A batch groups its frames by size.  Each size gets one or more contexts (the reference's
// CudaStuff, one per frame size, cuda_bimp.cu:222-397); each context is one branch of the
// graph and runs its frames back to back, so frames of one size share buffers and frames of
// different sizes overlap on the device.  Every frame's keypoints are those of bimp_detect_u8
// on a context of its size.
//
// The plan search that rebuilt the graph for trial plans (tune_plan != 0) is switched off:
// bimp_batch_create returns BIMP_E_UNSUPPORTED for it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/bimp_hip.h"
#include "context.hpp"
#include "kernels.hpp"

namespace {

struct SizeClass {
    int w = 0, h = 0;
    std::vector<int> frames;       // frame indices of this size, in order
    std::vector<bimp_ctx*> ctx;    // contexts (graph branches) of this size
    int invBatch = 0;              // inverse-pass sub-batch for its contexts (0 = budget)
};

}  // namespace

struct bimp_batch {
    int nframes = 0, cap = 0;
    std::vector<float> lambdas;
    bimp_options opt{};
    std::vector<SizeClass> cls;
    std::vector<int> frameClass;
    std::vector<uint8_t*> in;
    std::vector<size_t> pitch;
    bimp_keypoint* out = nullptr;  // nframes x cap
    int* counts = nullptr;         // nframes
    hipStream_t cs = nullptr;      // capture stream
    std::vector<hipStream_t> br;   // branch streams
    std::vector<hipEvent_t> ev;    // fork + one join per branch
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipGraphExec_t> frameExec;  // one graph per frame (latency of a frame alone)
    std::string err;
};

namespace {

int berr(bimp_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    return code;
}

// Phase markers of a batch build on stderr (set BIMP_TRACE to a non-empty value): a failure
// inside the runtime then names the step it happened in.
void phase(const char* what, int a = -1, int c = -1) {
    const char* on = std::getenv("BIMP_TRACE");
    if (!on || !*on) return;
    if (a >= 0)
        std::fprintf(stderr, "bimp_batch: %s %d/%d\n", what, a, c);
    else
        std::fprintf(stderr, "bimp_batch: %s\n", what);
    std::fflush(stderr);
}

#define BCHK(b, expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return berr(b, BIMP_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

void drop_graphs(bimp_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    for (auto e : b->frameExec)
        if (e) (void)hipGraphExecDestroy(e);
    b->frameExec.clear();
}

// Context slot k of class c, created on demand.
int ensure_ctx(bimp_batch* b, SizeClass& c, size_t k) {
    while (c.ctx.size() <= k) {
        bimp_ctx* x = nullptr;
        const int st = bimp_create(c.w, c.h, b->lambdas.data(), (int)b->lambdas.size(), &b->opt, &x);
        if (st != BIMP_OK)
            return berr(b, st, std::string("context ") + std::to_string(c.w) + "x" +
                                   std::to_string(c.h) + ": " + bimp_last_error(nullptr));
        c.ctx.push_back(x);
    }
    return BIMP_OK;
}

int enqueue_frame(bimp_batch* b, bimp_ctx* x, int f, hipStream_t s) {
    const int st = bimpi::run_frame(x, b->in[f], b->pitch[f], b->out + (size_t)f * b->cap, b->cap,
                                    0, s);
    if (st != BIMP_OK) return berr(b, st, std::string("frame ") + std::to_string(f) + ": " + x->err);
    BCHK(b, bimpk::launch_store_count(x->counter, b->counts + f, s));
    return BIMP_OK;
}

// Captures work enqueued by `body` on the capture stream into an executable graph.
template <class F>
int capture(bimp_batch* b, hipGraphExec_t* exec, hipGraph_t* keep, F body) {
    BCHK(b, hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal));
    const int st = body();
    phase("end capture");
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cs, &g);
    if (st != BIMP_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return berr(b, BIMP_E_DEVICE, std::string("capture: ") + hipGetErrorString(e));
    phase("instantiate");
    const hipError_t ie = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        return berr(b, BIMP_E_DEVICE, std::string("instantiate: ") + hipGetErrorString(ie));
    }
    if (keep)
        *keep = g;
    else
        (void)hipGraphDestroy(g);
    return BIMP_OK;
}

// The batch graph: fork on the capture stream, one branch per context, join; then one graph
// per frame (the latency of a frame alone).
int build(bimp_batch* b) {
    drop_graphs(b);
    std::vector<std::pair<bimp_ctx*, std::vector<int>>> branches;
    for (auto& c : b->cls) {
        const size_t K = std::max<size_t>(1, c.ctx.size());
        for (size_t k = 0; k < K; ++k) {
            const int st = ensure_ctx(b, c, k);
            if (st) return st;
            c.ctx[k]->invBatch = c.invBatch;
            std::vector<int> fr;
            for (size_t i = k; i < c.frames.size(); i += K) fr.push_back(c.frames[i]);
            if (!fr.empty()) branches.push_back({c.ctx[k], fr});
        }
    }
    while (b->br.size() < branches.size()) {
        hipStream_t s = nullptr;
        BCHK(b, hipStreamCreate(&s));
        b->br.push_back(s);
    }
    while (b->ev.size() < branches.size() + 1) {
        hipEvent_t e = nullptr;
        BCHK(b, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->ev.push_back(e);
    }
    phase("capture batch graph, branches", (int)branches.size(), (int)branches.size());
    int st = capture(b, &b->exec, &b->graph, [&]() -> int {
        BCHK(b, hipEventRecord(b->ev[0], b->cs));
        for (size_t j = 0; j < branches.size(); ++j) {
            hipStream_t s = b->br[j];
            BCHK(b, hipStreamWaitEvent(s, b->ev[0], 0));
            for (int f : branches[j].second) {
                const int r = enqueue_frame(b, branches[j].first, f, s);
                if (r) return r;
            }
            BCHK(b, hipEventRecord(b->ev[j + 1], s));
            BCHK(b, hipStreamWaitEvent(b->cs, b->ev[j + 1], 0));
        }
        return BIMP_OK;
    });
    if (st) return st;
    b->frameExec.assign(b->nframes, nullptr);
    for (int f = 0; f < b->nframes; ++f) {
        phase("capture frame graph", f, b->nframes);
        bimp_ctx* x = b->cls[b->frameClass[f]].ctx[0];
        st = capture(b, &b->frameExec[f], nullptr, [&]() { return enqueue_frame(b, x, f, b->cs); });
        if (st) return st;
    }
    return BIMP_OK;
}

int time_batch(bimp_batch* b, int iters, hipStream_t s, float* ms) {
    hipEvent_t a, z;
    BCHK(b, hipEventCreate(&a));
    BCHK(b, hipEventCreate(&z));
    BCHK(b, hipGraphLaunch(b->exec, s));  // warm-up
    BCHK(b, hipEventRecord(a, s));
    for (int i = 0; i < iters; ++i) BCHK(b, hipGraphLaunch(b->exec, s));
    BCHK(b, hipEventRecord(z, s));
    BCHK(b, hipEventSynchronize(z));
    float t = 0;
    BCHK(b, hipEventElapsedTime(&t, a, z));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(z);
    *ms = t / std::max(1, iters);
    return BIMP_OK;
}

}  // namespace

extern "C" {

int bimp_batch_create(const int* widths, const int* heights, int nframes, const float* lambdas,
                      int n, const bimp_options* opt, int cap, int tune_plan, bimp_batch** out) {
    if (!out) return BIMP_E_ARG;
    *out = nullptr;
    if (!widths || !heights || nframes <= 0 || !lambdas || n <= 0 || cap < 0) return BIMP_E_ARG;
    // The plan search (trial graphs rebuilt with more contexts per size) ended twice in a host
    // segmentation fault inside this call, cause not found: it stays switched off.
    if (tune_plan) return BIMP_E_UNSUPPORTED;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BIMP_E_DEVICE;
    bimp_batch* b = new bimp_batch;
    b->nframes = nframes;
    b->cap = cap;
    b->lambdas.assign(lambdas, lambdas + n);
    if (opt)
        b->opt = *opt;
    else
        bimp_default_options(&b->opt);
    b->opt.flags &= ~BIMP_FLAG_PROFILE;
    // graph branches: one stream per context, no forked second stream inside a frame
    b->opt.flags |= BIMP_FLAG_SINGLE_STREAM;
    std::map<std::pair<int, int>, int> idx;
    b->frameClass.resize(nframes);
    for (int f = 0; f < nframes; ++f) {
        const auto key = std::make_pair(widths[f], heights[f]);
        auto it = idx.find(key);
        if (it == idx.end()) {
            SizeClass c;
            c.w = widths[f];
            c.h = heights[f];
            it = idx.emplace(key, (int)b->cls.size()).first;
            b->cls.push_back(c);
        }
        b->frameClass[f] = it->second;
        b->cls[it->second].frames.push_back(f);
    }
    auto bail = [&](int code) {
        bimp_batch_destroy(b);
        return code;
    };
    b->in.assign(nframes, nullptr);
    b->pitch.assign(nframes, 0);
    for (int f = 0; f < nframes; ++f) {
        b->pitch[f] = (size_t)(widths[f] + 255) / 256 * 256;
        if (widths[f] <= 0 || heights[f] <= 0 ||
            hipMalloc(&b->in[f], b->pitch[f] * heights[f]) != hipSuccess ||
            hipMemset(b->in[f], 0, b->pitch[f] * heights[f]) != hipSuccess)
            return bail(BIMP_E_ALLOC);
    }
    if (hipMalloc(&b->out, (size_t)nframes * std::max(cap, 1) * sizeof(bimp_keypoint)) != hipSuccess ||
        hipMalloc(&b->counts, (size_t)nframes * sizeof(int)) != hipSuccess ||
        hipMemset(b->counts, 0, (size_t)nframes * sizeof(int)) != hipSuccess ||
        hipStreamCreate(&b->cs) != hipSuccess)
        return bail(BIMP_E_ALLOC);
    phase("create contexts", (int)b->cls.size(), (int)b->cls.size());
    for (auto& c : b->cls) {
        const int st = ensure_ctx(b, c, 0);
        if (st) {
            std::fprintf(stderr, "bimp_batch_create: %s\n", b->err.c_str());
            return bail(st);
        }
    }
    phase("build graphs");
    const int st = build(b);
    if (st) {
        std::fprintf(stderr, "bimp_batch_create: %s\n", b->err.c_str());
        return bail(st);
    }
    (void)hipStreamSynchronize(b->cs);
    phase("created");
    *out = b;
    return BIMP_OK;
}

int bimp_batch_destroy(bimp_batch* b) {
    if (!b) return BIMP_E_ARG;
    if (b->cs) (void)hipStreamSynchronize(b->cs);
    drop_graphs(b);
    for (auto& c : b->cls)
        for (auto* x : c.ctx) bimp_destroy(x);
    for (auto* p : b->in) (void)hipFree(p);
    (void)hipFree(b->out);
    (void)hipFree(b->counts);
    for (auto s : b->br) (void)hipStreamDestroy(s);
    for (auto e : b->ev) (void)hipEventDestroy(e);
    if (b->cs) (void)hipStreamDestroy(b->cs);
    delete b;
    return BIMP_OK;
}

int bimp_batch_frame(bimp_batch* b, int i, uint8_t** d_frame, size_t* pitch) {
    if (!b || i < 0 || i >= b->nframes || !d_frame || !pitch) return BIMP_E_ARG;
    *d_frame = b->in[i];
    *pitch = b->pitch[i];
    return BIMP_OK;
}

int bimp_batch_run(bimp_batch* b, bimp_stream_t stream) {
    if (!b || !b->exec) return BIMP_E_ARG;
    BCHK(b, hipGraphLaunch(b->exec, (hipStream_t)stream));
    return BIMP_OK;
}

int bimp_batch_results(bimp_batch* b, int i, bimp_keypoint* out, int cap, int* count) {
    if (!b || i < 0 || i >= b->nframes || !count || (cap > 0 && !out)) return BIMP_E_ARG;
    int total = 0;
    BCHK(b, hipMemcpy(&total, b->counts + i, sizeof(int), hipMemcpyDeviceToHost));
    const int n = std::min(std::min(total, b->cap), cap);
    if (n > 0)
        BCHK(b, hipMemcpy(out, b->out + (size_t)i * b->cap, (size_t)n * sizeof(bimp_keypoint),
                          hipMemcpyDeviceToHost));
    *count = total;
    return total > std::min(cap, b->cap) ? BIMP_W_OVERFLOW : BIMP_OK;
}

int bimp_batch_outputs(bimp_batch* b, int i, bimp_keypoint** d_keypoints, int** d_count) {
    if (!b || i < 0 || i >= b->nframes || !d_keypoints || !d_count) return BIMP_E_ARG;
    *d_keypoints = b->out + (size_t)i * b->cap;
    *d_count = b->counts + i;
    return BIMP_OK;
}

int bimp_batch_time(bimp_batch* b, int iters, bimp_stream_t stream, float* batch_ms,
                    float* frame_ms) {
    if (!b || !b->exec || iters <= 0 || !batch_ms) return BIMP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    int st = time_batch(b, iters, s, batch_ms);
    if (st) return st;
    if (!frame_ms) return BIMP_OK;
    hipEvent_t a, z;
    BCHK(b, hipEventCreate(&a));
    BCHK(b, hipEventCreate(&z));
    for (int f = 0; f < b->nframes; ++f) {
        BCHK(b, hipGraphLaunch(b->frameExec[f], s));  // warm-up
        float best = 1e30f;
        for (int r = 0; r < 3; ++r) {
            BCHK(b, hipEventRecord(a, s));
            BCHK(b, hipGraphLaunch(b->frameExec[f], s));
            BCHK(b, hipEventRecord(z, s));
            BCHK(b, hipEventSynchronize(z));
            float t = 0;
            BCHK(b, hipEventElapsedTime(&t, a, z));
            best = std::min(best, t);
        }
        frame_ms[f] = best;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(z);
    return BIMP_OK;
}

int bimp_batch_plan(const bimp_batch* b, char* buf, int cap) {
    if (!b || !buf || cap <= 0) return BIMP_E_ARG;
    std::string p;
    for (const auto& c : b->cls) {
        if (!p.empty()) p += "; ";
        p += std::to_string(c.w) + "x" + std::to_string(c.h) + ": " +
             std::to_string(c.frames.size()) + " frames, " + std::to_string(c.ctx.size()) +
             " contexts, inverse sub-batch " + (c.invBatch ? std::to_string(c.invBatch) : "auto");
    }
    std::snprintf(buf, (size_t)cap, "%s", p.c_str());
    return (int)p.size();
}

const char* bimp_batch_last_error(const bimp_batch* b) { return b ? b->err.c_str() : ""; }

}  // extern "C"
