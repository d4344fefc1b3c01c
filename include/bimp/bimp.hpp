// bimp/bimp.hpp -- header-only C++ facade of libbimp_hip.so with the reference's API shape.
//
// Mirrors daniel-koudouna/cuda-bimp's host interface (bimp.hpp:19-51 Context,
// cuda_bimp.hpp:92-146 CudaStuff / getKeypoints / makeLambdasLog / snapTransformSize,
// utils.hpp:32 downloadKeypoints) over the C ABI in bimp_hip.h.  Frames are 8-bit grayscale;
// keypoints come back in cv::KeyPoint field order, and the 7-row SURF_CUDA matrix layout is
// available for callers that consumed the reference's GpuMat.  When OpenCV's core header is
// on the include path, cv::Mat / cv::KeyPoint overloads are compiled as well.
//
// Errors: the reference exit()s inside its CUDA error macros; this facade throws
// std::runtime_error carrying bimp_last_error().  Keypoint overflow is not an error: the
// reference warns and clamps to KP_MAX_POINTS (cuda_bimp.cu:132-138), and so does this
// facade (the warning goes to stderr, once per call).
#ifndef BIMP_BIMP_HPP
#define BIMP_BIMP_HPP

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../bimp_hip.h"

#if defined(__has_include)
#if __has_include(<opencv2/core.hpp>)
#include <opencv2/core.hpp>
#define BIMP_HAVE_OPENCV 1
#endif
#endif

namespace bimp {

using KeyPoint = bimp_keypoint;

inline void check(int status, const bimp_ctx* ctx = nullptr) {
    if (status < 0) {
        const char* m = bimp_last_error(ctx);
        throw std::runtime_error(std::string("bimp: ") + bimp_status_string(status) +
                                 (m && *m ? std::string(": ") + m : std::string()));
    }
}

namespace utils {
// utils::uploadKeypoints layout (utils.cpp:109-128): 7 x n matrix, leading dimension n
inline std::vector<float> uploadKeypoints(const std::vector<KeyPoint>& pts) {
    std::vector<float> mat((size_t)BIMP_KP_VECTOR_LENGTH * (pts.empty() ? 1 : pts.size()));
    check(bimp_keypoints_to_matrix(pts.data(), (int)pts.size(), mat.data(),
                                   (int)(pts.empty() ? 1 : pts.size())));
    return mat;
}
// utils::downloadKeypoints (utils.cpp:130-150)
inline std::vector<KeyPoint> downloadKeypoints(const float* mat, int ld, int n) {
    std::vector<KeyPoint> pts((size_t)n);
    check(bimp_matrix_to_keypoints(mat, ld, n, pts.data()));
    return pts;
}
}  // namespace utils

namespace cuda {

// makeLambdasLog (cuda_bimp.cu:927-937)
inline std::vector<float> makeLambdasLog(float lambda_start, float lambda_end, int scalesPerOctave) {
    const int n = bimp_make_lambdas_log(lambda_start, lambda_end, scalesPerOctave, nullptr, 0);
    check(n);
    std::vector<float> out((size_t)n);
    bimp_make_lambdas_log(lambda_start, lambda_end, scalesPerOctave, out.data(), n);
    return out;
}

// snapTransformSize (cuda_bimp.cu:200-219)
inline int snapTransformSize(int dataSize) { return bimp_snap_transform_size(dataSize); }

// CudaStuff (cuda_bimp.hpp:92-100): per-frame-size state, one per stream/thread.
class CudaStuff {
public:
    std::vector<bimp_scale_info> ks;
    bimp_ctx* ctx = nullptr;

    CudaStuff(int width, int height, std::vector<float> lambdas,
              const bimp_options* opt = nullptr) {
        check(bimp_create(width, height, lambdas.data(), (int)lambdas.size(), opt, &ctx));
        const int n = bimp_num_scales(ctx);
        ks.resize((size_t)n);
        for (int i = 0; i < n; ++i) check(bimp_get_scale_info(ctx, i, &ks[(size_t)i]), ctx);
    }
    ~CudaStuff() {
        if (ctx) bimp_destroy(ctx);
    }
    CudaStuff(const CudaStuff&) = delete;
    CudaStuff& operator=(const CudaStuff&) = delete;
    CudaStuff(CudaStuff&& o) noexcept : ks(std::move(o.ks)), ctx(o.ctx) { o.ctx = nullptr; }
};

// getKeypoints (cuda_bimp.cu:109-140) + downloadKeypoints: keypoints of one 8-bit frame in
// (scale, row, column) order, clamped to max_points with a stderr warning on overflow.
inline std::vector<KeyPoint> getKeypoints(const uint8_t* img, size_t pitch, CudaStuff& cs,
                                          bool img_on_device = false,
                                          bimp_stream_t stream = nullptr,
                                          int max_points = BIMP_KP_MAX_POINTS) {
    std::vector<KeyPoint> out((size_t)max_points);
    int count = 0;
    const int st = bimp_detect_u8(cs.ctx, img, pitch, img_on_device ? 1 : 0, out.data(),
                                  max_points, 0, &count, stream);
    check(st, cs.ctx);
    if (st == BIMP_W_OVERFLOW)
        std::fprintf(stderr, "Warning: Keypoints (%d) exceed current keypoint maximum (%d).\n",
                     count, max_points);
    out.resize((size_t)(count < max_points ? count : max_points));
    return out;
}

// int getKeypoints(input, output, cs) (cuda_bimp.hpp:111): fills the 7-row SURF_CUDA matrix
// (leading dimension max_points) and returns the number of keypoints found (may exceed it).
inline int getKeypoints(const uint8_t* img, size_t pitch, std::vector<float>& mat, CudaStuff& cs,
                        bool img_on_device = false, bimp_stream_t stream = nullptr,
                        int max_points = BIMP_KP_MAX_POINTS) {
    mat.assign((size_t)BIMP_KP_VECTOR_LENGTH * max_points, 0.f);
    int count = 0;
    check(bimp_detect_u8_matrix(cs.ctx, img, pitch, img_on_device ? 1 : 0, mat.data(),
                                max_points, max_points, 0, &count, stream),
          cs.ctx);
    return count;
}

}  // namespace cuda

// Context (bimp.hpp:19-51, bimp.cpp:7-129), grayscale frames.  The reference constructor
// uses makeLambdasLog(4, 64, 2) (bimp.cpp:8); a bank can be given explicitly.
class Context {
public:
    struct Response {
        bimp_scale_info info;
        std::vector<float> doubles, singles;  // data_h x data_w maps (d_ResultDouble/Single)
    };

    Context(const uint8_t* gray, int width, int height, size_t pitch,
            std::vector<float> lambdas = cuda::makeLambdasLog(4, 64, 2),
            const bimp_options* opt = nullptr)
        : cuda_data(width, height, lambdas, opt), lambdas_(std::move(lambdas)),
          width_(width), height_(height) {
        loadImage(gray, pitch);
    }

    int width() const { return width_; }
    int height() const { return height_; }

    // The frame is copied: the reference keeps its own GpuMat (bimp.cpp:119-123).
    void loadImage(const uint8_t* gray, size_t pitch) {
        image_.resize((size_t)width_ * height_);
        for (int y = 0; y < height_; ++y)
            for (int x = 0; x < width_; ++x)
                image_[(size_t)y * width_ + x] = gray[(size_t)y * pitch + x];
        has_changed_keypoints = true;
    }

    // 7-row SURF_CUDA keypoint matrix (leading dimension = number of keypoints)
    const std::vector<float>& getKeypoints() {
        refresh();
        return matrix_;
    }

    std::vector<KeyPoint> getKeypointsAndDownload() {
        refresh();
        return keypoints_;
    }

    std::vector<Response> getResponses() {
        std::vector<Response> out;
        for (int i = 0; i < (int)cuda_data.ks.size(); ++i) {
            Response r;
            r.info = cuda_data.ks[(size_t)i];
            r.doubles.resize((size_t)r.info.data_h * r.info.data_w);
            r.singles.resize(r.doubles.size());
            check(bimp_get_map(cuda_data.ctx, i, BIMP_MAP_DOUBLE, 0, r.doubles.data(),
                               (size_t)r.info.data_w, 0, nullptr), cuda_data.ctx);
            check(bimp_get_map(cuda_data.ctx, i, BIMP_MAP_SINGLE, 0, r.singles.data(),
                               (size_t)r.info.data_w, 0, nullptr), cuda_data.ctx);
            out.push_back(std::move(r));
        }
        return out;
    }

    // Context::setCPUMode (bimp.hpp:43-44): keypoints from the reference's CPU extraction rules
    // (bimp_keypoints_cpu_mode) instead of the device NMS.
    void setCPUMode(bool cpu_mode) {
        if (cpu_mode != cpu_mode_) has_changed_keypoints = true;
        cpu_mode_ = cpu_mode;
    }
    bool getCPUMode() const { return cpu_mode_; }

#ifdef BIMP_HAVE_OPENCV
    explicit Context(const cv::Mat& gray, std::vector<float> lambdas = cuda::makeLambdasLog(4, 64, 2))
        : Context(gray.ptr<uint8_t>(), gray.cols, gray.rows, gray.step, std::move(lambdas)) {}
    void loadImage(const cv::Mat& gray) { loadImage(gray.ptr<uint8_t>(), gray.step); }
    std::vector<cv::KeyPoint> getKeypointsAndDownloadCV() {
        std::vector<cv::KeyPoint> out;
        for (const KeyPoint& k : getKeypointsAndDownload())
            out.emplace_back(cv::Point2f(k.x, k.y), k.size, k.angle, k.response, k.octave, k.class_id);
        return out;
    }
#endif

    cuda::CudaStuff cuda_data;
    bool has_changed_keypoints = true;

private:
    void refresh() {
        if (!has_changed_keypoints) return;
        if (cpu_mode_) {
            check(bimp_filter_u8(cuda_data.ctx, image_.data(), (size_t)width_, 0, nullptr),
                  cuda_data.ctx);
            check(bimp_cells(cuda_data.ctx, nullptr), cuda_data.ctx);
            keypoints_.resize((size_t)BIMP_KP_MAX_POINTS);
            int count = 0;
            const int st = bimp_keypoints_cpu_mode(cuda_data.ctx, keypoints_.data(),
                                                   BIMP_KP_MAX_POINTS, &count, nullptr);
            check(st, cuda_data.ctx);
            keypoints_.resize((size_t)(count < BIMP_KP_MAX_POINTS ? count : BIMP_KP_MAX_POINTS));
        } else {
            keypoints_ = cuda::getKeypoints(image_.data(), (size_t)width_, cuda_data);
        }
        matrix_ = utils::uploadKeypoints(keypoints_);
        has_changed_keypoints = false;
    }
    bool cpu_mode_ = false;
    std::vector<float> lambdas_;
    int width_ = 0, height_ = 0;
    std::vector<uint8_t> image_;
    std::vector<KeyPoint> keypoints_;
    std::vector<float> matrix_;
};

}  // namespace bimp

#endif
