// bimp/bimp.hpp -- header-only C++ facade of libbimp_hip.so with the reference's API shape.
//
// Mirrors daniel-koudouna/cuda-bimp's host interface (bimp.hpp:19-51 Context,
// cuda_bimp.hpp:92-146 CudaStuff / getKeypoints / makeLambdasLog / snapTransformSize,
// utils.hpp:32 downloadKeypoints) over the C ABI in bimp_hip.h.  Frames are 8-bit gray or BGR;
// keypoints come back in cv::KeyPoint field order, and the 7-row SURF_CUDA matrix is kept on
// the device for callers that consumed the reference's GpuMat.  When OpenCV's core header is
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

// Device memory owned by the facade (freed on destruction).
class DeviceBytes {
public:
    DeviceBytes() = default;
    explicit DeviceBytes(size_t n) { reset(n); }
    ~DeviceBytes() {
        if (p_) bimp_device_free(p_);
    }
    DeviceBytes(const DeviceBytes&) = delete;
    DeviceBytes& operator=(const DeviceBytes&) = delete;
    void reset(size_t n) {
        if (n <= n_) return;
        if (p_) bimp_device_free(p_);
        p_ = nullptr;
        n_ = 0;
        check(bimp_device_malloc(&p_, n));
        n_ = n;
    }
    template <class T = uint8_t>
    T* get() const { return static_cast<T*>(p_); }
    size_t size() const { return n_; }

private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// Context (bimp.hpp:19-51, bimp.cpp:7-129).  The reference constructor uses
// makeLambdasLog(4, 64, 2) (bimp.cpp:8); a bank can be given explicitly.  Frames are 8-bit
// gray (channels 1) or BGR (channels 3).  As getGPUImage does (cuda_bimp.cu:156-177), a frame
// whose rows (with cols >= 350) or cols are below 350 is upscaled bilinearly on the device and
// a colour frame is converted to gray there; the CudaStuff keeps the size of the frame given
// to the constructor (bimp.cpp:8-10), so detection reads the top-left width() x height() of the
// prepared frame, as the reference's does.
class Context {
public:
    // Device view of one scale's response maps (CudaKernels, cuda_bimp.hpp:38-60): the
    // context's own device memory, as Context::getResponses hands out (bimp.cpp:94-104).
    struct ResponseView {
        bimp_scale_info info;
        const float* d_double = nullptr;  // data_h x data_w, pitch floats per row
        const float* d_single = nullptr;
        size_t pitch = 0;
    };
    // Host copies of one scale's response maps (downloadResultsToCPU, bimp.cpp:75-80).
    struct Response {
        bimp_scale_info info;
        std::vector<float> doubles, singles;  // data_h x data_w maps (d_ResultDouble/Single)
    };
    // The 7-row SURF_CUDA keypoint matrix on the device (getKeypoints' GpuMat, bimp.hpp:36):
    // row r of keypoint i at d_matrix[r * ld + i], count columns.
    struct DeviceKeypoints {
        const float* d_matrix = nullptr;
        int ld = 0;
        int count = 0;
    };

    Context(const uint8_t* img, int width, int height, size_t pitch,
            std::vector<float> lambdas = cuda::makeLambdasLog(4, 64, 2),
            const bimp_options* opt = nullptr, int channels = 1)
        : cuda_data(width, height, lambdas, opt), lambdas_(std::move(lambdas)),
          width_(width), height_(height), base_flags_(initial_flags(opt)),
          flags_(base_flags_) {
        loadImage(img, width, height, pitch, channels);
    }

    int width() const { return width_; }
    int height() const { return height_; }
    // Size of the prepared frame detection reads from (>= width() x height()).
    int frameWidth() const { return frame_w_; }
    int frameHeight() const { return frame_h_; }

    // loadImage(Mat) (bimp.cpp:106-109): a host frame, prepared as by getGPUImage.
    void loadImage(const uint8_t* img, size_t pitch) { loadImage(img, width_, height_, pitch, 1); }
    void loadImage(const uint8_t* img, int w, int h, size_t pitch, int channels) {
        if (channels != 1 && channels != 3) check(BIMP_E_ARG);
        int pw = 0, ph = 0;
        float sf = 1.f;
        check(bimp_min_size_shape(w, h, &pw, &ph, &sf));
        if (pw < width_ || ph < height_) check(BIMP_E_ARG);
        const size_t row = (size_t)w * channels;
        raw_.reset(row * h);
        check(bimp_memcpy_h2d_2d(raw_.get(), row, img, pitch, row, h));
        image_.reset((size_t)pw * ph);
        check(bimp_get_gpu_image(raw_.get(), row, w, h, channels, 1, image_.get(), (size_t)pw,
                                 nullptr));
        if (channels == 3) {
            color_.reset((size_t)pw * ph * 3);
            check(bimp_get_gpu_image(raw_.get(), row, w, h, 3, 0, color_.get(), (size_t)pw * 3,
                                     nullptr));
        }
        check(bimp_stream_sync(nullptr));
        d_image_ = image_.get();
        image_pitch_ = (size_t)pw;
        frame_w_ = pw;
        frame_h_ = ph;
        has_color_ = channels == 3;
        has_changed_keypoints = true;
    }
    // loadImage(GpuMat) (bimp.cpp:111-117): a device gray frame used as it is (no resize); it
    // must stay valid while the context uses it and be at least width() x height().
    void loadImageDevice(const uint8_t* d_gray, int w, int h, size_t pitch) {
        if (!d_gray || w < width_ || h < height_ || pitch < (size_t)w) check(BIMP_E_ARG);
        d_image_ = d_gray;
        image_pitch_ = pitch;
        frame_w_ = w;
        frame_h_ = h;
        has_color_ = false;
        has_changed_keypoints = true;
    }

    // getKeypoints (bimp.cpp:61-73): the device keypoint matrix, at most KP_MAX_POINTS columns
    // (overflow warns on stderr and clamps, cuda_bimp.cu:132-138).
    const DeviceKeypoints& getKeypoints() {
        refresh();
        return device_kp_;
    }
    // getKeypointsAndDownload (bimp.cpp:89-92)
    std::vector<KeyPoint> getKeypointsAndDownload() {
        refresh();
        return keypoints_;
    }

    std::vector<ResponseView> getResponses() const {
        std::vector<ResponseView> out;
        for (int i = 0; i < (int)cuda_data.ks.size(); ++i) {
            ResponseView r;
            r.info = cuda_data.ks[(size_t)i];
            int rows = 0, cols = 0;
            check(bimp_map_view(cuda_data.ctx, i, BIMP_MAP_DOUBLE, 0, &r.d_double, &r.pitch,
                                &rows, &cols), cuda_data.ctx);
            check(bimp_map_view(cuda_data.ctx, i, BIMP_MAP_SINGLE, 0, &r.d_single, &r.pitch,
                                &rows, &cols), cuda_data.ctx);
            out.push_back(r);
        }
        return out;
    }
    std::vector<Response> downloadResultsToCPU() {
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

    // getDescriptors (bimp.cpp:32-59): one BIMP_DIMENSIONS-float row per keypoint on the
    // device: 64 SURF floats (bimp_surf_descriptors), zeros up to BIMP_SURF_DIMENSIONS, then
    // the colour histogram (zeros for a gray frame).  The reference leaves the columns between
    // its 64 SURF floats and the colour part unwritten.
    struct DeviceDescriptors {
        const float* d_desc = nullptr;
        int ld = BIMP_DIMENSIONS;
        int rows = 0;
    };
    const DeviceDescriptors& getDescriptors() {
        refresh();
        if (!has_changed_descriptors_) return device_desc_;
        const int n = device_kp_.count;
        desc_.reset((size_t)(n > 0 ? n : 1) * BIMP_DIMENSIONS * sizeof(float));
        std::vector<float> zero((size_t)(n > 0 ? n : 1) * BIMP_DIMENSIONS, 0.f);
        check(bimp_memcpy_h2d(desc_.get(), zero.data(), zero.size() * sizeof(float)));
        if (n > 0) {
            surf_descriptors_into(desc_.get<float>(), BIMP_DIMENSIONS);
            if (has_color_)
                check(bimp_color_features(device_kp_.d_matrix, device_kp_.ld, n, color_.get(),
                                          (size_t)frame_w_ * 3, frame_w_, frame_h_,
                                          desc_.get<float>() + BIMP_SURF_DIMENSIONS,
                                          BIMP_DIMENSIONS, nullptr));
            check(bimp_stream_sync(nullptr));
        }
        device_desc_ = DeviceDescriptors{desc_.get<float>(), BIMP_DIMENSIONS, n};
        has_changed_descriptors_ = false;
        return device_desc_;
    }
    std::vector<float> getDescriptorsAndDownload() {
        const DeviceDescriptors& d = getDescriptors();
        std::vector<float> out((size_t)d.rows * BIMP_DIMENSIONS);
        if (!out.empty())
            check(bimp_memcpy_d2h(out.data(), d.d_desc, out.size() * sizeof(float)));
        return out;
    }
    // The 64 SURF floats of each keypoint on the device (what matchContexts compares).
    const DeviceDescriptors& getSurfDescriptors() {
        refresh();
        if (!has_changed_surf_) return device_surf_;
        const int n = device_kp_.count;
        surf_.reset((size_t)(n > 0 ? n : 1) * 64 * sizeof(float));
        if (n > 0) {
            surf_descriptors_into(surf_.get<float>(), 64);
            check(bimp_stream_sync(nullptr));
        }
        device_surf_ = DeviceDescriptors{surf_.get<float>(), 64, n};
        has_changed_surf_ = false;
        return device_surf_;
    }

    // Context::setCPUMode (bimp.hpp:43-44): keypoints from the reference's CPU extraction rules
    // (bimp_keypoints_cpu_mode) instead of the device NMS.
    void setCPUMode(bool cpu_mode) {
        if (cpu_mode != cpu_mode_) has_changed_keypoints = true;
        cpu_mode_ = cpu_mode;
    }
    bool getCPUMode() const { return cpu_mode_; }

#ifdef BIMP_HAVE_OPENCV
    explicit Context(const cv::Mat& image, std::vector<float> lambdas = cuda::makeLambdasLog(4, 64, 2))
        : Context(image.ptr<uint8_t>(), image.cols, image.rows, image.step, std::move(lambdas),
                  nullptr, image.channels()) {}
    void loadImage(const cv::Mat& image) {
        loadImage(image.ptr<uint8_t>(), image.cols, image.rows, image.step, image.channels());
    }
    std::vector<cv::KeyPoint> getKeypointsAndDownloadCV() {
        std::vector<cv::KeyPoint> out;
        for (const KeyPoint& k : getKeypointsAndDownload())
            out.emplace_back(cv::Point2f(k.x, k.y), k.size, k.angle, k.response, k.octave, k.class_id);
        return out;
    }
#endif

    cuda::CudaStuff cuda_data;
    bool has_changed_keypoints = true;
    // Context::resize_cpu (bimp.hpp:24): in the CPU keypoint mode, pyramid levels rounded as
    // the host cv::pyrDown rounds (BIMP_FLAG_RESIZE_CPU, computed on the device); the
    // reference reads it in that mode only (cuda_bimp.cu:113-124).
    bool resize_cpu = false;

private:
    static int initial_flags(const bimp_options* opt) {
        if (opt) return opt->flags;
        bimp_options d;
        bimp_default_options(&d);
        return d.flags;
    }
    void refresh() {
        if (resize_cpu != last_resize_cpu_) {
            last_resize_cpu_ = resize_cpu;
            if (cpu_mode_) has_changed_keypoints = true;
        }
        if (!has_changed_keypoints) return;
        const int want = base_flags_ | (cpu_mode_ && resize_cpu ? BIMP_FLAG_RESIZE_CPU : 0);
        if (want != flags_) {
            check(bimp_set_flags(cuda_data.ctx, want), cuda_data.ctx);
            flags_ = want;
        }
        const int cap = BIMP_KP_MAX_POINTS;
        if (cpu_mode_) {
            check(bimp_filter_u8(cuda_data.ctx, d_image_, image_pitch_, 1, nullptr), cuda_data.ctx);
            check(bimp_cells(cuda_data.ctx, nullptr), cuda_data.ctx);
            keypoints_.resize((size_t)cap);
            int count = 0;
            check(bimp_keypoints_cpu_mode(cuda_data.ctx, keypoints_.data(), cap, &count, nullptr),
                  cuda_data.ctx);
            if (count > cap) warn(count);
            keypoints_.resize((size_t)(count < cap ? count : cap));
            matrix_d_.reset((size_t)BIMP_KP_VECTOR_LENGTH * cap * sizeof(float));
            std::vector<float> mat((size_t)BIMP_KP_VECTOR_LENGTH * cap, 0.f);
            check(bimp_keypoints_to_matrix(keypoints_.data(), (int)keypoints_.size(), mat.data(), cap));
            check(bimp_memcpy_h2d(matrix_d_.get(), mat.data(), mat.size() * sizeof(float)));
        } else {
            matrix_d_.reset((size_t)BIMP_KP_VECTOR_LENGTH * cap * sizeof(float));
            int count = 0;
            const int st = bimp_detect_u8_matrix(cuda_data.ctx, d_image_, image_pitch_, 1,
                                                 matrix_d_.get<float>(), cap, cap, 1, &count,
                                                 nullptr);
            check(st, cuda_data.ctx);
            if (st == BIMP_W_OVERFLOW) warn(count);
            const int n = count < cap ? count : cap;
            std::vector<float> mat((size_t)BIMP_KP_VECTOR_LENGTH * cap);
            check(bimp_memcpy_d2h(mat.data(), matrix_d_.get(), mat.size() * sizeof(float)));
            keypoints_ = utils::downloadKeypoints(mat.data(), cap, n);
        }
        device_kp_.d_matrix = matrix_d_.get<float>();
        device_kp_.ld = cap;
        device_kp_.count = (int)keypoints_.size();
        has_changed_keypoints = false;
        has_changed_descriptors_ = true;
        has_changed_surf_ = true;
    }
    void surf_descriptors_into(float* d_out, int ld) {
        integral_.reset((size_t)(frame_w_ + 1) * (frame_h_ + 1) * sizeof(unsigned));
        check(bimp_integral_image(d_image_, image_pitch_, frame_w_, frame_h_,
                                  integral_.get<unsigned>(), nullptr));
        check(bimp_surf_descriptors(integral_.get<unsigned>(), frame_w_, frame_h_,
                                    device_kp_.d_matrix, device_kp_.ld, device_kp_.count, 0,
                                    d_out, ld, nullptr, nullptr));
    }
    static void warn(int count) {
        std::fprintf(stderr, "Warning: Keypoints (%d) exceed current keypoint maximum (%d).\n",
                     count, BIMP_KP_MAX_POINTS);
    }
    bool cpu_mode_ = false;
    std::vector<float> lambdas_;
    int width_ = 0, height_ = 0;
    int base_flags_ = 0, flags_ = 0;  // the options' flags; those the context has now
    bool last_resize_cpu_ = false;
    DeviceBytes raw_, image_, color_, matrix_d_, desc_, surf_, integral_;
    DeviceDescriptors device_desc_, device_surf_;
    bool has_changed_descriptors_ = true, has_changed_surf_ = true;
    const uint8_t* d_image_ = nullptr;
    size_t image_pitch_ = 0;
    int frame_w_ = 0, frame_h_ = 0;
    bool has_color_ = false;
    std::vector<KeyPoint> keypoints_;
    DeviceKeypoints device_kp_;
};

// matchContexts / matchContextsAndDownload / downloadMatches (bimp.cpp:131-240): the SURF
// descriptors of both contexts' keypoints, brute-force 2-nearest-neighbour matching on the
// device, then the ratio test dist1 < 0.6 * dist2 on the host.
inline std::vector<bimp_match> matchContextsAndDownload(Context& c1, Context& c2) {
    const Context::DeviceDescriptors& q = c1.getSurfDescriptors();
    const Context::DeviceDescriptors& t = c2.getSurfDescriptors();
    std::vector<bimp_match> out;
    if (q.rows == 0) return out;
    DeviceBytes idx((size_t)q.rows * 2 * sizeof(int)), dist((size_t)q.rows * 2 * sizeof(float));
    check(bimp_match_knn2(q.d_desc, q.rows, q.ld, t.d_desc, t.rows, t.ld, 64, idx.get<int>(),
                          dist.get<float>(), nullptr));
    check(bimp_stream_sync(nullptr));
    std::vector<int> hi((size_t)q.rows * 2);
    std::vector<float> hd((size_t)q.rows * 2);
    check(bimp_memcpy_d2h(hi.data(), idx.get(), hi.size() * sizeof(int)));
    check(bimp_memcpy_d2h(hd.data(), dist.get(), hd.size() * sizeof(float)));
    out.resize((size_t)q.rows);
    int n = 0;
    check(bimp_ratio_matches(hi.data(), hd.data(), q.rows, 0.6, out.data(), q.rows, &n));
    out.resize((size_t)n);
    return out;
}

}  // namespace bimp

#endif
