// Drives the header-only C++ facade (include/bimp/bimp.hpp) the way a reference caller
// drives bimp::Context: build a context on a frame, fetch keypoints, read responses.
// Usage: facade_main <width> <height> <channels> <frame.u8> <out.bin> [default]
//   channels 1 (gray) or 3 (BGR, interleaved); with "default", the Context's own bank
//   makeLambdasLog(4, 64, 2) (bimp.cpp:8), otherwise the 8-wavelength bank 4 * 2^(k/4).
//   writes: int32 prepared frame width and height; int32 count, count keypoint records, the
//   SURF_CUDA matrix (7 x count, from the device); the response views' scale count and, per
//   scale, the D map read through its device view; int32 count and records of the CPU
//   keypoint mode, then of the CPU mode with resize_cpu; count x 152 descriptor floats (SURF,
//   zeros, colour histogram); the matches of the context against itself.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "bimp/bimp.hpp"

int main(int argc, char** argv) {
    if (argc != 6 && argc != 7) return 2;
    const int w = std::atoi(argv[1]), h = std::atoi(argv[2]), cn = std::atoi(argv[3]);
    const bool default_bank = argc == 7 && std::string(argv[6]) == "default";
    std::vector<uint8_t> img((size_t)w * h * cn);
    std::ifstream(argv[4], std::ios::binary).read(reinterpret_cast<char*>(img.data()), img.size());
    std::vector<float> bank;
    for (int k = 0; k < 8; ++k) bank.push_back((float)(4.0 * std::pow(2.0, k / 4.0)));
    if (default_bank) bank = bimp::cuda::makeLambdasLog(4, 64, 2);
    try {
        std::unique_ptr<bimp::Context> owner(
            default_bank ? new bimp::Context(img.data(), w, h, (size_t)w * cn,
                                             bimp::cuda::makeLambdasLog(4, 64, 2), nullptr, cn)
                         : new bimp::Context(img.data(), w, h, (size_t)w * cn, bank, nullptr, cn));
        bimp::Context& ctx = *owner;
        std::ofstream out(argv[5], std::ios::binary);
        auto put_int = [&](int v) { out.write(reinterpret_cast<const char*>(&v), sizeof v); };
        put_int(ctx.frameWidth());
        put_int(ctx.frameHeight());
        const std::vector<bimp::KeyPoint> kps = ctx.getKeypointsAndDownload();
        const bimp::Context::DeviceKeypoints& dk = ctx.getKeypoints();
        const int n = (int)kps.size();
        if (dk.count != n) return 4;
        std::vector<float> full((size_t)BIMP_KP_VECTOR_LENGTH * dk.ld);
        bimp::check(bimp_memcpy_d2h(full.data(), dk.d_matrix, full.size() * sizeof(float)));
        put_int(n);
        out.write(reinterpret_cast<const char*>(kps.data()), (std::streamsize)(n * sizeof(bimp::KeyPoint)));
        for (int r = 0; r < BIMP_KP_VECTOR_LENGTH; ++r)
            out.write(reinterpret_cast<const char*>(full.data() + (size_t)r * dk.ld),
                      (std::streamsize)((size_t)n * sizeof(float)));
        const auto views = ctx.getResponses();
        if (views.size() != bank.size()) return 3;
        put_int((int)views.size());
        for (const auto& v : views) {
            std::vector<float> d((size_t)v.info.data_h * v.pitch);
            bimp::check(bimp_memcpy_d2h(d.data(), v.d_double, d.size() * sizeof(float)));
            for (int y = 0; y < v.info.data_h; ++y)
                out.write(reinterpret_cast<const char*>(d.data() + (size_t)y * v.pitch),
                          (std::streamsize)((size_t)v.info.data_w * sizeof(float)));
        }
        // the reference's CPU keypoint mode (Context::setCPUMode)
        ctx.setCPUMode(true);
        const std::vector<bimp::KeyPoint> cpu = ctx.getKeypointsAndDownload();
        const int nc = (int)cpu.size();
        put_int(nc);
        out.write(reinterpret_cast<const char*>(cpu.data()), (std::streamsize)(nc * sizeof(bimp::KeyPoint)));
        // Context::resize_cpu in the CPU mode: host-rounded pyramid levels
        ctx.resize_cpu = true;
        const std::vector<bimp::KeyPoint> cpu_r = ctx.getKeypointsAndDownload();
        put_int((int)cpu_r.size());
        out.write(reinterpret_cast<const char*>(cpu_r.data()),
                  (std::streamsize)(cpu_r.size() * sizeof(bimp::KeyPoint)));
        ctx.resize_cpu = false;
        ctx.setCPUMode(false);
        // getDescriptors: count x BIMP_DIMENSIONS (SURF, zeros, colour histogram)
        const std::vector<float> desc = ctx.getDescriptorsAndDownload();
        if (desc.size() != (size_t)n * BIMP_DIMENSIONS) return 5;
        out.write(reinterpret_cast<const char*>(desc.data()), (std::streamsize)(desc.size() * sizeof(float)));
        // matchContexts against itself: every keypoint's nearest neighbour is itself
        const std::vector<bimp_match> self = bimp::matchContextsAndDownload(ctx, ctx);
        put_int((int)self.size());
        out.write(reinterpret_cast<const char*>(self.data()), (std::streamsize)(self.size() * sizeof(bimp_match)));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
