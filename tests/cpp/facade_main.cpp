// Drives the header-only C++ facade (include/bimp/bimp.hpp) the way a reference caller
// drives bimp::Context: build a context on a frame, fetch keypoints, read responses.
// Usage: facade_main <width> <height> <frame.u8> <out.bin> [default]
//   with "default", the Context's own bank makeLambdasLog(4, 64, 2) (bimp.cpp:8); otherwise
//   the 8-wavelength bank 4 * 2^(k/4)
//   writes: int32 count, then count keypoint records, then the SURF_CUDA matrix (7 x count),
//   then int32 count and records of the CPU keypoint mode
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "bimp/bimp.hpp"

int main(int argc, char** argv) {
    if (argc != 5 && argc != 6) return 2;
    const bool default_bank = argc == 6 && std::string(argv[5]) == "default";
    const int w = std::atoi(argv[1]), h = std::atoi(argv[2]);
    std::vector<uint8_t> img((size_t)w * h);
    std::ifstream(argv[3], std::ios::binary).read(reinterpret_cast<char*>(img.data()), img.size());
    std::vector<float> bank;
    for (int k = 0; k < 8; ++k) bank.push_back((float)(4.0 * std::pow(2.0, k / 4.0)));
    try {
        if (default_bank) bank = bimp::cuda::makeLambdasLog(4, 64, 2);
        std::unique_ptr<bimp::Context> owner(
            default_bank ? new bimp::Context(img.data(), w, h, (size_t)w)
                         : new bimp::Context(img.data(), w, h, (size_t)w, bank));
        bimp::Context& ctx = *owner;
        const std::vector<bimp::KeyPoint> kps = ctx.getKeypointsAndDownload();
        const std::vector<float>& mat = ctx.getKeypoints();
        const auto resp = ctx.getResponses();
        if (resp.size() != bank.size()) return 3;
        std::ofstream out(argv[4], std::ios::binary);
        const int n = (int)kps.size();
        out.write(reinterpret_cast<const char*>(&n), sizeof n);
        out.write(reinterpret_cast<const char*>(kps.data()), (std::streamsize)(n * sizeof(bimp::KeyPoint)));
        out.write(reinterpret_cast<const char*>(mat.data()), (std::streamsize)(7 * (size_t)n * sizeof(float)));
        // the reference's CPU keypoint mode (Context::setCPUMode)
        ctx.setCPUMode(true);
        const std::vector<bimp::KeyPoint> cpu = ctx.getKeypointsAndDownload();
        const int nc = (int)cpu.size();
        out.write(reinterpret_cast<const char*>(&nc), sizeof nc);
        out.write(reinterpret_cast<const char*>(cpu.data()), (std::streamsize)(nc * sizeof(bimp::KeyPoint)));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
