// Seeded, portable synthetic frame generator used for benchmark and test inputs.
//
// The reference ships no sample images, so frames are synthesised: a constant background
// level drawn from U[0,255], N shapes (rectangle or ellipse with probability 1/2 each;
// centre uniform over the frame, width and height each U[10,120] px, rotation U[0,pi),
// intensity U[0,255]) painted in order with a pixel-centre inside test, then additive
// Gaussian noise N(0, 5^2) per pixel, round-half-even and clamp to uint8.
//
// Random numbers: SplitMix64 expands the 64-bit seed into the 256-bit state of
// xoshiro256**; a uniform double is the top 53 bits times 2^-53; normals come from
// Box-Muller in FP64 (one normal per pair of uniforms, the cosine branch).
// Everything is plain IEEE FP64 plus libm sqrt/log/cos/sin, so two machines running the
// same image regenerate byte-identical frames (the tests pin SHA-256 hashes).
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace {

struct Xoshiro256ss {
    uint64_t s[4];
    static uint64_t splitmix64(uint64_t& x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Xoshiro256ss(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    // uniform in [0, 1)
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

extern "C" int bimp_synth_frame(int width, int height, int nshapes, uint64_t seed,
                                uint8_t* out, size_t pitch) {
    if (width <= 0 || height <= 0 || nshapes < 0 || !out || pitch < (size_t)width) return -1;
    const double kPi = 3.14159265358979323846;
    Xoshiro256ss rng(seed);
    std::vector<double> canvas((size_t)width * height, 255.0 * rng.uniform());

    for (int s = 0; s < nshapes; ++s) {
        const bool ellipse = rng.uniform() < 0.5;
        const double cx = width * rng.uniform();
        const double cy = height * rng.uniform();
        const double sw = 10.0 + 110.0 * rng.uniform();
        const double sh = 10.0 + 110.0 * rng.uniform();
        const double rot = kPi * rng.uniform();
        const double val = 255.0 * rng.uniform();
        const double c = std::cos(rot), sn = std::sin(rot);
        const double hw = 0.5 * sw, hh = 0.5 * sh;
        // conservative axis-aligned bounding box of the rotated shape
        const double ext = std::sqrt(hw * hw + hh * hh) + 1.0;
        int x0 = (int)std::floor(cx - ext), x1 = (int)std::ceil(cx + ext);
        int y0 = (int)std::floor(cy - ext), y1 = (int)std::ceil(cy + ext);
        if (x0 < 0) x0 = 0;
        if (y0 < 0) y0 = 0;
        if (x1 > width - 1) x1 = width - 1;
        if (y1 > height - 1) y1 = height - 1;
        for (int y = y0; y <= y1; ++y) {
            for (int x = x0; x <= x1; ++x) {
                const double dx = (x + 0.5) - cx, dy = (y + 0.5) - cy;
                const double u = dx * c + dy * sn;
                const double v = -dx * sn + dy * c;
                bool inside;
                if (ellipse) {
                    const double a = u / hw, b = v / hh;
                    inside = a * a + b * b <= 1.0;
                } else {
                    inside = std::fabs(u) <= hw && std::fabs(v) <= hh;
                }
                if (inside) canvas[(size_t)y * width + x] = val;
            }
        }
    }

    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            const double u1 = 1.0 - rng.uniform();  // (0, 1]
            const double u2 = rng.uniform();
            const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
            double v = std::nearbyint(canvas[(size_t)y * width + x] + 5.0 * z);  // half-even
            if (v < 0.0) v = 0.0;
            if (v > 255.0) v = 255.0;
            out[(size_t)y * pitch + x] = (uint8_t)v;
        }
    }
    return 0;
}
