// golden_mc_channel.cpp -- writes the channel LLRs that the reference's
// golden Monte Carlo mode draws with one host thread, so that the oracle can
// decode and tally the very codewords the reference decoded.
//
// The reference's mode (tests/test_circulant_dma_stepper.cpp:708-757, run as
// `stepper --sfpi --monte-carlo <ebn0> <blocks> <target_fe> <max_iter> 1`):
//   sigma     = 1 / sqrt(2 R g),  R = 1/2,  g = 10^(ebn0/10)        (:720-723)
//   llr_scale = 2 / sigma^2                                         (:724)
//   thread 0: std::mt19937_64 seeded 1337, std::normal_distribution<float>
//             (0, (float)sigma), one distribution object for the whole run
//             (:734-735)
//   per batch of 4 codewords, variables v = 0..2047 in order, codewords
//   0..3 inside each variable: y = 1 + n, LLR = bf16_rne((float)llr_scale * y);
//   variables 2048..2559 (punctured) are +0                          (:744-757)
//
// The draws come from the C++ standard library itself, so the stream is the
// one the reference's build on this toolchain (libstdc++) produces.
//
// Usage: golden_mc_channel <ebn0_db> <n_codewords (multiple of 4)> <out.bin>
// Output: n_codewords x 2560 little-endian bf16 bit patterns, codeword-major.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

namespace {

constexpr int kVars = 2560;
constexpr int kTx = 2048;

uint16_t bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, sizeof u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <ebn0_db> <n_codewords> <out.bin>\n", argv[0]);
        return 2;
    }
    const double ebn0_db = std::atof(argv[1]);
    const long n_cw = std::atol(argv[2]);
    if (n_cw <= 0 || n_cw % 4) {
        std::fprintf(stderr, "n_codewords must be a positive multiple of 4\n");
        return 2;
    }
    const double rate = 0.5;
    const double g = std::pow(10.0, ebn0_db / 10.0);
    const double sigma = 1.0 / std::sqrt(2.0 * rate * g);
    const float scale = static_cast<float>(2.0 / (sigma * sigma));
    std::mt19937_64 rng(1337);
    std::normal_distribution<float> dist(0.0f, static_cast<float>(sigma));

    std::vector<uint16_t> out(static_cast<size_t>(n_cw) * kVars, 0);
    for (long b = 0; b < n_cw / 4; ++b) {
        uint16_t *cw0 = out.data() + static_cast<size_t>(b) * 4 * kVars;
        for (int v = 0; v < kTx; ++v)
            for (int c = 0; c < 4; ++c) {
                const float y = 1.0f + dist(rng);
                cw0[static_cast<size_t>(c) * kVars + v] = bf16_rne(scale * y);
            }
    }
    FILE *f = std::fopen(argv[3], "wb");
    if (!f) {
        std::perror(argv[3]);
        return 1;
    }
    const size_t w = std::fwrite(out.data(), sizeof(uint16_t), out.size(), f);
    std::fclose(f);
    return w == out.size() ? 0 : 1;
}
