// ref_scalar_driver.cpp -- fixture writer around the reference's scalar_5w
// decoder (engines/scalar_5w/ldpc_decoder_core.h, included where it lies).
//
// TEST INFRASTRUCTURE ONLY (built into oracle/_ref by oracle/Makefile.ref).
// It decodes one pair of codewords with the reference's own
// run_ldpc_worker (ldpc_decoder_core.h:214-586), exactly as one RISC-V worker
// does on silicon, and records:
//   - the channel LLRs the worker generated for the pair, reproduced by
//     drawing from the same Xoshiro128Plus state with the header's own
//     generate_awgn_llr_pair (ldpc_decoder_core.h:128-152) in the worker's
//     order (cw0 pair, cw1 pair, for each pair of variables) and clamp;
//   - the final packed LLRs left in channel_llrs_pair;
//   - the worker's stats (bit errors, frame errors, cw0 iterations).
// Build with -DUSE_NORMALIZED_MIN_SUM for the min-sum variant.
//
// usage: ref_scalar_driver <matrix.chinn.out> <ebn0_db> <punctured> <max_iter>
//                          <seed> <l_max> <r_max> <l_post_max> <out.bin>
// out.bin: u32 N, u32 M, u32 n_tx, u32 max_iter, u32 in[N], u32 out[N],
//          u32 stats[4]   (pairs: low half codeword 0, high half codeword 1)
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "ldpc_decoder_core.h"

int main(int argc, char **argv) {
    if (argc != 10) {
        std::fprintf(stderr, "usage: %s matrix ebn0 punctured max_iter seed l_max r_max l_post_max out.bin\n", argv[0]);
        return 2;
    }
    const std::string path = argv[1];
    const double ebn0 = std::atof(argv[2]);
    const uint32_t P = (uint32_t)std::atoi(argv[3]);
    const uint32_t max_iter = (uint32_t)std::atoi(argv[4]);
    uint32_t seed = (uint32_t)std::strtoul(argv[5], nullptr, 0);
    const float l_max = (float)std::atof(argv[6]);
    const float r_max = (float)std::atof(argv[7]);
    const float l_post = (float)std::atof(argv[8]);

    // .chinn.out triplets, last line = dimensions; rows keep file order
    std::ifstream f(path);
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", path.c_str());
        return 2;
    }
    std::vector<uint32_t> R, C;
    std::vector<float> V;
    std::string line;
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        uint32_t r, c;
        float v;
        if (ss >> r >> c >> v) {
            R.push_back(r);
            C.push_back(c);
            V.push_back(v);
        }
    }
    const uint32_t M = R.back(), N = C.back();
    std::vector<std::vector<uint16_t>> rows(M);
    for (size_t i = 0; i + 1 < R.size(); ++i)
        if (std::fabs(V[i]) > 1e-6f) rows[R[i] - 1].push_back((uint16_t)(C[i] - 1));
    uint32_t dmax = 0;
    for (auto &r : rows) dmax = r.size() > dmax ? (uint32_t)r.size() : dmax;
    if (dmax > 8) {
        std::fprintf(stderr, "degree %u exceeds the worker's message stride\n", dmax);
        return 2;
    }
    const uint32_t stride = dmax + 1;
    std::vector<uint16_t> h(M * stride, 0xFFFF);
    for (uint32_t m = 0; m < M; ++m) {
        h[m * stride] = (uint16_t)rows[m].size();
        for (size_t d = 0; d < rows[m].size(); ++d) h[m * stride + 1 + d] = rows[m][d];
    }

    // mu / sigma as the engine computes them (ldpc_sim_scalar.cpp:352-355)
    const double rate = (double)(N - M) / (double)(N - P);
    const double g = std::pow(10.0, ebn0 / 10.0);
    const float mu = (float)(4.0 * rate * g), sigma = (float)std::sqrt(8.0 * rate * g);

    uint32_t s[4];
    for (int i = 0; i < 4; ++i) s[i] = splitmix32(seed);

    // reproduce the worker's channel draw for the pair
    const uint32_t ntx = N - P;
    std::vector<uint32_t> in(N, 0u);
    {
        Xoshiro128Plus rng;
        rng.init_state(s[0], s[1], s[2], s[3]);
        for (uint32_t i = 0; i < ntx; i += 2) {
            float a0, a1, b0, b1;
            generate_awgn_llr_pair(rng, mu, sigma, a0, a1);
            if (l_max > 0.0f) {
                if (a0 > l_max) a0 = l_max; else if (a0 < -l_max) a0 = -l_max;
                if (a1 > l_max) a1 = l_max; else if (a1 < -l_max) a1 = -l_max;
            }
            generate_awgn_llr_pair(rng, mu, sigma, b0, b1);
            if (l_max > 0.0f) {
                if (b0 > l_max) b0 = l_max; else if (b0 < -l_max) b0 = -l_max;
                if (b1 > l_max) b1 = l_max; else if (b1 < -l_max) b1 = -l_max;
            }
            in[i] = (uint32_t)fp32_to_bf16(a0) | ((uint32_t)fp32_to_bf16(b0) << 16);
            in[i + 1] = (uint32_t)fp32_to_bf16(a1) | ((uint32_t)fp32_to_bf16(b1) << 16);
        }
    }

    std::vector<uint32_t> llr(N, 0u), rmsg(M * MAX_DEG_STRIDE, 0u);
    uint32_t stats[4] = {0, 0, 0, 0};
    run_ldpc_worker(0, 2, N, M, dmax, P, max_iter, mu, sigma, l_max, r_max, l_post, s[0], s[1], s[2], s[3], h.data(),
                    llr.data(), rmsg.data(), stats);

    FILE *o = std::fopen(argv[9], "wb");
    if (!o) return 2;
    const uint32_t hdr[4] = {N, M, ntx, max_iter};
    std::fwrite(hdr, 4, 4, o);
    std::fwrite(in.data(), 4, N, o);
    std::fwrite(llr.data(), 4, N, o);
    std::fwrite(stats, 4, 4, o);
    std::fclose(o);
    return 0;
}
