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
//
// CPU baseline mode (the reference's scalar worker timed on the host):
//   ref_scalar_driver --monte-carlo <matrix.chinn.out> <ebn0_db> <punctured>
//                     <max_iter> <codewords_per_thread> <threads> <seed>
// runs run_ldpc_worker once per OpenMP thread with worker_id = thread index
// (the worker jumps its Xoshiro128Plus stream by worker_id * 2^64, as the
// engine's workers do) and prints one JSON line: codewords, frame and bit
// errors, seconds, Msps (transmitted symbols per second).
#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "ldpc_decoder_core.h"

namespace {

struct Matrix {
    uint32_t M = 0, N = 0, dmax = 0, stride = 0;
    std::vector<uint16_t> h;  // [M][stride]: degree, then variables (the worker's h_col_idx layout)
};

// .chinn.out triplets (1-indexed), last line = dimensions; rows keep file order
bool read_matrix(const std::string &path, Matrix &mx) {
    std::ifstream f(path);
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", path.c_str());
        return false;
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
    mx.M = R.back();
    mx.N = C.back();
    std::vector<std::vector<uint16_t>> rows(mx.M);
    for (size_t i = 0; i + 1 < R.size(); ++i)
        if (std::fabs(V[i]) > 1e-6f) rows[R[i] - 1].push_back((uint16_t)(C[i] - 1));
    for (auto &r : rows) mx.dmax = r.size() > mx.dmax ? (uint32_t)r.size() : mx.dmax;
    if (mx.dmax > 8) {
        std::fprintf(stderr, "degree %u exceeds the worker's message stride\n", mx.dmax);
        return false;
    }
    mx.stride = mx.dmax + 1;
    mx.h.assign(mx.M * mx.stride, 0xFFFF);
    for (uint32_t m = 0; m < mx.M; ++m) {
        mx.h[m * mx.stride] = (uint16_t)rows[m].size();
        for (size_t d = 0; d < rows[m].size(); ++d) mx.h[m * mx.stride + 1 + d] = rows[m][d];
    }
    return true;
}

// mu / sigma as the engine computes them (ldpc_sim_scalar.cpp:352-355)
void channel(const Matrix &mx, uint32_t P, double ebn0, float &mu, float &sigma) {
    const double rate = (double)(mx.N - mx.M) / (double)(mx.N - P);
    const double g = std::pow(10.0, ebn0 / 10.0);
    mu = (float)(4.0 * rate * g);
    sigma = (float)std::sqrt(8.0 * rate * g);
}

int monte_carlo(int argc, char **argv) {
    if (argc != 9) {
        std::fprintf(stderr, "usage: %s --monte-carlo matrix ebn0 punctured max_iter cw_per_thread threads seed\n",
                     argv[0]);
        return 2;
    }
    Matrix mx;
    if (!read_matrix(argv[2], mx)) return 2;
    const double ebn0 = std::atof(argv[3]);
    const uint32_t P = (uint32_t)std::atoi(argv[4]);
    const uint32_t max_iter = (uint32_t)std::atoi(argv[5]);
    const uint32_t per_thread = (uint32_t)std::strtoul(argv[6], nullptr, 0);
    const int threads = std::atoi(argv[7]);
    uint32_t seed = (uint32_t)std::strtoul(argv[8], nullptr, 0);
    float mu, sigma;
    channel(mx, P, ebn0, mu, sigma);
    uint32_t s[4];
    for (int i = 0; i < 4; ++i) s[i] = splitmix32(seed);
    unsigned long long be = 0, fe = 0;
    const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel num_threads(threads) reduction(+ : be, fe)
    {
        std::vector<uint32_t> llr(mx.N, 0u), rmsg(mx.M * MAX_DEG_STRIDE, 0u);
        uint32_t stats[4] = {0, 0, 0, 0};
        run_ldpc_worker((uint32_t)omp_get_thread_num(), per_thread, mx.N, mx.M, mx.dmax, P, max_iter, mu, sigma, 0.0f,
                        0.0f, 0.0f, s[0], s[1], s[2], s[3], mx.h.data(), llr.data(), rmsg.data(), stats);
        be += stats[0];
        fe += stats[1];
    }
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const unsigned long long cw = (unsigned long long)per_thread * (unsigned long long)threads;
    std::printf("{\"codewords\": %llu, \"frame_errors\": %llu, \"bit_errors\": %llu, \"seconds\": %.6f, "
                "\"msps\": %.6f, \"threads\": %d}\n",
                cw, fe, be, sec, (double)cw * (double)(mx.N - P) / sec / 1e6, threads);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc > 1 && std::string(argv[1]) == "--monte-carlo") return monte_carlo(argc, argv);
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

    Matrix mx;
    if (!read_matrix(path, mx)) return 2;
    const uint32_t M = mx.M, N = mx.N, dmax = mx.dmax;
    float mu, sigma;
    channel(mx, P, ebn0, mu, sigma);

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
    run_ldpc_worker(0, 2, N, M, dmax, P, max_iter, mu, sigma, l_max, r_max, l_post, s[0], s[1], s[2], s[3],
                    mx.h.data(), llr.data(), rmsg.data(), stats);

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
