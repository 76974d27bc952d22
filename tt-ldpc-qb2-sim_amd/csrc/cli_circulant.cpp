// cli_circulant.cpp -- ldpc_sim_circulant: the circulant DMA engine's command
// line on top of libldpc (Profile T, AR4JA r1/2 K=1024, one MI355X).
//
// Arguments as the reference binary (ldpc_sim_circulant_dma.cpp:526-568):
//   ldpc_sim_circulant [ebn0=2.5] [max_blocks=1e10] [max_iter=200]
//       [--single-core | --mesh] [--min-fe N=25] [--chunk-batches N=2500]
//       [--out results.csv] [--seed N | random] [--beta B=0.9375] [--trap]
// A chunk is chunk_batches x 4 codewords x 440 cores (x 1 with
// --single-core), as on the mesh; the run stops at max_blocks or once
// min_fe frame errors are found.  Per chunk: one console line and, with
// --out, one cumulative CSV row in the engine's format
// (ldpc_sim_circulant_dma.cpp:282-290, 395-406):
//   EbN0_dB,Total_Blocks,Total_Bit_Errors,Total_Block_Errors,BER,FER,
//   Mean_Iterations,Throughput_Msps,Host_Wall_s
// with BER over the 1024 information bits and Msps over 2048 symbols.
// --trap writes the Packed4CW dumps tensix_initial_vnodes.bin and
// tensix_final_vnodes.bin of the group of four codewords holding the first
// (lowest id) failing codeword, or of codewords 0..3 if max_blocks <= 100,
// and prints the trap report (ldpc_sim_circulant_dma.cpp:435-472).
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <random>
#include <string>
#include <sys/stat.h>
#include <utility>
#include <vector>

#include "../../include/ldpc_mi355x.h"

int main(int argc, char **argv) {
    double ebn0 = 2.5;
    uint64_t max_blocks = 10000000000ULL, min_fe = 25;
    uint32_t max_iter = 200, chunk_batches = 2500;
    bool single = false, trap = false;
    std::string csv;
    uint64_t seed = 0;
    float beta = 0.9375f;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--single-core") single = true;
        else if (a == "--mesh") single = false;
        else if (a == "--min-fe" && i + 1 < argc) min_fe = std::stoull(argv[++i]);
        else if (a == "--chunk-batches" && i + 1 < argc) chunk_batches = (uint32_t)std::stoul(argv[++i]);
        else if (a == "--out" && i + 1 < argc) csv = argv[++i];
        else if (a == "--beta" && i + 1 < argc) beta = std::stof(argv[++i]);
        else if (a == "--trap") trap = true;
        else if (a == "--seed" && i + 1 < argc) {
            const std::string s = argv[++i];
            seed = (s == "random" || s == "rand") ? std::random_device{}() : std::stoull(s);
        } else if (a == "--random-seed") seed = std::random_device{}();
    }
    if (argc > 1 && argv[1][0] != '-') ebn0 = std::stod(argv[1]);
    if (argc > 2 && argv[2][0] != '-') max_blocks = std::stoull(argv[2]);
    if (argc > 3 && argv[3][0] != '-') max_iter = (uint32_t)std::stoul(argv[3]);

    ldpc_code_desc code{};
    code.family = LDPC_CODE_AR4JA;
    code.k_info = 1024;
    code.rate_num = 1;
    code.rate_den = 2;
    ldpc_decoder_cfg cfg{};
    cfg.profile = LDPC_PROFILE_T;
    cfg.max_iter = max_iter;
    cfg.beta = beta;
    ldpc_ctx *ctx = nullptr;
    int rc = ldpc_create(&code, &cfg, 0, &ctx);
    if (rc) {
        std::cerr << "libldpc error " << rc << ": " << ldpc_last_error(nullptr) << std::endl;
        return 1;
    }
    const uint64_t cores = single ? 1 : 440;
    const uint64_t chunk = (uint64_t)chunk_batches * 4 * cores;
    std::cout << "MI355X circulant engine (libldpc Profile T, beta " << beta << "): Eb/N0 " << ebn0 << " dB, max "
              << max_blocks << " blocks, " << max_iter << " iterations, chunk " << chunk << " codewords, seed "
              << seed << std::endl;

    std::ofstream out;
    if (!csv.empty()) {
        struct stat st;
        const bool exists = stat(csv.c_str(), &st) == 0;
        out.open(csv, std::ios::app);
        if (!exists)
            out << "EbN0_dB,Total_Blocks,Total_Bit_Errors,Total_Block_Errors,BER,FER,Mean_Iterations,"
                   "Throughput_Msps,Host_Wall_s\n";
        out.flush();
    }
    if (trap && (rc = ldpc_set_capture(ctx, 64))) {
        std::cerr << "libldpc error " << rc << ": " << ldpc_last_error(ctx) << std::endl;
        return 1;
    }
    uint64_t cws = 0, fe = 0, be = 0, its = 0, next = 0;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t k = 0; cws < max_blocks && fe < min_fe; ++k) {
        const uint64_t n = std::min(chunk, max_blocks - cws);
        ldpc_point_cfg pt{ebn0, seed, next, n, 0, std::min<uint64_t>(n, 1ull << 22)};
        ldpc_point_result r{};
        if ((rc = ldpc_run_point(ctx, &pt, &r, nullptr))) {
            std::cerr << "libldpc error " << rc << ": " << ldpc_last_error(ctx) << std::endl;
            return 1;
        }
        cws += r.blocks;
        fe += r.frame_errors;
        be += r.bit_errors;
        its += r.iter_sum;
        next = r.next_cw;
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const double f = cws ? (double)fe / cws : 0.0, b = cws ? be / (cws * 1024.0) : 0.0;
        const double mi = cws ? (double)its / cws : 0.0, msps = cws * 2048.0 / (sec * 1e6);
        std::cout << "[CHUNK " << std::setw(4) << k + 1 << "] Blocks: " << std::setw(11) << cws << " / " << max_blocks
                  << " | FE: " << std::setw(4) << fe << " / " << min_fe << " | BE: " << std::setw(6) << be
                  << " | FER: " << std::scientific << std::setprecision(2) << f << " | MeanIter: " << std::fixed
                  << std::setprecision(2) << mi << " | Rate: " << std::setprecision(1) << msps << " Msps"
                  << " | Wall: " << std::setprecision(1) << sec << "s" << std::endl;
        if (out.is_open()) {
            out << std::fixed << std::setprecision(2) << ebn0 << "," << cws << "," << be << "," << fe << ","
                << std::scientific << std::setprecision(6) << b << "," << f << "," << std::fixed
                << std::setprecision(3) << mi << "," << std::setprecision(2) << msps << "," << sec << "\n";
            out.flush();
        }
    }
    if (trap) {
        uint64_t ids[64], seen = 0;
        const int nc = ldpc_get_captures(ctx, ids, 64, &seen);
        if (nc < 0) {
            std::cerr << "libldpc error " << nc << ": " << ldpc_last_error(ctx) << std::endl;
            return 1;
        }
        if (nc > 0 || max_blocks <= 100) {
            const uint64_t first = nc > 0 ? ids[0] : 0, g = first / 4;
            const uint32_t n = 2560;
            std::vector<uint16_t> init(4 * n), fin(4 * n);
            std::vector<uint32_t> it(4);
            std::vector<uint8_t> ok(4);
            if ((rc = ldpc_gen_channel(ctx, seed, 4 * g, 4, ebn0, init.data())) ||
                (rc = ldpc_decode_llrs(ctx, init.data(), 4, fin.data(), it.data(), ok.data()))) {
                std::cerr << "libldpc error " << rc << ": " << ldpc_last_error(ctx) << std::endl;
                return 1;
            }
            uint32_t mask = 0, bits = 0;
            for (int c = 0; c < 4; ++c)
                if (!ok[c]) {
                    mask |= 1u << c;
                    for (uint32_t k = 0; k < 1024; ++k) bits += (fin[c * n + k] & 0x8000u) || fin[c * n + k] == 0;
                }
            std::cout << "-------------------------------------------------------" << std::endl;
            if (nc > 0) {
                std::cout << " >>> [FIRST BLOCK ERROR TRAPPED] <<<" << std::endl;
                std::cout << "  Failing Batch:     " << g << "  (codewords " << 4 * g << ".." << 4 * g + 3 << ")"
                          << std::endl;
                std::cout << "  Failing Codeword:  " << first - 4 * g << "  (id " << first << ")" << std::endl;
                std::cout << "  Bit Errors:        " << bits << std::endl;
                std::cout << "  Active Mask:       0x" << std::hex << mask << std::dec << std::endl;
                std::cout << "  Failures seen:     " << seen << std::endl;
            } else {
                std::cout << " >>> [CAPTURING BATCH 0 VNODE STATE] <<<" << std::endl;
            }
            for (const auto &f : {std::make_pair(std::string("tensix_initial_vnodes.bin"), &init),
                                  std::make_pair(std::string("tensix_final_vnodes.bin"), &fin)}) {
                std::ofstream o(f.first, std::ios::binary);
                for (uint32_t v = 0; v < n; ++v)
                    for (int c = 0; c < 4; ++c) {
                        const uint16_t h = (*f.second)[c * n + v];
                        const char b[2] = {(char)(h & 0xFF), (char)(h >> 8)};
                        o.write(b, 2);
                    }
                std::cout << (o ? "  [SUCCESS] Saved " : "  [ERROR] Failed to write ") << f.first << " (20,480 Bytes)"
                          << std::endl;
            }
        }
    }
    ldpc_destroy(ctx);
    return 0;
}
