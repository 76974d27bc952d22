// cli_scalar.cpp -- ldpc_sim_scalar: the scalar_5w engine's command line on
// top of libldpc (Profile R / R_NMS on one MI355X).
//
// Same getopt flags and stdout protocol as the reference engine binary
// (engines/scalar_5w/ldpc_sim_scalar.cpp:271-641), so run_simulation.py's
// --binary path drives it unchanged:
//   -c matrix (.chinn.out triplets or .chinn rows)   -e Eb/N0 dB
//   -p punctured   -t target frame errors   -k max blocks per core
//   -m max blocks (total)   -b batch per core   -i max iterations
//   -s single core   -n normalized min-sum   -v lockstep verify
//   -L / -R / -V clamps on channel LLR / message / posterior, -G gamma
// Block accounting keeps the reference's 440-core mesh (1 with -s): the
// engine simulates blocks_per_core x cores codewords in batches of
// batch_per_core x cores.  stdout: non-JSON [CONFIG] lines, one progress
// JSON line per batch, a final JSON line; non-zero exit on errors.
// -v runs every batch twice on the same codeword ids and fails on any
// difference (the reference's lockstep check, ldpc_sim_scalar.cpp:545-605).
#include <getopt.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ldpc_mi355x.h"

namespace {

struct Matrix {
    uint32_t n = 0, m = 0;
    std::vector<int32_t> row_ptr, col;
};

bool next_line(std::istream &f, std::string &line) {
    while (std::getline(f, line)) {
        const size_t s = line.find_first_not_of(" \t\r\n");
        if (s == std::string::npos || line[s] == '#') continue;
        line = line.substr(s);
        return true;
    }
    return false;
}

// .chinn.out: "row col val" triplets (1-based), last line = dimensions;
// rows keep file order (ldpc_sim_scalar.cpp:116-202).
// .chinn: "N M [dv dc]" header, then per row "deg v1 .. vdeg" (:204-268).
Matrix parse_matrix(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot read the matrix " + path);
    std::string line;
    if (!next_line(f, line)) throw std::runtime_error("the matrix file " + path + " holds no lines");
    std::vector<std::string> tok;
    {
        std::stringstream ss(line);
        std::string t;
        while (ss >> t) tok.push_back(t);
    }
    Matrix M;
    if (tok.size() == 3) {
        std::vector<uint32_t> R, C;
        std::vector<float> V;
        do {
            std::stringstream ss(line);
            uint32_t r, c;
            float v;
            if (ss >> r >> c >> v) {
                R.push_back(r);
                C.push_back(c);
                V.push_back(v);
            }
        } while (next_line(f, line));
        M.m = R.back();
        M.n = C.back();
        std::vector<std::vector<int32_t>> rows(M.m);
        for (size_t i = 0; i + 1 < R.size(); ++i) {
            if (std::fabs(V[i]) <= 1e-6f) continue;
            if (R[i] == 0 || R[i] > M.m || C[i] == 0 || C[i] > M.n)
                throw std::runtime_error("an entry of " + path + " lies outside the matrix");
            rows[R[i] - 1].push_back((int32_t)(C[i] - 1));
        }
        M.row_ptr.push_back(0);
        for (auto &r : rows) {
            M.col.insert(M.col.end(), r.begin(), r.end());
            M.row_ptr.push_back((int32_t)M.col.size());
        }
    } else {
        std::stringstream hs(line);
        uint32_t dv = 0, dc = 0;
        hs >> M.n >> M.m;
        if (!(hs >> dv >> dc)) {
            if (!next_line(f, line)) throw std::runtime_error("no degree line in the .chinn file " + path);
        }
        M.row_ptr.push_back(0);
        for (uint32_t r = 0; r < M.m; ++r) {
            if (!next_line(f, line)) throw std::runtime_error("the .chinn file ends before check row " + std::to_string(r));
            std::stringstream rs(line);
            uint32_t deg = 0, v = 0;
            rs >> deg;
            for (uint32_t d = 0; d < deg; ++d) {
                rs >> v;
                M.col.push_back((int32_t)(v > 0 ? v - 1 : 0));
            }
            M.row_ptr.push_back((int32_t)M.col.size());
        }
    }
    if (M.n > 16384 || M.m > 8192)
        throw std::runtime_error("matrix too large for the engine limits (N <= 16384, M <= 8192): N = " + std::to_string(M.n) +
                                 ", M = " + std::to_string(M.m));
    return M;
}

void die(ldpc_ctx *ctx, int rc) {
    std::cerr << "libldpc error " << rc << ": " << ldpc_last_error(ctx) << std::endl;
    std::exit(1);
}

}  // namespace

int main(int argc, char **argv) {
    std::string chinn;
    double ebn0 = 2.0;
    uint32_t P = 0;
    uint64_t target_fe = 50, blocks_per_core = 100000, max_blocks = 0;
    uint32_t batch_per_core = 1000, max_iter = 16;
    float l_max = 0, r_max = 0, l_post = 0, gamma = 0;
    bool single = false, lockstep = false, nms = false;
    int opt;
    try {
        while ((opt = getopt(argc, argv, "c:e:p:t:m:b:k:i:svnL:R:G:V:")) != -1) {
            switch (opt) {
                case 'c': chinn = optarg; break;
                case 'e': ebn0 = std::stod(optarg); break;
                case 'p': P = (uint32_t)std::stoul(optarg); break;
                case 't': target_fe = std::stoull(optarg); break;
                case 'm': max_blocks = std::stoull(optarg); break;
                case 'b': batch_per_core = (uint32_t)std::stoul(optarg); break;
                case 'k': blocks_per_core = std::stoull(optarg); break;
                case 'i': max_iter = (uint32_t)std::stoul(optarg); break;
                case 's': single = true; break;
                case 'v': lockstep = true; break;
                case 'n': nms = true; break;
                case 'L': l_max = std::stof(optarg); break;
                case 'R': r_max = std::stof(optarg); break;
                case 'V': l_post = std::stof(optarg); break;
                case 'G': gamma = std::stof(optarg); break;
                default: break;
            }
        }
        const Matrix H = parse_matrix(chinn);
        uint32_t maxd = 0;
        for (uint32_t r = 0; r < H.m; ++r) maxd = std::max<uint32_t>(maxd, H.row_ptr[r + 1] - H.row_ptr[r]);
        std::cout << "[CONFIG] Matrix loaded: " << chinn << " (N=" << H.n << ", M=" << H.m
                  << ", max_check_deg=" << maxd << ")" << std::endl;
        if (maxd > 12) throw std::runtime_error("check degree " + std::to_string(maxd) + " is above the engine's limit of 12");

        const uint64_t cores = single ? 1 : 440;
        if (max_blocks == 0 && blocks_per_core > 0) max_blocks = blocks_per_core * cores;
        if (blocks_per_core > 0 && batch_per_core > blocks_per_core) batch_per_core = (uint32_t)blocks_per_core;
        std::cout << "[CONFIG] MI355X (libldpc Profile " << (nms ? "R_NMS" : "R") << "), block accounting for "
                  << cores << " cores; batch " << (uint64_t)batch_per_core * cores << " codewords" << std::endl;
        std::cout << "[CONFIG] Decoder algorithm: "
                  << (nms ? "Normalized Min-Sum (alpha=0.75)" : "Approximate-Min* (Christopher Jones MILCOM 2003)")
                  << std::endl;

        ldpc_code_desc code{};
        code.family = LDPC_CODE_MATRIX;
        code.n = H.n;
        code.m = H.m;
        code.n_punct = P;
        code.row_ptr = H.row_ptr.data();
        code.col = H.col.data();
        ldpc_decoder_cfg cfg{};
        cfg.profile = nms ? LDPC_PROFILE_R_NMS : LDPC_PROFILE_R;
        cfg.max_iter = max_iter;
        cfg.l_max = l_max;
        cfg.r_max = r_max;
        cfg.l_post_max = l_post;
        cfg.gamma = gamma;
        ldpc_ctx *ctx = nullptr;
        int rc = ldpc_create(&code, &cfg, 0, &ctx);
        if (rc) die(nullptr, rc);

        uint64_t blocks = 0, be = 0, fe = 0, next = 0;
        const uint64_t seed = 133742;  // the engine's root seed (ldpc_sim_scalar.cpp:405)
        const auto t0 = std::chrono::high_resolution_clock::now();
        while (fe < target_fe && blocks < max_blocks) {
            const uint64_t remaining = max_blocks - blocks;
            uint64_t batch = (uint64_t)batch_per_core * cores;
            if (remaining < batch) batch = std::max<uint64_t>(1, remaining / cores) * cores;
            batch = std::min(batch, remaining);
            ldpc_point_cfg pt{ebn0, seed, next, batch, 0, batch};
            ldpc_point_result r{};
            if ((rc = ldpc_run_point(ctx, &pt, &r, nullptr))) die(ctx, rc);
            if (lockstep) {
                ldpc_point_result r2{};
                if ((rc = ldpc_run_point(ctx, &pt, &r2, nullptr))) die(ctx, rc);
                if (r2.blocks != r.blocks || r2.bit_errors != r.bit_errors || r2.frame_errors != r.frame_errors ||
                    r2.iter_sum != r.iter_sum) {
                    std::cerr << "[VERIFICATION FAIL] repeated batch differs" << std::endl;
                    return 1;
                }
                std::cout << "[VERIFICATION PASS] repeated batch identical: bit_err=" << r.bit_errors
                          << ", frame_err=" << r.frame_errors << std::endl;
            }
            blocks += r.blocks;
            be += r.bit_errors;
            fe += r.frame_errors;
            next = r.next_cw;
            const double el =
                std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
            const double ntx = (double)(H.n - P);
            std::cout << "{\"progress\": true, \"ebn0\": " << ebn0 << ", \"blocks\": " << blocks
                      << ", \"max_blocks\": " << max_blocks << ", \"bit_errors\": " << be
                      << ", \"frame_errors\": " << fe << ", \"target_errors\": " << target_fe
                      << ", \"ber\": " << (blocks ? be / (blocks * ntx) : 0.0)
                      << ", \"fer\": " << (blocks ? (double)fe / blocks : 0.0)
                      << ", \"throughput_msps\": " << (el > 0 ? blocks * ntx / el / 1e6 : 0.0)
                      << ", \"elapsed_sec\": " << el << "}" << std::endl;
        }
        const double el = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
        std::cout << "{\"ebn0\": " << ebn0 << ", \"blocks\": " << blocks << ", \"bit_errors\": " << be
                  << ", \"frame_errors\": " << fe << ", \"elapsed_sec\": " << el << "}" << std::endl;
        ldpc_destroy(ctx);
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
