// kernels.h -- host-side launchers of the HIP kernels (internal to libldpc).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ldpc {

// Trapping-set capture: Monte Carlo kernels append the id of every codeword
// that fails to ids[] (the first `max` of them, in completion order); the
// channel and the decode of a captured id are reproduced from (seed, id).
struct Capture {
    uint32_t *count;
    unsigned long long *ids;
    uint32_t max;  // 0: off
};

__device__ __forceinline__ void capture_failure(const Capture &c, unsigned long long id) {
    if (c.max) {
        const uint32_t s = atomicAdd(c.count, 1u);
        if (s < c.max) c.ids[s] = id;
    }
}

// Profile T, AR4JA r1/2 K=1024 (decode_t.hip)
int t12_waves_per_wg();
hipError_t t12_occupancy(int *wg_per_cu);
hipError_t t12_occupancy_decode(int *wg_per_cu);
hipError_t t12_launch_decode(const uint16_t *llr_in, uint16_t *llr_out, uint32_t *iters, uint8_t *ok, uint32_t n_cw,
                             int max_iter, float beta, uint32_t *queue, int grid, hipStream_t st);
// cw_gen: null for the all-zero codeword; else random-info-bit codewords with
// the quasi-cyclic generator words of the 8 transmitted parity blocks in
// device memory ([8][8][4] uint32, g_{r,i} of ldpc_get_generator for r < 8).
hipError_t t12_launch_mc(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, int max_iter,
                         float beta, uint32_t *queue, unsigned long long *counters, unsigned long long *hist, int grid,
                         const Capture &cap, const uint32_t *cw_gen, hipStream_t st);
hipError_t t12_launch_gen_cw(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, uint16_t *out,
                             const uint32_t *cw_gen, hipStream_t st);
int t12_cw_gen_words();
hipError_t t12_launch_gen(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, uint16_t *out,
                          hipStream_t st);
int t12_schedule(int32_t *out, int cap);
// Tail counts of the channel normals (|z| > thr[k], k < n_thr <= NOISE_TAIL_MAX)
// of codewords first_cw .. first_cw + n_cw - 1 with n_tx symbols each; adds to counts.
constexpr int NOISE_TAIL_MAX = 16;
hipError_t launch_noise_tails(uint64_t seed, uint64_t first_cw, uint64_t n_cw, int n_tx, const float *thr, int n_thr,
                              unsigned long long *counts, int cus, hipStream_t st);

// Profile R: list-scheduled rows in passes.  Two kernels:
//  * decode_r.hip: one codeword per wave, one row per lane and pass of up to
//    R_PASS_ROWS rows, messages in LDS (any degree <= 12);
//  * decode_r2.hip: two codewords per wave (one per 32-lane half), passes of
//    up to R2_PASS_ROWS rows, messages compressed in VGPRs: degree <= 6 and at
//    most R2_MAX_PASSES passes.  r2_mode >= 0 selects it; the r_launch_*
//    entry points dispatch.
constexpr int R_PASS_ROWS = 64;
constexpr int R2_PASS_ROWS = 32;
constexpr int R2_MAX_PASSES = 64;
// pass kinds of the two-codeword kernel
constexpr uint32_t R2_PK_NONE = 0;    // past the last pass
constexpr uint32_t R2_PK_EXACT3 = 3;  // every row of the pass has degree 3
constexpr uint32_t R2_PK_EXACT6 = 6;  // every row of the pass has degree 6
constexpr uint32_t R2_PK_MASKED = 7;  // rows of degree 1..6, masked datapath
struct RLaunch {
    const void *table;  // n_pass * R_PASS_ROWS entries of 16 B (maxd <= 6) or 32 B (maxd <= 12)
    int n_pass, n, n_tx, max_iter;
    int n_edges;  // message slots in LDS: the code's edges plus MAXD - maxd spare
    float r_max, l_post_max, l_max;
    int maxd;
    bool nms;
    int grid;
    uint32_t *queue;
    Capture cap;
    // two-codeword kernel
    int r2_mode = -1;                  // -1: not used; 0: passes of kind EXACT3 / EXACT6; 1: MASKED passes
    const void *table2 = nullptr;      // (4 n_groups + 1) * 64 entries of 16 B
    const uint32_t *kinds2 = nullptr;  // n_groups words, byte j of word g = kind of pass 4 g + j
    const uint16_t *slot2 = nullptr;   // LDS slot of every variable (transmitted ones below n_tx), padded to 4
    int n_groups = 0;                  // ceil(passes / 4)
    int hs = 0;                        // bytes of LDS per half (the LLRs, rounded up to 16)
};
int r_waves_per_wg();
size_t r_lds_bytes(int n, int n_edges);
// LDS per workgroup of the kernel L selects
size_t r_launch_lds(const RLaunch &L);
hipError_t r_occupancy(const RLaunch &L, size_t lds, int *wg_per_cu);
hipError_t r_launch_decode(const RLaunch &L, const uint16_t *llr_in, uint16_t *llr_out, uint32_t *iters, uint8_t *ok,
                           uint32_t n_cw, hipStream_t st);
hipError_t r_launch_gen(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, float l_max, int n_tx,
                        int n, uint16_t *out, hipStream_t st);
hipError_t r_launch_mc(const RLaunch &L, uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma,
                       unsigned long long *counters, unsigned long long *hist, hipStream_t st);
size_t r2_lds_bytes(const RLaunch &L);
hipError_t r2_occupancy(const RLaunch &L, int *wg_per_cu);
hipError_t r2_launch_decode(const RLaunch &L, const uint16_t *llr_in, uint16_t *llr_out, uint32_t *iters, uint8_t *ok,
                            uint32_t n_cw, hipStream_t st);
hipError_t r2_launch_mc(const RLaunch &L, uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma,
                        unsigned long long *counters, unsigned long long *hist, hipStream_t st);

// Profile T for the nine AR4JA codes (decode_multi.hip, one object per code).
// Per-launch values a captured graph changes between replays, read by the
// kernel from device memory when MultiParams::dyn is set.
struct MultiDyn {
    uint64_t seed, first_cw;
    uint32_t n_cw;
    float mu, sigma;
    uint32_t reserved;
};
struct MultiParams {
    const uint16_t *llr_in;  // decode mode: n_cw x N bf16 input; null for Monte Carlo
    uint16_t *llr_out;       // decode mode output
    uint32_t *iters;
    uint8_t *ok;
    uint32_t *queue;               // one u32, zero at launch
    unsigned long long *counters;  // blocks, frame errors, bit errors, iteration sum
    unsigned long long *hist;      // max_iter + 1 bins
    const MultiDyn *dyn;
    uint64_t seed, first_cw;
    uint32_t n_cw;
    int max_iter;
    float mu, sigma, beta, beta3;
    // Profile D (learned min-sum): scale / offset rows [n_groups][P], P = sum
    // of the layer degrees; iteration t uses row min(t, n_groups - 1)
    const float *d_alpha, *d_beta;
    int d_groups, d_corr;
    float d_rmax;
    Capture cap;    // Monte Carlo mode only
    uint32_t *work;  // per-workgroup workspace (MultiVariant::work_per_wg bytes each), when the kernel needs one
};
struct MultiVariant {
    int kid, rid, c, order;  // code (K = 1024 << 2 kid, rate 1/2, 2/3, 4/5), checks per thread, layer order
    int threads, cw_per_wg;  // workgroup size, codewords in flight per workgroup
    hipError_t (*launch)(const MultiParams &p, int grid, hipStream_t st);
    hipError_t (*occupancy)(int *wg_per_cu);
    hipError_t (*attributes)(hipFuncAttributes *a);
    int profile_d;  // 1: the learned min-sum kernel (Profile D), 0: Profile T
    size_t work_per_wg = 0;  // bytes of HBM workspace per workgroup of the launch grid (0: none)
};
const MultiVariant *multi_table_k0_r0(int *n);
const MultiVariant *multi_table_k0_r1(int *n);
const MultiVariant *multi_table_k0_r2(int *n);
const MultiVariant *multi_table_k1_r0(int *n);
const MultiVariant *multi_table_k1_r1(int *n);
const MultiVariant *multi_table_k1_r2(int *n);
const MultiVariant *multi_table_k2_r0(int *n);
const MultiVariant *multi_table_k2_r1(int *n);
const MultiVariant *multi_table_k2_r2(int *n);
// Profile D kernel of a code
const MultiVariant *multi_table_d_k0_r0();
const MultiVariant *multi_table_d_k0_r1();
const MultiVariant *multi_table_d_k0_r2();
const MultiVariant *multi_table_d_k1_r0();
const MultiVariant *multi_table_d_k1_r1();
const MultiVariant *multi_table_d_k1_r2();
const MultiVariant *multi_table_d_k2_r0();
const MultiVariant *multi_table_d_k2_r1();
const MultiVariant *multi_table_d_k2_r2();

hipError_t launch_fill_u64(unsigned long long *p, size_t n, unsigned long long v, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t st);

}  // namespace ldpc
