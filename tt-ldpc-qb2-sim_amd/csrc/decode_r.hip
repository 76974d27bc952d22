// decode_r.hip -- Profile R decoder (the scalar_5w engine behind
// run_simulation.py) for gfx950.
//
// Bit-exact restatement of engines/scalar_5w/ldpc_decoder_core.h:
//   q = L - R (fp32, unrounded)                                   (:422-436)
//   sign parity and min / argmin, first index wins ties          (:424-440)
//   recursive Approx-Min* with fp32 constants                    (:176-201)
//     (the degree-3 variant :155-173 is the same function: for finite
//      floats a > b <=> a - b > 0, so "(m > c) ? m - c : 0" and
//      "max(m - c, 0)" agree)
//   or normalized min-sum, alpha = 0.75 (USE_NORMALIZED_MIN_SUM) (:320-411)
//   optional clamps |R| <= r_max, |L'| <= l_post_max             (:490-506)
//   L' = bf16(q + r), R' = bf16(r)                                (:507-508)
//   syndrome on the raw sign bit, from the third iteration on    (:526-553)
//   frame error <=> a transmitted sign bit set; bit errors count them
//                                                                 (:559-577)
// Rows are decoded in file order, sequential Gauss-Seidel.  The host splits
// that order into ASAP levels: rows of a level share no variable, and every
// row's earlier neighbours sit in earlier levels, so processing level by
// level reads exactly the values the sequential order reads.  A level is cut
// into passes of up to 64 rows, one row per lane.
//
// Layout: one codeword per wave64; the wave's variable LLRs and per-edge
// messages (bf16) live in LDS; the pass table (one entry per lane and pass:
// degree, edge offset, variable list) is read from global memory and stays
// in the caches.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.h"
#include "kernels.h"

#pragma clang fp contract(off)

namespace ldpc {
namespace {

constexpr int WAVES = 2;  // 2 x 20 KiB of LDS per workgroup for r1/2 K=1024

__device__ __forceinline__ float bf(uint16_t h) { return u2f((uint32_t)h << 16); }
__device__ __forceinline__ float absbits(float x) { return u2f(f2u(x) & 0x7FFFFFFFu); }
__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// Correction of the recursive min* (ldpc_decoder_core.h:162):
// 0 for diff >= 2; 0.69315 - 0.40 diff below 0.5; 0.49315 - 0.328 (diff - 0.5)
__device__ __forceinline__ float corr_r(float diff) {
    const float l1 = __fsub_rn(0.69315f, __fmul_rn(0.40f, diff));
    const float l2 = __fsub_rn(0.49315f, __fmul_rn(0.328f, __fsub_rn(diff, 0.5f)));
    const float c = diff < 0.5f ? l1 : l2;
    return diff >= 2.0f ? 0.0f : c;
}

// One fold step: max(min(m, a) - g(|m - a|), 0).  m, a >= +0 and the
// difference is never -0, so fmaxf agrees with the reference's select.
__device__ __forceinline__ float fold_r(float m, float a) {
    const float diff = fabsf(__fsub_rn(m, a));
    const float cur = fminf(m, a);
    return fmaxf(__fsub_rn(cur, corr_r(diff)), 0.0f);
}

template <int MAXD>
struct Row {
    uint32_t meta;  // degree | edge_start << 4 (degree 0: idle lane)
    uint16_t v[MAXD];
};

template <int MAXD>
__device__ __forceinline__ Row<MAXD> load_row(const uint4 *table, int pass, int lane) {
    Row<MAXD> r;
    const uint4 a = table[(size_t)(pass * 64 + lane) * (MAXD / 6)];
    r.meta = a.x;
    r.v[0] = (uint16_t)a.y;
    r.v[1] = (uint16_t)(a.y >> 16);
    r.v[2] = (uint16_t)a.z;
    r.v[3] = (uint16_t)(a.z >> 16);
    r.v[4] = (uint16_t)a.w;
    r.v[5] = (uint16_t)(a.w >> 16);
    if constexpr (MAXD == 12) {
        const uint4 b = table[(size_t)(pass * 64 + lane) * 2 + 1];
        r.v[6] = (uint16_t)b.x;
        r.v[7] = (uint16_t)(b.x >> 16);
        r.v[8] = (uint16_t)b.y;
        r.v[9] = (uint16_t)(b.y >> 16);
        r.v[10] = (uint16_t)b.z;
        r.v[11] = (uint16_t)(b.z >> 16);
    }
    return r;
}

struct RCfg {
    float r_max, l_post_max;
};

// One row on one lane (inactive lanes have deg == 0 and do nothing).
template <int MAXD, bool NMS>
__device__ __forceinline__ void row_update(uint16_t *Ls, uint16_t *Rs, const Row<MAXD> &row, const RCfg &cfg) {
    const int deg = (int)(row.meta & 15u);
    const int es = (int)(row.meta >> 4);
    if (deg == 0) return;
    float q[MAXD];
    uint32_t sign = 0;
    int idx = 0;
    float r_min, r_all;
    if constexpr (NMS) {
        float min1 = 999.0f, min2 = 999.0f;
#pragma unroll
        for (int d = 0; d < MAXD; ++d) {
            if (d < deg) {
                q[d] = __fsub_rn(bf(Ls[row.v[d]]), bf(Rs[es + d]));
                sign ^= f2u(q[d]) >> 31;
                const float a = absbits(q[d]);
                if (a < min1) {
                    min2 = min1;
                    min1 = a;
                    idx = d;
                } else if (a < min2) {
                    min2 = a;
                }
            }
        }
        r_min = __fmul_rn(0.75f, min2);
        r_all = __fmul_rn(0.75f, min1);
    } else {
        float min_mag = 9999.0f;
#pragma unroll
        for (int d = 0; d < MAXD; ++d) {
            if (d < deg) {
                q[d] = __fsub_rn(bf(Ls[row.v[d]]), bf(Rs[es + d]));
                sign ^= f2u(q[d]) >> 31;
                const float a = absbits(q[d]);
                if (a < min_mag) {
                    min_mag = a;
                    idx = d;
                }
            }
        }
        // fold the non-minimum edges in index order (ldpc_decoder_core.h:180-190)
        const int first = idx == 0 ? 1 : 0;
        float m = deg > 1 ? absbits(q[first]) : min_mag;
#pragma unroll
        for (int d = 1; d < MAXD; ++d)
            if (d > first && d < deg && d != idx) m = fold_r(m, absbits(q[d]));
        r_min = m;
        r_all = fold_r(m, min_mag);
    }
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
        if (d < deg) {
            float mag = d == idx ? r_min : r_all;
            if (cfg.r_max > 0.0f && mag > cfg.r_max) mag = cfg.r_max;
            const uint32_t s = sign ^ (f2u(q[d]) >> 31);
            const float rn = u2f(f2u(mag) | (s << 31));
            float ln = __fadd_rn(q[d], rn);
            if (cfg.l_post_max > 0.0f) {
                if (ln > cfg.l_post_max)
                    ln = cfg.l_post_max;
                else if (ln < -cfg.l_post_max)
                    ln = -cfg.l_post_max;
            }
            Ls[row.v[d]] = (uint16_t)(rne_bits(ln) >> 16);
            Rs[es + d] = (uint16_t)(rne_bits(rn) >> 16);
        }
    }
}

// Syndrome with the hard decision on the raw sign bit; stops at the first
// pass that holds an unsatisfied row.
template <int MAXD>
__device__ __forceinline__ bool syndrome_ok(const uint16_t *Ls, const uint4 *table, int n_pass, int lane) {
    for (int p = 0; p < n_pass; ++p) {
        const Row<MAXD> row = load_row<MAXD>(table, p, lane);
        const int deg = (int)(row.meta & 15u);
        uint32_t par = 0;
#pragma unroll
        for (int d = 0; d < MAXD; ++d)
            if (d < deg) par ^= (uint32_t)Ls[row.v[d]] >> 15;
        if (__ballot(par & 1u)) return false;
    }
    return true;
}

struct RArgs {
    const uint4 *table;
    int n_pass, n, n_tx, n_edges;
    int max_iter;
    float r_max, l_post_max, l_max;
};

template <int MAXD, bool NMS>
__device__ int decode_wave_r(uint16_t *Ls, uint16_t *Rs, int lane, const RArgs &a) {
    for (int e = lane; e < a.n_edges; e += 64) Rs[e] = 0;
    wave_fence();
    const RCfg cfg{a.r_max, a.l_post_max};
    for (int it = 0; it < a.max_iter; ++it) {
        Row<MAXD> row = load_row<MAXD>(a.table, 0, lane);
        for (int p = 0; p < a.n_pass; ++p) {
            const Row<MAXD> cur = row;
            if (p + 1 < a.n_pass) row = load_row<MAXD>(a.table, p + 1, lane);  // prefetch
            row_update<MAXD, NMS>(Ls, Rs, cur, cfg);
            wave_fence();
        }
        if (it >= 2 && syndrome_ok<MAXD>(Ls, a.table, a.n_pass, lane)) return it + 1;
    }
    return 0;
}

// Channel LLRs (all-zero codeword) of codeword w into Ls with the clamp
// |L| <= l_max before rounding (ldpc_decoder_core.h:266-275); same Philox /
// Box-Muller stream as the Profile T generator.
__device__ void wave_gen_channel_r(uint16_t *Ls, int lane, uint64_t seed, uint64_t w, float mu, float sigma,
                                   float l_max, int n_tx, int n) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const int nblk = (n_tx + 3) / 4;
    for (int blk = lane; blk < nblk; blk += 64) {
        const u32x4 x = philox4x32_10((uint32_t)w, (uint32_t)(w >> 32), (uint32_t)blk, 0u, k0, k1);
        float z[4];
        box_muller(x.x, x.y, z[0], z[1]);
        box_muller(x.z, x.w, z[2], z[3]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = 4 * blk + j;
            if (i < n_tx) {
                float l = __fadd_rn(mu, __fmul_rn(sigma, z[j]));
                if (l_max > 0.0f) {
                    if (l > l_max)
                        l = l_max;
                    else if (l < -l_max)
                        l = -l_max;
                }
                Ls[i] = (uint16_t)(rne_bits(l) >> 16);
            }
        }
    }
    for (int i = n_tx + lane; i < n; i += 64) Ls[i] = 0;
}

// Dynamic LDS per wave: n variables then n_edges messages (bf16), rounded
// up to 16 B.
__device__ __forceinline__ int wave_lds_halfs(int n, int n_edges) { return ((n + n_edges) + 7) & ~7; }

__device__ __forceinline__ uint32_t next_cw(uint32_t *queue, int lane) {
    uint32_t v = 0;
    if (lane == 0) v = atomicAdd(queue, 1u);
    return __builtin_amdgcn_readfirstlane(v);
}

template <int MAXD, bool NMS>
__global__ void __launch_bounds__(64 * WAVES)
    k_decode_llrs_r(RArgs a, const uint16_t *__restrict__ llr_in, uint16_t *__restrict__ llr_out,
                    uint32_t *__restrict__ iters, uint8_t *__restrict__ ok, uint32_t n_cw, uint32_t *queue) {
    extern __shared__ __attribute__((aligned(16))) uint16_t smem_r[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint16_t *Ls = smem_r + wv * wave_lds_halfs(a.n, a.n_edges);
    uint16_t *Rs = Ls + a.n;
    for (;;) {
        const uint32_t cw = next_cw(queue, lane);
        if (cw >= n_cw) break;
        const uint16_t *src = llr_in + (size_t)cw * a.n;
        for (int i = lane; i < a.n; i += 64) Ls[i] = src[i];
        wave_fence();
        const int it = decode_wave_r<MAXD, NMS>(Ls, Rs, lane, a);
        wave_fence();
        uint16_t *dst = llr_out + (size_t)cw * a.n;
        for (int i = lane; i < a.n; i += 64) dst[i] = Ls[i];
        if (lane == 0) {
            iters[cw] = it ? (uint32_t)it : (uint32_t)a.max_iter;
            ok[cw] = it ? 1 : 0;
        }
        wave_fence();
    }
}

template <int MAXD, bool NMS>
__global__ void __launch_bounds__(64 * WAVES)
    k_monte_carlo_r(RArgs a, uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, uint32_t *queue,
                    unsigned long long *counters, unsigned long long *hist, Capture cap) {
    extern __shared__ __attribute__((aligned(16))) uint16_t smem_r[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint16_t *Ls = smem_r + wv * wave_lds_halfs(a.n, a.n_edges);
    uint16_t *Rs = Ls + a.n;
    uint64_t blocks = 0, fe = 0, be = 0, its = 0;
    for (;;) {
        const uint32_t cw = next_cw(queue, lane);
        if (cw >= n_cw) break;
        wave_gen_channel_r(Ls, lane, seed, first_cw + cw, mu, sigma, a.l_max, a.n_tx, a.n);
        wave_fence();
        const int it = decode_wave_r<MAXD, NMS>(Ls, Rs, lane, a);
        wave_fence();
        uint32_t b = 0;
        for (int i0 = 0; i0 < a.n_tx; i0 += 64) {
            const int i = i0 + lane;
            b += (uint32_t)__popcll(__ballot(i < a.n_tx && (Ls[i] >> 15)));
        }
        blocks += 1;
        its += it ? (uint64_t)it : (uint64_t)a.max_iter;
        be += b;
        fe += b > 0;
        if (lane == 0) {
            atomicAdd(hist + it, 1ull);
            if (b > 0) capture_failure(cap, first_cw + cw);
        }
        wave_fence();
    }
    if (lane == 0 && blocks) {
        atomicAdd(counters + 0, (unsigned long long)blocks);
        if (fe) atomicAdd(counters + 1, (unsigned long long)fe);
        if (be) atomicAdd(counters + 2, (unsigned long long)be);
        atomicAdd(counters + 3, (unsigned long long)its);
    }
}

template <int MAXD, bool NMS>
hipError_t launch_decode_t(const RLaunch &L, const uint16_t *llr_in, uint16_t *llr_out, uint32_t *iters, uint8_t *ok,
                           uint32_t n_cw, hipStream_t st) {
    RArgs a{(const uint4 *)L.table, L.n_pass, L.n, L.n_tx, L.n_edges, L.max_iter, L.r_max, L.l_post_max, L.l_max};
    const size_t lds = (size_t)WAVES * 2 * (((L.n + L.n_edges) + 7) & ~7);
    hipLaunchKernelGGL((k_decode_llrs_r<MAXD, NMS>), dim3(L.grid), dim3(64 * WAVES), lds, st, a, llr_in, llr_out, iters,
                       ok, n_cw, L.queue);
    return hipGetLastError();
}

template <int MAXD, bool NMS>
hipError_t launch_mc_t(const RLaunch &L, uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma,
                       unsigned long long *counters, unsigned long long *hist, hipStream_t st) {
    RArgs a{(const uint4 *)L.table, L.n_pass, L.n, L.n_tx, L.n_edges, L.max_iter, L.r_max, L.l_post_max, L.l_max};
    const size_t lds = (size_t)WAVES * 2 * (((L.n + L.n_edges) + 7) & ~7);
    hipLaunchKernelGGL((k_monte_carlo_r<MAXD, NMS>), dim3(L.grid), dim3(64 * WAVES), lds, st, a, seed, first_cw, n_cw,
                       mu, sigma, L.queue, counters, hist, L.cap);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256)
    k_gen_channel_r(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, float l_max, int n_tx,
                    int n, uint16_t *__restrict__ out) {
    // one thread per Philox block of 4 symbols; punctured tail zeroed by
    // the threads of block index >= ceil(n_tx / 4)
    const int nblk = (n_tx + 3) / 4, nz = (n - 4 * nblk + 3) / 4;
    const int per_cw = nblk + (nz > 0 ? nz : 0);
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (uint64_t)n_cw * per_cw) return;
    const uint32_t c = (uint32_t)(gid / per_cw);
    const int blk = (int)(gid % per_cw);
    uint16_t *L = out + (size_t)c * n;
    if (blk >= nblk) {
        for (int i = 4 * blk; i < 4 * blk + 4 && i < n; ++i) L[i] = 0;
        return;
    }
    const uint64_t w = first_cw + c;
    const u32x4 x = philox4x32_10((uint32_t)w, (uint32_t)(w >> 32), (uint32_t)blk, 0u, (uint32_t)seed,
                                  (uint32_t)(seed >> 32));
    float z[4];
    box_muller(x.x, x.y, z[0], z[1]);
    box_muller(x.z, x.w, z[2], z[3]);
    for (int j = 0; j < 4; ++j) {
        const int i = 4 * blk + j;
        if (i >= n) break;
        if (i >= n_tx) {
            L[i] = 0;
            continue;
        }
        float l = __fadd_rn(mu, __fmul_rn(sigma, z[j]));
        if (l_max > 0.0f) {
            if (l > l_max)
                l = l_max;
            else if (l < -l_max)
                l = -l_max;
        }
        L[i] = (uint16_t)(rne_bits(l) >> 16);
    }
}

}  // namespace

hipError_t r_launch_gen(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, float l_max, int n_tx,
                        int n, uint16_t *out, hipStream_t st) {
    const int nblk = (n_tx + 3) / 4, nz = (n - 4 * nblk + 3) / 4;
    const uint64_t threads = (uint64_t)n_cw * (nblk + (nz > 0 ? nz : 0));
    if (!threads) return hipSuccess;
    hipLaunchKernelGGL(k_gen_channel_r, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, seed, first_cw,
                       n_cw, mu, sigma, l_max, n_tx, n, out);
    return hipGetLastError();
}

int r_waves_per_wg() { return WAVES; }

size_t r_lds_bytes(int n, int n_edges) { return (size_t)WAVES * 2 * (((n + n_edges) + 7) & ~7); }

hipError_t r_occupancy(int maxd, bool nms, size_t lds, int *wg_per_cu) {
#define OCC(M, N) \
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, k_monte_carlo_r<M, N>, 64 * WAVES, lds)
    if (maxd <= 6) {
        if (nms) OCC(6, true);
        OCC(6, false);
    }
    if (nms) OCC(12, true);
    OCC(12, false);
#undef OCC
}

hipError_t r_launch_decode(const RLaunch &L, const uint16_t *llr_in, uint16_t *llr_out, uint32_t *iters, uint8_t *ok,
                           uint32_t n_cw, hipStream_t st) {
    if (L.maxd <= 6)
        return L.nms ? launch_decode_t<6, true>(L, llr_in, llr_out, iters, ok, n_cw, st)
                     : launch_decode_t<6, false>(L, llr_in, llr_out, iters, ok, n_cw, st);
    return L.nms ? launch_decode_t<12, true>(L, llr_in, llr_out, iters, ok, n_cw, st)
                 : launch_decode_t<12, false>(L, llr_in, llr_out, iters, ok, n_cw, st);
}

hipError_t r_launch_mc(const RLaunch &L, uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma,
                       unsigned long long *counters, unsigned long long *hist, hipStream_t st) {
    if (L.maxd <= 6)
        return L.nms ? launch_mc_t<6, true>(L, seed, first_cw, n_cw, mu, sigma, counters, hist, st)
                     : launch_mc_t<6, false>(L, seed, first_cw, n_cw, mu, sigma, counters, hist, st);
    return L.nms ? launch_mc_t<12, true>(L, seed, first_cw, n_cw, mu, sigma, counters, hist, st)
                 : launch_mc_t<12, false>(L, seed, first_cw, n_cw, mu, sigma, counters, hist, st);
}

}  // namespace ldpc
