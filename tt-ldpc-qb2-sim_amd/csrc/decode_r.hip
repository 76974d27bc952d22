// decode_r.hip -- Profile R decoder (the scalar_5w engine behind
// run_simulation.py) for gfx950.
//
// Bit-exact restatement of engines/scalar_5w/ldpc_decoder_core.h:
//   q = L - R (fp32, unrounded)                                   (:422-436)
//   sign parity and min / argmin, first index wins ties          (:424-440)
//   recursive Approx-Min* with fp32 constants (profile_r.h)      (:176-201)
//   or normalized min-sum, alpha = 0.75 (USE_NORMALIZED_MIN_SUM) (:320-411)
//   optional clamps |R| <= r_max, |L'| <= l_post_max             (:490-506)
//   L' = bf16(q + r), R' = bf16(r)                                (:507-508)
//   syndrome on the raw sign bit, from the third iteration on    (:526-553)
//   frame error <=> a transmitted sign bit set; bit errors count them
//                                                                 (:559-577)
// Rows are decoded in file order, sequential Gauss-Seidel.  The host packs
// that order into passes of up to 64 rows, one row per lane: a row joins a
// pass only once every earlier row sharing a variable with it sits in an
// earlier pass, so the rows of a pass touch disjoint variables and each
// reads exactly the values the sequential order reads (37 passes for the
// r1/2 K=1024 matrix, the length of its dependency chain).
//
// Layout: one codeword per wave64; the wave's variable LLRs and per-edge
// messages (bf16) live in LDS; the pass table (one entry per lane and pass:
// degree, edge offset, variable list) is read from global memory and stays
// in the caches.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.h"
#include "kernels.h"
#include "profile_r.h"

#pragma clang fp contract(off)

namespace ldpc {
namespace {

constexpr int WAVES = 2;  // 2 x 20 KiB of LDS per workgroup for r1/2 K=1024

using pr::absbits;
using pr::fold_r;
using pr::RCfg;
__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// One lane's entry of the pass table as loaded: word 0 = degree |
// edge_start << 4, then the variable indices as 16-bit pairs (4 words per
// entry for MAXD 6, 8 for MAXD 12).  Slots past
// the degree (and idle lanes, degree 0) hold variable 0: a valid address
// whose value is read and then masked out, so every gather of a row can be
// issued before the first one is used.
template <int MAXD>
struct RowRaw {
    uint4 w[MAXD / 6];
};

template <int MAXD>
__device__ __forceinline__ RowRaw<MAXD> load_row(const uint4 *table, int pass, int lane) {
    RowRaw<MAXD> r;
#pragma unroll
    for (int k = 0; k < MAXD / 6; ++k) r.w[k] = table[(size_t)(pass * R_PASS_ROWS + lane) * (MAXD / 6) + k];
    return r;
}

// Variable d sits in 32-bit word 1 + d / 2 of the entry (words 1..3 for
// d < 6, words 4..6 for 6 <= d < 12), low half for even d.
template <int MAXD>
__device__ __forceinline__ int var_of(const RowRaw<MAXD> &r, int d) {
    const int k = 1 + d / 2;
    const uint4 &w = r.w[k / 4];
    const uint32_t word = (k % 4) == 0 ? w.x : (k % 4) == 1 ? w.y : (k % 4) == 2 ? w.z : w.w;
    return (int)((d & 1) ? (word >> 16) : (word & 0xFFFFu));
}

// One row on one lane.  Straight-line code: all 2 * MAXD gathers first,
// then the arithmetic with slots d >= deg masked by selects, then the
// scatters of the deg real slots.  Idle lanes (deg == 0) compute on
// variable 0 / edge 0 and write nothing.  The host lays the messages out so
// that es + MAXD - 1 is inside the message array for every row (a row of
// the largest degree is stored last, plus MAXD - maxd spare slots), so the
// message reads of slots past the degree need no address select.
// CLAMP: the optional |R| <= r_max and |L'| <= l_post_max clamps.
template <int MAXD, bool NMS, bool CLAMP>
__device__ __forceinline__ void row_update(uint16_t *Ls, uint16_t *Rs, const RowRaw<MAXD> &row, const RCfg &cfg) {
    const uint32_t meta = row.w[0].x;
    const int deg = (int)(meta & 15u);
    const int es = (int)(meta >> 4);
    int v[MAXD];
    uint32_t lh[MAXD], rh[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d) v[d] = var_of<MAXD>(row, d);
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
        lh[d] = Ls[v[d]];
        rh[d] = Rs[es + d];
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the 2 * MAXD reads in one batch
    float q[MAXD];
    uint32_t sign = 0;  // bit 31: parity of the signs of q over the row
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
        q[d] = __fsub_rn(u2f(lh[d] << 16), u2f(rh[d] << 16));
        sign ^= d < deg ? f2u(q[d]) : 0u;
    }
    int idx = 0;
    float r_min, r_all;
    if constexpr (NMS) {
        float min1 = 999.0f, min2 = 999.0f;
#pragma unroll
        for (int d = 0; d < MAXD; ++d) {
            const float a = absbits(q[d]);
            const bool c1 = d < deg && a < min1;
            const bool c2 = d < deg && !c1 && a < min2;
            min2 = c1 ? min1 : (c2 ? a : min2);
            min1 = c1 ? a : min1;
            idx = c1 ? d : idx;
        }
        r_min = __fmul_rn(0.75f, min2);
        r_all = __fmul_rn(0.75f, min1);
    } else {
        float min_mag = 9999.0f;
#pragma unroll
        for (int d = 0; d < MAXD; ++d) {
            const float a = absbits(q[d]);
            const bool c = d < deg && a < min_mag;
            min_mag = c ? a : min_mag;
            idx = c ? d : idx;
        }
        // fold the non-minimum edges in index order (ldpc_decoder_core.h:180-190)
        const int first = idx == 0 ? 1 : 0;
        float m = deg > 1 ? absbits(first ? q[1] : q[0]) : min_mag;
#pragma unroll
        for (int d = 1; d < MAXD; ++d) {
            const float f = fold_r(m, absbits(q[d]));
            m = (d > first && d < deg && d != idx) ? f : m;
        }
        r_min = m;
        r_all = fold_r(m, min_mag);
    }
    if (CLAMP && cfg.r_max > 0.0f) {
        r_min = r_min > cfg.r_max ? cfg.r_max : r_min;
        r_all = r_all > cfg.r_max ? cfg.r_max : r_all;
    }
    // R' = bf16(rn) = sign | bf16(|rn|): two magnitudes per row
    const uint32_t rmin_b = f2u(r_min), rall_b = f2u(r_all);
    const uint32_t rmin_h = rne_bits(r_min), rall_h = rne_bits(r_all);
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
        const bool is_min = d == idx;
        const uint32_t s = (sign ^ f2u(q[d])) & 0x80000000u;
        const float rn = u2f((is_min ? rmin_b : rall_b) | s);
        float ln = __fadd_rn(q[d], rn);
        if (CLAMP && cfg.l_post_max > 0.0f) ln = __builtin_amdgcn_fmed3f(ln, -cfg.l_post_max, cfg.l_post_max);
        if (d < deg) {
            Ls[v[d]] = (uint16_t)(rne_bits(ln) >> 16);
            Rs[es + d] = (uint16_t)(((is_min ? rmin_h : rall_h) | s) >> 16);
        }
    }
}

// Parity of a row on the raw sign bits (slots past the degree masked).
template <int MAXD>
__device__ __forceinline__ uint32_t row_parity(const uint16_t *Ls, const RowRaw<MAXD> &row) {
    const int deg = (int)(row.w[0].x & 15u);
    uint32_t h[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d) h[d] = Ls[var_of<MAXD>(row, d)];
    __builtin_amdgcn_sched_barrier(0);
    uint32_t par = 0;
#pragma unroll
    for (int d = 0; d < MAXD; ++d) par ^= d < deg ? h[d] : 0u;
    return (par >> 15) & 1u;
}

// Syndrome with the hard decision on the raw sign bit; stops at the first
// pass that holds an unsatisfied row.  The table entry of the next pass is
// loaded while the current one is checked (two named buffers, no copies).
template <int MAXD>
__device__ __forceinline__ bool syndrome_ok(const uint16_t *Ls, const uint4 *table, int n_pass, int lane) {
    const int last = n_pass - 1;
    RowRaw<MAXD> x = load_row<MAXD>(table, 0, lane), y;
    for (int p = 0; p <= last; p += 2) {
        y = load_row<MAXD>(table, min(p + 1, last), lane);
        if (__ballot(row_parity<MAXD>(Ls, x))) return false;
        if (p + 1 > last) break;
        x = load_row<MAXD>(table, min(p + 2, last), lane);
        if (__ballot(row_parity<MAXD>(Ls, y))) return false;
    }
    return true;
}

struct RArgs {
    const uint4 *table;
    int n_pass, n, n_tx, n_edges;
    int max_iter;
    float r_max, l_post_max, l_max;
};

template <int MAXD, bool NMS, bool CLAMP>
__device__ int decode_wave_r(uint16_t *Ls, uint16_t *Rs, int lane, const RArgs &a) {
    for (int e = lane; e < a.n_edges; e += 64) Rs[e] = 0;
    wave_fence();
    const RCfg cfg{a.r_max, a.l_post_max};
    const int last = a.n_pass - 1;
    for (int it = 0; it < a.max_iter; ++it) {
        // pass p + 1's entry is in flight while pass p runs
        RowRaw<MAXD> x = load_row<MAXD>(a.table, 0, lane), y;
        for (int p = 0; p <= last; p += 2) {
            y = load_row<MAXD>(a.table, min(p + 1, last), lane);
            row_update<MAXD, NMS, CLAMP>(Ls, Rs, x, cfg);
            wave_fence();
            if (p + 1 > last) break;
            x = load_row<MAXD>(a.table, min(p + 2, last), lane);
            row_update<MAXD, NMS, CLAMP>(Ls, Rs, y, cfg);
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

template <int MAXD, bool NMS, bool CLAMP>
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
        const int it = decode_wave_r<MAXD, NMS, CLAMP>(Ls, Rs, lane, a);
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

template <int MAXD, bool NMS, bool CLAMP>
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
        const int it = decode_wave_r<MAXD, NMS, CLAMP>(Ls, Rs, lane, a);
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

template <int MAXD, bool NMS, bool CLAMP>
hipError_t launch_decode_t(const RLaunch &L, const uint16_t *llr_in, uint16_t *llr_out, uint32_t *iters, uint8_t *ok,
                           uint32_t n_cw, hipStream_t st) {
    RArgs a{(const uint4 *)L.table, L.n_pass, L.n, L.n_tx, L.n_edges, L.max_iter, L.r_max, L.l_post_max, L.l_max};
    const size_t lds = (size_t)WAVES * 2 * (((L.n + L.n_edges) + 7) & ~7);
    hipLaunchKernelGGL((k_decode_llrs_r<MAXD, NMS, CLAMP>), dim3(L.grid), dim3(64 * WAVES), lds, st, a, llr_in, llr_out, iters,
                       ok, n_cw, L.queue);
    return hipGetLastError();
}

template <int MAXD, bool NMS, bool CLAMP>
hipError_t launch_mc_t(const RLaunch &L, uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma,
                       unsigned long long *counters, unsigned long long *hist, hipStream_t st) {
    RArgs a{(const uint4 *)L.table, L.n_pass, L.n, L.n_tx, L.n_edges, L.max_iter, L.r_max, L.l_post_max, L.l_max};
    const size_t lds = (size_t)WAVES * 2 * (((L.n + L.n_edges) + 7) & ~7);
    hipLaunchKernelGGL((k_monte_carlo_r<MAXD, NMS, CLAMP>), dim3(L.grid), dim3(64 * WAVES), lds, st, a, seed, first_cw, n_cw,
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

template <int MAXD, bool NMS, bool CLAMP>
hipError_t occ_t(size_t lds, int *wg_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, k_monte_carlo_r<MAXD, NMS, CLAMP>, 64 * WAVES, lds);
}

// Returns CALL(MAXD, NMS, CLAMP) for the launch's configuration.
#define R_DISPATCH(L, CALL)                                             \
    do {                                                                \
        const bool c_ = (L).r_max > 0.0f || (L).l_post_max > 0.0f;      \
        if ((L).maxd <= 6) {                                            \
            if ((L).nms) return c_ ? CALL(6, true, true) : CALL(6, true, false);    \
            return c_ ? CALL(6, false, true) : CALL(6, false, false);               \
        }                                                               \
        if ((L).nms) return c_ ? CALL(12, true, true) : CALL(12, true, false);      \
        return c_ ? CALL(12, false, true) : CALL(12, false, false);                 \
    } while (0)

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

size_t r_launch_lds(const RLaunch &L) { return L.r2_mode >= 0 ? r2_lds_bytes(L) : r_lds_bytes(L.n, L.n_edges); }

hipError_t r_occupancy(const RLaunch &L, size_t lds, int *wg_per_cu) {
    if (L.r2_mode >= 0) return r2_occupancy(L, wg_per_cu);
#define OCC(M, N, C) occ_t<M, N, C>(lds, wg_per_cu)
    R_DISPATCH(L, OCC);
#undef OCC
}

hipError_t r_launch_decode(const RLaunch &L, const uint16_t *llr_in, uint16_t *llr_out, uint32_t *iters, uint8_t *ok,
                           uint32_t n_cw, hipStream_t st) {
    if (L.r2_mode >= 0) return r2_launch_decode(L, llr_in, llr_out, iters, ok, n_cw, st);
#define DEC(M, N, C) launch_decode_t<M, N, C>(L, llr_in, llr_out, iters, ok, n_cw, st)
    R_DISPATCH(L, DEC);
#undef DEC
}

hipError_t r_launch_mc(const RLaunch &L, uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma,
                       unsigned long long *counters, unsigned long long *hist, hipStream_t st) {
    if (L.r2_mode >= 0) return r2_launch_mc(L, seed, first_cw, n_cw, mu, sigma, counters, hist, st);
#define MC(M, N, C) launch_mc_t<M, N, C>(L, seed, first_cw, n_cw, mu, sigma, counters, hist, st)
    R_DISPATCH(L, MC);
#undef MC
}

}  // namespace ldpc
