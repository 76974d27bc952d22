// decode_r2.hip -- Profile R decoder (the scalar_5w engine behind
// run_simulation.py) with two codewords per wave and the row messages in
// VGPRs, for gfx950.
//
// Arithmetic: the same bit-exact restatement as decode_r.hip of
// engines/scalar_5w/ldpc_decoder_core.h (q = L - R in fp32; sign parity;
// min / argmin with the first index winning ties; recursive Approx-Min* with
// fp32 constants, profile_r.h, or normalized min-sum with alpha = 0.75;
// optional |R| <= r_max and |L'| <= l_post_max clamps; L' = bf16(q + r),
// R' = bf16(r); syndrome on the raw sign bit from the third iteration on;
// :155-201, 320-411, 413-586).  Rows are decoded in file order, sequential
// Gauss-Seidel: the host packs that order into passes of up to 32 rows whose
// rows touch disjoint variables (a row joins a pass only once every earlier
// row sharing a variable with it sits in an earlier pass), so every row reads
// exactly the values the sequential order reads.
//
// Mapping:
//  * A wave decodes two codewords, one per 32-lane half; lane l of half h
//    runs row l of every pass on codeword h.  A chain of c dependent rows
//    takes c passes whatever their width, so with 64-row passes (decode_r.hip)
//    lanes idle while the chain drains; passes of 32 fill.
//  * Where every row has degree 3 or 6, each pass holds rows of one degree
//    (kind PK_EXACT3 / PK_EXACT6) and runs that degree's datapath with no
//    masking; other matrices run masked passes of up to 6 edges (PK_MASKED).
//  * A lane owns one row per pass, so its rows' messages stay in VGPRs across
//    iterations, compressed to 39 bits a row: the bf16 magnitudes of the two
//    distinct outgoing messages (the argmin edge's and the others'), the
//    argmin and the six signs.  Pass p = 4 g + j keeps them in P_j[g] and in
//    bits 7 j .. 7 j + 6 of Mt[g] (five 16-dword vectors, 80 VGPRs, up to 64
//    passes), read and written at the uniform index g through s_set_gpr_idx.
//    LDS holds only the two codewords' LLRs, each variable in the slot the
//    host chose so that the lanes of a gather fall in distinct banks.
//  * Each half converges on its own: after an iteration a half whose codeword
//    passed the syndrome check, or ran out of iterations, is tallied and
//    refilled from the work queue while the other half keeps decoding.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.h"
#include "kernels.h"
#include "profile_r.h"

#pragma clang fp contract(off)

namespace ldpc {
namespace {

using pr::absbits;
using pr::fold_r;
using pr::RCfg;

typedef uint32_t v16u __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

struct R2Args {
    const uint4 *table;      // (4 n_groups + 1) * 64 entries: degree, then six 16-bit LDS byte offsets
    const uint32_t *kinds;   // n_groups words; byte j of word g = kind of pass 4 g + j
    const uint16_t *slot;    // LDS slot of every variable; transmitted ones below n_tx
    int n_groups, n, n_tx, max_iter, hs;
    float r_max, l_post_max, l_max;
};

struct R2Work {
    // Monte Carlo
    uint64_t seed, first_cw;
    float mu, sigma;
    unsigned long long *counters, *hist;
    Capture cap;
    // decode of given LLRs
    const uint16_t *llr_in;
    uint16_t *llr_out;
    uint32_t *iters;
    uint8_t *ok;
    // both
    uint32_t n_cw;
    uint32_t *queue;
};

// One row on one lane.  e: the lane's table entry (degree in bits 0..3, 0 on
// idle lanes; LDS byte offsets of the row's variables in 16-bit fields, the
// half's slot 0 past the degree and on idle lanes: read, never used).  pair / meta: the
// row's compressed messages, updated in place:
//   pair: bits 0..14 bf16(r_min), bit 15 s_0, bits 16..30 bf16(r_all), bit 31 s_1
//   meta: bits 0..2 argmin, bit d + 1 s_d for d >= 2
// where r_min goes to the argmin edge, r_all to the others, and s_d is the
// sign of edge d's message.  All-zero words are the all-zero messages.
template <int D, bool MASKED, bool NMS, bool CLAMP>
__device__ __forceinline__ void row2(uint8_t *lds, const uint4 &e, uint32_t &pair, uint32_t &meta, const RCfg &cfg) {
    const int deg = (int)(e.x & 15u);
    const uint32_t wv[3] = {e.y, e.z, e.w};
    uint32_t off[D], lh[D];
#pragma unroll
    for (int d = 0; d < D; ++d) off[d] = (d & 1) ? (wv[d / 2] >> 16) : (wv[d / 2] & 0xFFFFu);
#pragma unroll
    for (int d = 0; d < D; ++d) lh[d] = *(const uint16_t *)(lds + off[d]);
    __builtin_amdgcn_sched_barrier(0);  // keep the D reads in one batch
    // the old messages R'_d = (d == argmin ? bf16(r_min) : bf16(r_all)) | s_d
    const uint32_t t = pair << 16;
    const uint32_t rmin_o = t & 0x7FFF0000u, rall_o = pair & 0x7FFF0000u;
    const uint32_t idx_o = meta & 7u;
    float q[D];
    uint32_t sign = 0;  // bit 31: parity of the signs of q over the row
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const uint32_t sw = d == 0 ? t : d == 1 ? pair : (meta << (30 - d));
        const uint32_t r = ((uint32_t)d == idx_o ? rmin_o : rall_o) | (sw & 0x80000000u);
        q[d] = __fsub_rn(u2f(lh[d] << 16), u2f(r));
        sign ^= (!MASKED || d < deg) ? f2u(q[d]) : 0u;
    }
    // c[d]: edge d lowers the running minimum (strict <, so the first index
    // wins ties); the argmin is the last such edge, or edge 0 when none does.
    // The per-edge facts below are lane masks, combined by scalar ops.
    bool c[D];
    float r_min, r_all, min_mag;
    float a[D];
#pragma unroll
    for (int d = 0; d < D; ++d) a[d] = absbits(q[d]);
    if constexpr (NMS) {
        float min1 = 999.0f, min2 = 999.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            c[d] = (!MASKED || d < deg) && a[d] < min1;
            const bool c2 = (!MASKED || d < deg) && !c[d] && a[d] < min2;
            min2 = c[d] ? min1 : (c2 ? a[d] : min2);
            min1 = c[d] ? a[d] : min1;
        }
        min_mag = min1;
        r_min = __fmul_rn(0.75f, min2);
        r_all = __fmul_rn(0.75f, min1);
    } else {
        min_mag = 9999.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            c[d] = (!MASKED || d < deg) && a[d] < min_mag;
            min_mag = c[d] ? a[d] : min_mag;
        }
    }
    bool later[D];  // later[d]: some edge after d lowers the minimum, i.e. d < argmin
    later[D - 1] = false;
#pragma unroll
    for (int d = D - 2; d >= 0; --d) later[d] = later[d + 1] || c[d + 1];
    bool is_min[D];
    is_min[0] = !later[0];
#pragma unroll
    for (int d = 1; d < D; ++d) is_min[d] = c[d] && !later[d];
    uint32_t idx = 0;
#pragma unroll
    for (int d = 1; d < D; ++d) idx = is_min[d] ? (uint32_t)d : idx;
    if constexpr (!NMS) {
        if constexpr (!MASKED) {
            // the non-minimum edges in index order (ldpc_decoder_core.h:180-190)
            float o[D - 1];
#pragma unroll
            for (int k = 0; k < D - 1; ++k) o[k] = later[k] ? a[k] : a[k + 1];
            float m = o[0];
#pragma unroll
            for (int k = 1; k < D - 1; ++k) m = fold_r(m, o[k]);
            r_min = m;
        } else {
            const bool first1 = is_min[0];  // the fold starts at edge 1 when edge 0 is the argmin
            float m = deg > 1 ? (first1 ? a[1] : a[0]) : min_mag;
#pragma unroll
            for (int d = 1; d < D; ++d) {
                const float f = fold_r(m, a[d]);
                m = ((d > 1 || !first1) && d < deg && !is_min[d]) ? f : m;
            }
            r_min = m;
        }
        r_all = fold_r(r_min, min_mag);
    }
    if (CLAMP && cfg.r_max > 0.0f) {
        r_min = r_min > cfg.r_max ? cfg.r_max : r_min;
        r_all = r_all > cfg.r_max ? cfg.r_max : r_all;
    }
    // new messages, compressed (r_min, r_all >= +0: their sign bits are free)
    uint32_t s[D];
#pragma unroll
    for (int d = 0; d < D; ++d) s[d] = (sign ^ f2u(q[d])) & 0x80000000u;
    uint32_t mt = (uint32_t)idx;
#pragma unroll
    for (int d = 2; d < D; ++d) mt |= s[d] >> (30 - d);
    pair = cvt_pk_bf16(r_min, r_all) | (s[0] >> 16) | s[1];
    meta = mt;
    // L' = bf16(q + r), r the unrounded message with its sign
    const uint32_t rmin_b = f2u(r_min), rall_b = f2u(r_all);
    float ln[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        ln[d] = __fadd_rn(q[d], u2f((is_min[d] ? rmin_b : rall_b) | s[d]));
        if (CLAMP && cfg.l_post_max > 0.0f) ln[d] = __builtin_amdgcn_fmed3f(ln[d], -cfg.l_post_max, cfg.l_post_max);
    }
    uint32_t lw[(D + 1) / 2];
#pragma unroll
    for (int k = 0; k < (D + 1) / 2; ++k) lw[k] = cvt_pk_bf16(ln[2 * k], ln[2 * k + 1 < D ? 2 * k + 1 : 2 * k]);
    if constexpr (MASKED) {
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (d < deg) *(uint16_t *)(lds + off[d]) = (uint16_t)((d & 1) ? (lw[d / 2] >> 16) : lw[d / 2]);
    } else if (deg) {
#pragma unroll
        for (int d = 0; d < D; ++d) *(uint16_t *)(lds + off[d]) = (uint16_t)((d & 1) ? (lw[d / 2] >> 16) : lw[d / 2]);
    }
}

// Parity of a row on the raw sign bits; 0 on idle lanes.  MASKED: the
// slots past the degree are left out.
template <int D, bool MASKED>
__device__ __forceinline__ uint32_t parity2(const uint8_t *lds, const uint4 &e) {
    const int deg = (int)(e.x & 15u);
    const uint32_t wv[3] = {e.y, e.z, e.w};
    uint32_t h[D];
#pragma unroll
    for (int d = 0; d < D; ++d)
        h[d] = *(const uint16_t *)(lds + ((d & 1) ? (wv[d / 2] >> 16) : (wv[d / 2] & 0xFFFFu)));
    __builtin_amdgcn_sched_barrier(0);
    uint32_t par = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) par ^= (!MASKED || d < deg) ? h[d] : 0u;
    return deg ? (par >> 15) & 1u : 0u;
}

// MODE 0: every pass is PK_EXACT3 or PK_EXACT6 (or PK_NONE past the last);
// MODE 1: every pass is PK_MASKED (or PK_NONE).
template <int MODE, int J, bool NMS, bool CLAMP>
__device__ __forceinline__ void pass_step(uint8_t *lds, const uint4 &e, uint32_t kw, v16u &PJ, uint32_t &mw, int g,
                                          const RCfg &cfg) {
    const uint32_t kind = (kw >> (8 * J)) & 0xFFu;
    if (kind == R2_PK_NONE) return;
    uint32_t pr = PJ[g];
    uint32_t mt = (mw >> (7 * J)) & 0x7Fu;
    if constexpr (MODE == 0) {
        if (kind == R2_PK_EXACT6)
            row2<6, false, NMS, CLAMP>(lds, e, pr, mt, cfg);
        else
            row2<3, false, NMS, CLAMP>(lds, e, pr, mt, cfg);
    } else {
        row2<6, true, NMS, CLAMP>(lds, e, pr, mt, cfg);
    }
    PJ[g] = pr;
    mw = (mw & ~(0x7Fu << (7 * J))) | (mt << (7 * J));
    wave_fence();
}

// Which of the halves in chk have every row satisfied.
template <int MODE>
__device__ uint32_t syndrome2(const uint8_t *lds, const R2Args &a, int lane, uint32_t chk) {
    uint32_t bad = 0;  // halves with an unsatisfied row
    uint4 ea = a.table[lane], eb;
    for (int g = 0; g < a.n_groups; ++g) {
        const uint32_t kw = a.kinds[g];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint4 &e = (j & 1) ? eb : ea;
            uint4 &nx = (j & 1) ? ea : eb;
            nx = a.table[(size_t)(4 * g + j + 1) * 64 + lane];
            const uint32_t kind = (kw >> (8 * j)) & 0xFFu;
            if (kind == R2_PK_NONE) continue;
            const uint32_t par = MODE == 1 ? parity2<6, true>(lds, e)
                                 : kind == R2_PK_EXACT3 ? parity2<3, false>(lds, e) : parity2<6, false>(lds, e);
            const uint64_t b = __ballot(par);
            bad |= ((uint32_t)b ? 1u : 0u) | ((uint32_t)(b >> 32) ? 2u : 0u);
            if ((bad & chk) == chk) return 0;
        }
    }
    return chk & ~bad;
}

// Channel LLRs (all-zero codeword) of codeword w into one half's LLRs Lt,
// with the clamp |L| <= l_max before rounding (ldpc_decoder_core.h:266-275):
// the stream of decode_r.hip, generated by gs lanes (gl = 0 .. gs - 1);
// variable i goes to LDS slot slot[i].  Zeroes the punctured variables'
// slots (n_tx .. n - 1).
__device__ void gen_channel_r2(uint16_t *Lt, const uint16_t *__restrict__ slot, int gl, int gs, uint64_t seed,
                               uint64_t w, float mu, float sigma, float l_max, int n_tx, int n) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const int nblk = (n_tx + 3) / 4;
    for (int blk = gl; blk < nblk; blk += gs) {
        const uint2 sl = *(const uint2 *)(slot + 4 * blk);  // slots of variables 4 blk .. 4 blk + 3
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
                const uint32_t sw = j < 2 ? sl.x : sl.y;
                Lt[(j & 1) ? (sw >> 16) : (sw & 0xFFFFu)] = (uint16_t)(rne_bits(l) >> 16);
            }
        }
    }
    for (int i = n_tx + gl; i < n; i += gs) Lt[i] = 0;
}

template <int MODE, bool NMS, bool CLAMP, bool MC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_r2(R2Args a, R2Work w) {
    extern __shared__ __attribute__((aligned(16))) uint8_t lds_r2[];
    const int lane = threadIdx.x;
    const int h = lane >> 5, hl = lane & 31;
    uint16_t *Lh = (uint16_t *)(lds_r2 + h * a.hs);
    const RCfg cfg{a.r_max, a.l_post_max};
    v16u P0 = 0, P1 = 0, P2 = 0, P3 = 0, Mt = 0;
    uint32_t cw0 = 0, cw1 = 0;  // codeword of each half (uniform)
    int it0 = 0, it1 = 0;       // iterations completed by each half
    uint32_t act = 0;           // bit h: half h holds a codeword
    uint32_t need = 3u;         // bit h: half h takes the next codeword
    uint32_t blocks = 0, fe = 0;
    uint64_t be = 0, its = 0;
    for (;;) {
        if (need) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(w.queue, (uint32_t)__builtin_popcount(need));
            base = __builtin_amdgcn_readfirstlane(base);
            uint32_t fill = 0;
            if (need & 1u) {
                cw0 = base++;
                it0 = 0;
                fill |= cw0 < w.n_cw ? 1u : 0u;
            }
            if (need & 2u) {
                cw1 = base;
                it1 = 0;
                fill |= cw1 < w.n_cw ? 2u : 0u;
            }
            act = (act & ~need) | fill;
            if (fill) {
                // both halves refilled: each half loads its own codeword; one
                // half refilled: all 64 lanes load it
                const bool both = fill == 3u;
                const uint32_t th = both ? (uint32_t)h : fill >> 1;  // the half this lane loads
                const int gl = both ? hl : lane, gs = both ? 32 : 64;
                const uint32_t my = th ? cw1 : cw0;
                uint16_t *Lt = (uint16_t *)(lds_r2 + th * a.hs);
                if constexpr (MC) {
                    gen_channel_r2(Lt, a.slot, gl, gs, w.seed, w.first_cw + my, w.mu, w.sigma, a.l_max, a.n_tx, a.n);
                } else {
                    const uint16_t *src = w.llr_in + (size_t)my * a.n;
                    for (int i = gl; i < a.n; i += gs) Lt[a.slot[i]] = src[i];
                }
            }
            if ((fill >> h) & 1u) {
                P0 = 0;
                P1 = 0;
                P2 = 0;
                P3 = 0;
                Mt = 0;
            }
            need = 0;
            wave_fence();
        }
        if (!act) break;
        // one iteration of both halves; pass p + 1's entry is in flight while pass p runs
        {
            uint4 ea = a.table[lane], eb;
            for (int g = 0; g < a.n_groups; ++g) {
                const uint32_t kw = a.kinds[g];
                uint32_t mw = Mt[g];
                const size_t t0 = (size_t)(4 * g) * 64 + lane;
                eb = a.table[t0 + 64];
                pass_step<MODE, 0, NMS, CLAMP>(lds_r2, ea, kw, P0, mw, g, cfg);
                ea = a.table[t0 + 128];
                pass_step<MODE, 1, NMS, CLAMP>(lds_r2, eb, kw, P1, mw, g, cfg);
                eb = a.table[t0 + 192];
                pass_step<MODE, 2, NMS, CLAMP>(lds_r2, ea, kw, P2, mw, g, cfg);
                ea = a.table[t0 + 256];
                pass_step<MODE, 3, NMS, CLAMP>(lds_r2, eb, kw, P3, mw, g, cfg);
                Mt[g] = mw;
            }
        }
        it0 += (int)(act & 1u);
        it1 += (int)((act >> 1) & 1u);
        const uint32_t chk = ((act & 1u) && it0 >= 3 ? 1u : 0u) | ((act & 2u) && it1 >= 3 ? 2u : 0u);
        const uint32_t conv = chk ? syndrome2<MODE>(lds_r2, a, lane, chk) : 0u;
        const uint32_t done =
            conv | ((act & 1u) && it0 >= a.max_iter ? 1u : 0u) | ((act & 2u) && it1 >= a.max_iter ? 2u : 0u);
        if (done) {
            if ((done >> h) & 1u) {
                const uint32_t my = h ? cw1 : cw0;
                const int it = ((conv >> h) & 1u) ? (h ? it1 : it0) : 0;  // 0: not converged
                if constexpr (MC) {
                    uint32_t b = 0;  // transmitted variables sit in slots 0 .. n_tx - 1
                    for (int i = hl; i < a.n_tx; i += 32) b += (uint32_t)(Lh[i] >> 15);
#pragma unroll
                    for (int o = 16; o; o >>= 1) b += (uint32_t)__shfl_xor((int)b, o);
                    if (hl == 0) {
                        blocks += 1;
                        fe += b > 0;
                        be += b;
                        its += it ? (uint64_t)it : (uint64_t)a.max_iter;
                        atomicAdd(w.hist + it, 1ull);
                        if (b > 0) capture_failure(w.cap, w.first_cw + my);
                    }
                } else {
                    uint16_t *dst = w.llr_out + (size_t)my * a.n;
                    for (int i = hl; i < a.n; i += 32) dst[i] = Lh[a.slot[i]];
                    if (hl == 0) {
                        w.iters[my] = it ? (uint32_t)it : (uint32_t)a.max_iter;
                        w.ok[my] = it ? 1 : 0;
                    }
                }
            }
            need = done;
            wave_fence();
        }
    }
    if constexpr (MC) {
        if (hl == 0 && blocks) {
            atomicAdd(w.counters + 0, (unsigned long long)blocks);
            if (fe) atomicAdd(w.counters + 1, (unsigned long long)fe);
            if (be) atomicAdd(w.counters + 2, (unsigned long long)be);
            atomicAdd(w.counters + 3, (unsigned long long)its);
        }
    }
}

R2Args args_of(const RLaunch &L) {
    return R2Args{(const uint4 *)L.table2, L.kinds2, L.slot2,    L.n_groups, L.n,          L.n_tx,
                  L.max_iter,            L.hs,     L.r_max, L.l_post_max, L.l_max};
}

template <int MODE, bool NMS, bool CLAMP, bool MC>
hipError_t launch_t(const RLaunch &L, const R2Work &w, hipStream_t st) {
    hipLaunchKernelGGL((k_r2<MODE, NMS, CLAMP, MC>), dim3(L.grid), dim3(64), (size_t)2 * L.hs, st, args_of(L), w);
    return hipGetLastError();
}

template <int MODE, bool NMS, bool CLAMP>
hipError_t occ_t(size_t lds, int *wg_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, k_r2<MODE, NMS, CLAMP, true>, 64, lds);
}

#define R2_DISPATCH(L, CALL)                                                        \
    do {                                                                            \
        const bool c_ = (L).r_max > 0.0f || (L).l_post_max > 0.0f;                  \
        if ((L).r2_mode == 0) {                                                     \
            if ((L).nms) return c_ ? CALL(0, true, true) : CALL(0, true, false);    \
            return c_ ? CALL(0, false, true) : CALL(0, false, false);               \
        }                                                                           \
        if ((L).nms) return c_ ? CALL(1, true, true) : CALL(1, true, false);        \
        return c_ ? CALL(1, false, true) : CALL(1, false, false);                   \
    } while (0)

}  // namespace

size_t r2_lds_bytes(const RLaunch &L) { return (size_t)2 * L.hs; }

hipError_t r2_occupancy(const RLaunch &L, int *wg_per_cu) {
    const size_t lds = r2_lds_bytes(L);
#define OCC(M, N, C) occ_t<M, N, C>(lds, wg_per_cu)
    R2_DISPATCH(L, OCC);
#undef OCC
}

hipError_t r2_launch_decode(const RLaunch &L, const uint16_t *llr_in, uint16_t *llr_out, uint32_t *iters, uint8_t *ok,
                            uint32_t n_cw, hipStream_t st) {
    R2Work w{};
    w.llr_in = llr_in;
    w.llr_out = llr_out;
    w.iters = iters;
    w.ok = ok;
    w.n_cw = n_cw;
    w.queue = L.queue;
#define DEC(M, N, C) launch_t<M, N, C, false>(L, w, st)
    R2_DISPATCH(L, DEC);
#undef DEC
}

hipError_t r2_launch_mc(const RLaunch &L, uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma,
                        unsigned long long *counters, unsigned long long *hist, hipStream_t st) {
    R2Work w{};
    w.seed = seed;
    w.first_cw = first_cw;
    w.mu = mu;
    w.sigma = sigma;
    w.counters = counters;
    w.hist = hist;
    w.cap = L.cap;
    w.n_cw = n_cw;
    w.queue = L.queue;
#define MC(M, N, C) launch_t<M, N, C, true>(L, w, st)
    R2_DISPATCH(L, MC);
#undef MC
}

}  // namespace ldpc
