// decode_t.hip -- Profile T layered Approximate-Min* decoder for gfx950.
//
// Bit-exact restatement of the reference's golden layered decoder
// (tests/test_circulant_dma_stepper.cpp:107-327 check math, :743-815 schedule,
// syndrome, latch and tallies) for AR4JA rate-1/2 K=1024, mapped to CDNA4:
//
//  * one codeword per wave64; the wave's 2560 variable LLRs (bf16) live in
//    LDS (5 KiB), the 7680 check-to-variable messages R live in VGPRs as
//    packed bf16 pairs (60 dwords per lane);
//  * a layer is one quarter of a block row (128 checks, no shared variable);
//    lane l updates checks l and l+64 of the layer, so every edge slot of a
//    layer is one cyclic-shifted 128-variable read/modify/write in LDS;
//  * layers run in Gauss-Seidel order with the 12-layer schedule unrolled at
//    compile time; LDS operations of one wave execute in issue order, so no
//    barrier separates layers (only a compiler fence);
//  * the syndrome is computed after every iteration from 40 wave ballots of
//    the hard decisions (one 128-bit mask per variable quarter) rotated and
//    XOR-ed on the scalar unit;
//  * converged codewords leave the loop (convergence latch); a wave then
//    takes the next codeword id from a device work queue.
//
// Arithmetic per edge (fp32, every op separately rounded, RNE to bf16 with
// the reference's integer bias formula):
//   q   = RNE(L - R)                          (stepper:112-116)
//   r_d = tree min* over the other edges, * beta (stepper:151-225; :300-304)
//   out = RNE(copysign(r_d, s ^ sign(q_d)))   (stepper:178-186)
//   L'  = RNE(q + out),  R' = out              (stepper:179-180)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "ar4ja_schedule.h"
#include "device_math.h"
#include "kernels.h"

#pragma clang fp contract(off)

namespace ldpc {
namespace {

constexpr ar4ja::Schedule S12 = ar4ja::make_schedule(0, 0);  // r1/2, K = 1024
constexpr int Z = 128;
constexpr int NV = 2560;
constexpr int NTX = 2048;
constexpr int KINFO = 1024;
constexpr int NLAYERS = 12;
constexpr int WAVES_PER_WG = 4;
static_assert(S12.Z == Z && S12.N == NV && S12.N - S12.n_punct == NTX && S12.k_info == KINFO, "schedule");

constexpr int r_offset(int li) {
    int o = 0;
    for (int i = 0; i < li; ++i) o += S12.layer[i].deg;
    return o;
}
constexpr int NR = r_offset(NLAYERS);  // 60 packed dwords per lane
static_assert(NR == 60, "R register budget");

__device__ __forceinline__ float ld_bf(const uint16_t *p) { return u2f((uint32_t)(*p) << 16); }

// Compiler-only ordering point between layers.  LDS requests of one
// wavefront are performed in issue order by the hardware, so a store by one
// lane is seen by a later load of another lane of the same wave; the fence
// keeps the compiler from hoisting the next layer's loads above the stores.
__device__ __forceinline__ void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

template <int D>
__device__ __forceinline__ void check_tree(const float (&q)[D], float beta, float (&r)[D]) {
    if constexpr (D == 6) {
        float m0 = fabsf(q[0]), m1 = fabsf(q[1]), m2 = fabsf(q[2]);
        float m3 = fabsf(q[3]), m4 = fabsf(q[4]), m5 = fabsf(q[5]);
        float p03 = amin_t(m0, m3), p14 = amin_t(m1, m4), p25 = amin_t(m2, m5);
        float o0 = amin_t(p14, p25), o1 = amin_t(p03, p25), o2 = amin_t(p03, p14);
        r[0] = __fmul_rn(amin_t(o0, m3), beta);
        r[3] = __fmul_rn(amin_t(o0, m0), beta);
        r[1] = __fmul_rn(amin_t(o1, m4), beta);
        r[4] = __fmul_rn(amin_t(o1, m1), beta);
        r[2] = __fmul_rn(amin_t(o2, m5), beta);
        r[5] = __fmul_rn(amin_t(o2, m2), beta);
    } else {
        static_assert(D == 3, "Profile T defines degree 3 and 6 checks");
        float m0 = fabsf(q[0]), m1 = fabsf(q[1]), m2 = fabsf(q[2]);
        r[0] = __fmul_rn(amin_t(m1, m2), beta);
        r[1] = __fmul_rn(amin_t(m0, m2), beta);
        r[2] = __fmul_rn(amin_t(m0, m1), beta);
    }
}

// One layer: lane updates checks j = lane (a) and j = lane + 64 (b).  q, R'
// and L' of the two checks travel as packed bf16 pairs; the signs of R' are
// applied to the packed word (RNE commutes with the sign bit).
template <int LI>
__device__ __forceinline__ void run_layer(uint16_t *Ls, uint32_t (&R)[NR], int lane, float beta6, float beta3) {
    constexpr ar4ja::Layer Ly = S12.layer[LI];
    constexpr int D = Ly.deg;
    constexpr int R0 = r_offset(LI);
    const float beta = (D == 6) ? beta6 : beta3;
    // Opaque copy of the lane id: keeps the shifted LDS addresses from being
    // hoisted out of the iteration loop (they would pin ~60 VGPRs).
    asm volatile("" : "+v"(lane));
    int ia[D], ib[D];
    uint32_t qp[D];
    float qa[D], qb[D];
    uint32_t sp = 0;  // sign parity of both checks: bit 15 (a), bit 31 (b)
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int base = Ly.slot[d].base, sh = Ly.slot[d].shift;
        int xa = sh == 0 ? lane : ((lane + sh) & (Z - 1));
        int xb = sh == 0 ? lane + 64 : (xa ^ 64);
        ia[d] = base + xa;
        ib[d] = base + xb;
        const uint32_t rp = R[R0 + d];
        qp[d] = cvt_pk_bf16(__fsub_rn(ld_bf(Ls + ia[d]), lo_f(rp)), __fsub_rn(ld_bf(Ls + ib[d]), hi_f(rp)));
        qa[d] = lo_f(qp[d]);
        qb[d] = hi_f(qp[d]);
        sp ^= qp[d];
    }
    float ra[D], rb[D];
    check_tree<D>(qa, beta, ra);
    check_tree<D>(qb, beta, rb);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const uint32_t rn = cvt_pk_bf16(ra[d], rb[d]) | ((sp ^ qp[d]) & 0x80008000u);
        const uint32_t ln = cvt_pk_bf16(__fadd_rn(qa[d], lo_f(rn)), __fadd_rn(qb[d], hi_f(rn)));
        Ls[ia[d]] = (uint16_t)ln;
        Ls[ib[d]] = (uint16_t)(ln >> 16);
        R[R0 + d] = rn;
    }
    wave_lds_fence();
}

template <int... LI>
__device__ __forceinline__ void run_iteration(uint16_t *Ls, uint32_t (&R)[NR], int lane, float beta6, float beta3,
                                              std::integer_sequence<int, LI...>) {
    (run_layer<LI>(Ls, R, lane, beta6, beta3), ...);
}

// 128-bit rotate right by a compile-time amount: bit j of the result is bit
// (j + s) mod 128 of (hi:lo).
template <int s>
__device__ __forceinline__ void rotr128(uint64_t lo, uint64_t hi, uint64_t &olo, uint64_t &ohi) {
    if constexpr (s == 0) {
        olo = lo;
        ohi = hi;
    } else if constexpr (s < 64) {
        olo = (lo >> s) | (hi << (64 - s));
        ohi = (hi >> s) | (lo << (64 - s));
    } else if constexpr (s == 64) {
        olo = hi;
        ohi = lo;
    } else {
        olo = (hi >> (s - 64)) | (lo << (128 - s));
        ohi = (lo >> (s - 64)) | (hi << (128 - s));
    }
}

// Hard decision "bf16 < 0.0f" (stepper:781): sign set and magnitude nonzero.
__device__ __forceinline__ bool hard_neg(uint16_t h) { return h > 0x8000u; }

// Parity of the layer's 128 checks from the hard decisions in Ls.  Each edge
// slot contributes the 128-bit hard-decision mask of its variable quarter
// (two wave ballots), rotated by the slot's shift; the XOR over slots is the
// layer's syndrome vector.  Masks are formed per slot so that only a few
// 64-bit scalars are live at a time.
template <int LI, int... D>
__device__ __forceinline__ uint64_t layer_syndrome(const uint16_t *Ls, int lane, std::integer_sequence<int, D...>) {
    constexpr ar4ja::Layer Ly = S12.layer[LI];
    uint64_t plo = 0, phi = 0;
    (
        [&] {
            constexpr int base = Ly.slot[D].base;
            const uint64_t lo = __ballot(hard_neg(Ls[base + lane]));
            const uint64_t hi = __ballot(hard_neg(Ls[base + 64 + lane]));
            uint64_t a, b;
            rotr128<Ly.slot[D].shift>(lo, hi, a, b);
            plo ^= a;
            phi ^= b;
        }(),
        ...);
    return plo | phi;
}

// Layers are tested one at a time and the test stops at the first layer with
// an unsatisfied check, so an iteration that has not converged usually costs
// one layer's ballots.  The memory clobber keeps the compiler from caching
// ballots across layers (which would pin dozens of scalar registers).
template <int LI>
__device__ __forceinline__ bool syndrome_from(const uint16_t *Ls, int lane) {
    if constexpr (LI == NLAYERS) {
        return true;
    } else {
        asm volatile("" ::: "memory");
        if (layer_syndrome<LI>(Ls, lane, std::make_integer_sequence<int, S12.layer[LI].deg>{}) != 0) return false;
        return syndrome_from<LI + 1>(Ls, lane);
    }
}

__device__ __forceinline__ bool wave_syndrome_ok(const uint16_t *Ls, int lane) { return syndrome_from<0>(Ls, lane); }

// Decode the codeword held in Ls.  Returns the iteration of convergence
// (1..max_iter) or 0 if the syndrome never cleared within max_iter.
__device__ __forceinline__ int decode_wave(uint16_t *Ls, int lane, int max_iter, float beta6, float beta3) {
    uint32_t R[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) R[i] = 0u;  // bf16 +0: iteration 0 uses q = L
    wave_lds_fence();
    for (int it = 0; it < max_iter; ++it) {
        run_iteration(Ls, R, lane, beta6, beta3, std::make_integer_sequence<int, NLAYERS>{});
        if (wave_syndrome_ok(Ls, lane)) return it + 1;
    }
    return 0;
}

// Bit errors of an unconverged codeword: info bits k < K with bf16 <= 0.0f
// (stepper:801-809).
__device__ __forceinline__ uint32_t wave_info_errors(const uint16_t *Ls, int lane) {
    uint32_t be = 0;
#pragma unroll
    for (int t = 0; t < KINFO / 64; ++t) {
        uint16_t h = Ls[t * 64 + lane];
        be += (uint32_t)__popcll(__ballot((h & 0x8000u) || h == 0));
    }
    return be;
}

// Random-info-bit codewords (no reference counterpart: the reference sends
// the all-zero codeword, ldpc_decoder_core.h:258-297, stepper:744-757).  The
// info bits of codeword w are Philox(ctr = (w_lo, w_hi, j, 1), key = seed),
// j = 0..7, 128 bits per call (bit b of word d = info bit 128 j + 32 d + b).
// Parity: the code is quasi-cyclic with 128-bit blocks; transmitted parity
// block r (variables 1024 + 128 r + t) is p_r = sum_i g_{r,i} u_i in
// GF(2)[x] / (x^128 + 1), i.e. bit t = parity(g_{r,i} AND w_{i,t}) summed over
// the info blocks i, with w_{i,t}[s] = u_i[(t - s) mod 128].  g comes from the
// host's encoder: 8 x 8 x 4 words in device memory ([transmitted parity block
// r][info block i][128 bits]), read at the point of use.
constexpr int CW_GEN_WORDS = 8 * 8 * 4;

// Info bits of codeword w into ub[0..31]; the doubled bit-reversed blocks
// v_i[k] = u_i[(-k) mod 128] into scr[8 i .. 8 i + 7] (scr: LDS scratch).
__device__ __forceinline__ void wave_info_bits(uint32_t *ub, uint32_t *scr, int lane, uint32_t k0, uint32_t k1,
                                               uint64_t w) {
    if (lane < KINFO / 128) {
        const u32x4 x = philox4x32_10((uint32_t)w, (uint32_t)(w >> 32), (uint32_t)lane, 1u, k0, k1);
        *reinterpret_cast<uint4 *>(ub + 4 * lane) = make_uint4(x.x, x.y, x.z, x.w);
        // bit-reverse the 128 bits (b[k] = u[127 - k]), then rotate left by one (v[k] = b[k - 1])
        const uint32_t b0 = __brev(x.w), b1 = __brev(x.z), b2 = __brev(x.y), b3 = __brev(x.x);
        const uint4 v = make_uint4((b0 << 1) | (b3 >> 31), (b1 << 1) | (b0 >> 31), (b2 << 1) | (b1 >> 31),
                                   (b3 << 1) | (b2 >> 31));
        *reinterpret_cast<uint4 *>(scr + 8 * lane) = v;
        *reinterpret_cast<uint4 *>(scr + 8 * lane + 4) = v;
    }
}

// w_{i,t} = bits [o, o + 128) of the doubled v_i, o = (128 - t) mod 128.
__device__ __forceinline__ void info_window(const uint32_t *vv, int t, uint32_t (&w)[4]) {
    const int o = (Z - t) & (Z - 1), d = o >> 5;
    const uint32_t sh = (uint32_t)(o & 31);
    uint32_t x[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) x[k] = vv[d + k];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = __builtin_amdgcn_alignbit(x[k + 1], x[k], sh);
}

// Transmitted parity bits of the codeword into scr[64 .. 95] (bit p of the
// 1024 = variable 1024 + p).  Lane l computes positions l and l + 64 of
// every parity block.
struct Words4 {
    uint32_t x, y, z, w;
};
typedef __attribute__((address_space(4))) const Words4 const_words4;

__device__ __forceinline__ void wave_encode(uint32_t *scr, int lane, const uint32_t *__restrict__ gen) {
    // The table is read through the constant address space (scalar loads:
    // uniform, no VGPRs), from an opaque copy of its pointer so that the
    // loads stay inside the codeword loop, and one info block at a time
    // (scheduling barriers): kept live across the decode, or loaded all at
    // once, its 256 words spill registers and cut the occupancy.
    asm volatile("" : "+s"(gen));
    const_words4 *g4 = (const_words4 *)(const void *)gen;
    uint32_t acc[8][2];
#pragma unroll
    for (int r = 0; r < 8; ++r) acc[r][0] = acc[r][1] = 0u;
#pragma unroll
    for (int i = 0; i < KINFO / 128; ++i) {
        __builtin_amdgcn_sched_barrier(0);
        uint32_t wa[4], wb[4];
        info_window(scr + 8 * i, lane, wa);
        info_window(scr + 8 * i, lane + 64, wb);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const uint32_t gx = g4[r * 8 + i].x, gy = g4[r * 8 + i].y, gz = g4[r * 8 + i].z, gw = g4[r * 8 + i].w;
            acc[r][0] ^= (gx & wa[0]) ^ (gy & wa[1]) ^ (gz & wa[2]) ^ (gw & wa[3]);
            acc[r][1] ^= (gx & wb[0]) ^ (gy & wb[1]) ^ (gz & wb[2]) ^ (gw & wb[3]);
        }
    }
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint64_t m = __ballot(__popc(acc[r][h]) & 1u);
            if (lane == 0) *reinterpret_cast<uint2 *>(scr + 64 + 4 * r + 2 * h) = make_uint2((uint32_t)m, (uint32_t)(m >> 32));
        }
}

// Channel LLRs of codeword w into Ls: Philox(ctr = (w_lo, w_hi, blk, 0),
// key = seed) -> 4 normals -> RNE_bf16(+-mu + sigma z); punctured = +0.  The
// all-zero codeword sends +mu everywhere.  A random codeword (RANDOM) sends
// -mu where its bit is 1; its info bits stay in ub for the tally, and the
// punctured tail of Ls serves as scratch for the encoder until it is zeroed.
template <bool RANDOM>
__device__ __forceinline__ void wave_gen_channel(uint16_t *Ls, uint32_t *ub, int lane, uint64_t seed, uint64_t w,
                                                 float mu, float sigma, const uint32_t *__restrict__ gen) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    uint32_t *scr = reinterpret_cast<uint32_t *>(Ls + NTX);
    if constexpr (RANDOM) {
        wave_info_bits(ub, scr, lane, k0, k1, w);
        wave_lds_fence();
        wave_encode(scr, lane, gen);
        wave_lds_fence();
    }
#pragma unroll 2
    for (int t = 0; t < NTX / 4 / 64; ++t) {
        const uint32_t blk = (uint32_t)(t * 64 + lane);
        u32x4 x = philox4x32_10((uint32_t)w, (uint32_t)(w >> 32), blk, 0u, k0, k1);
        float z0, z1, z2, z3;
        box_muller(x.x, x.y, z0, z1);
        box_muller(x.z, x.w, z2, z3);
        float m0 = mu, m1 = mu, m2 = mu, m3 = mu;
        if constexpr (RANDOM) {
            const uint32_t v = 4 * blk;  // four consecutive variables, one nibble of the codeword bits
            const uint32_t cw = (v < KINFO ? ub[v >> 5] : scr[64 + ((v - KINFO) >> 5)]) >> (v & 31);
            m0 = (cw & 1u) ? -mu : mu;
            m1 = (cw & 2u) ? -mu : mu;
            m2 = (cw & 4u) ? -mu : mu;
            m3 = (cw & 8u) ? -mu : mu;
        }
        uint32_t lo = (llr_bits(m1, sigma, z1) & 0xFFFF0000u) | (llr_bits(m0, sigma, z0) >> 16);
        uint32_t hi = (llr_bits(m3, sigma, z3) & 0xFFFF0000u) | (llr_bits(m2, sigma, z2) >> 16);
        *reinterpret_cast<uint2 *>(Ls + 4 * blk) = make_uint2(lo, hi);
    }
    if constexpr (RANDOM) wave_lds_fence();
    // punctured variables NTX..NV-1: 512 x bf16 +0 = 64 lanes x 16 B
    *reinterpret_cast<uint4 *>(Ls + NTX + 8 * lane) = make_uint4(0u, 0u, 0u, 0u);
}

// Tally of a random codeword (info bits in ub): mism = some info-bit hard
// decision (bf16 < 0 -> 1, stepper:781) differs from the sent bit; errs = info
// bits not strictly on the sent side (bf16 +-0, or the sign differs), which
// for a sent 0 is the golden's bf16 <= 0.0f (stepper:801-809).
__device__ __forceinline__ void wave_cw_errors(const uint16_t *Ls, const uint32_t *ub, int lane, bool &mism,
                                               uint32_t &errs) {
    uint64_t any = 0;
    errs = 0;
#pragma unroll
    for (int t = 0; t < KINFO / 64; ++t) {
        const int k = t * 64 + lane;
        const uint16_t h = Ls[k];
        const uint32_t c = (ub[k >> 5] >> (k & 31)) & 1u;
        any |= __ballot((uint32_t)hard_neg(h) != c);
        errs += (uint32_t)__popcll(__ballot((h & 0x7FFFu) == 0 || (uint32_t)(h >> 15) != c));
    }
    mism = any != 0;
}

// Work distribution: every codeword of a launch comes from a device queue,
// one atomic per codeword.  Waves that drew codewords needing few iterations
// (or that run on a faster CU) take more of them.  A round-robin static
// share of 7/8 of the launch took 17% more time on the r1/2 K=1024 point
// (60.3 against 70.5 Gsps): the per-wave work it fixes in advance cannot
// follow the wave's real speed.
__device__ __forceinline__ uint32_t next_codeword(uint32_t *queue, int lane) {
    uint32_t v = 0;
    if (lane == 0) v = atomicAdd(queue, 1u);
    return __builtin_amdgcn_readfirstlane(v);
}

__global__ void __launch_bounds__(64 * WAVES_PER_WG)
    k_decode_llrs(const uint16_t *__restrict__ llr_in, uint16_t *__restrict__ llr_out, uint32_t *__restrict__ iters,
                  uint8_t *__restrict__ ok, uint32_t n_cw, int max_iter, float beta6, float beta3, uint32_t *queue) {
    __shared__ __attribute__((aligned(16))) uint16_t smem[WAVES_PER_WG][NV];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint16_t *Ls = smem[wv];
    for (;;) {
        const uint32_t cw = next_codeword(queue, lane);
        if (cw >= n_cw) break;
        const uint4 *src = reinterpret_cast<const uint4 *>(llr_in + (size_t)cw * NV);
#pragma unroll
        for (int i = 0; i < NV / 8 / 64; ++i) reinterpret_cast<uint4 *>(Ls)[i * 64 + lane] = src[i * 64 + lane];
        const int it = decode_wave(Ls, lane, max_iter, beta6, beta3);
        wave_lds_fence();
        uint4 *dst = reinterpret_cast<uint4 *>(llr_out + (size_t)cw * NV);
#pragma unroll
        for (int i = 0; i < NV / 8 / 64; ++i) dst[i * 64 + lane] = reinterpret_cast<const uint4 *>(Ls)[i * 64 + lane];
        if (lane == 0) {
            iters[cw] = it ? (uint32_t)it : (uint32_t)max_iter;
            ok[cw] = it ? 1 : 0;
        }
        wave_lds_fence();
    }
}

// Fused Monte Carlo: channel generation -> decode -> tallies.
// counters: [0] blocks, [1] frame errors, [2] bit errors, [3] iteration sum;
// hist[i] (i = 1..max_iter) counts convergence at iteration i, hist[0] failures.
// The histogram is binned per workgroup in LDS (bins 0..HIST_LDS-1) and
// flushed once per workgroup at the end; iteration counts beyond the LDS bins
// go straight to global memory.  One global atomic per codeword would cost a
// 64-byte HBM write each.
constexpr int HIST_LDS = 256;

template <bool RANDOM>
__device__ __forceinline__ void monte_carlo(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma,
                                            int max_iter, float beta6, float beta3, uint32_t *queue,
                                            unsigned long long *counters, unsigned long long *hist, const Capture &cap,
                                            const uint32_t *__restrict__ gen) {
    // random codewords keep their info bits right after the wave's LLRs, so
    // that one base address serves both (a second one costs registers)
    __shared__ __attribute__((aligned(16))) uint16_t smem[WAVES_PER_WG][NV + (RANDOM ? 2 * KINFO / 32 : 0)];
    __shared__ uint32_t hist_s[HIST_LDS];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < HIST_LDS; i += blockDim.x) hist_s[i] = 0u;
    __syncthreads();
    uint16_t *Ls = smem[wv];
    uint32_t *ub = reinterpret_cast<uint32_t *>(Ls + NV);
    uint64_t blocks = 0, fe = 0, be = 0, its = 0;
    for (;;) {
        const uint32_t cw = next_codeword(queue, lane);
        if (cw >= n_cw) break;
        wave_gen_channel<RANDOM>(Ls, ub, lane, seed, first_cw + cw, mu, sigma, gen);
        wave_lds_fence();
        const int it = decode_wave(Ls, lane, max_iter, beta6, beta3);
        blocks += 1;
        its += it ? (uint64_t)it : (uint64_t)max_iter;
        bool failed = !it;
        if constexpr (RANDOM) {
            bool mism;
            uint32_t errs;
            wave_cw_errors(Ls, ub, lane, mism, errs);
            failed = failed || mism;
            if (failed) be += errs;
        } else {
            if (failed) be += wave_info_errors(Ls, lane);
        }
        if (failed) fe += 1;
        if (lane == 0) {
            if (it < HIST_LDS)
                atomicAdd(hist_s + it, 1u);
            else
                atomicAdd(hist + it, 1ull);
            if (failed) capture_failure(cap, first_cw + cw);
        }
        wave_lds_fence();
    }
    if (lane == 0 && blocks) {
        atomicAdd(counters + 0, (unsigned long long)blocks);
        if (fe) atomicAdd(counters + 1, (unsigned long long)fe);
        if (be) atomicAdd(counters + 2, (unsigned long long)be);
        atomicAdd(counters + 3, (unsigned long long)its);
    }
    __syncthreads();
    const int nb = max_iter + 1 < HIST_LDS ? max_iter + 1 : HIST_LDS;
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        if (hist_s[i]) atomicAdd(hist + i, (unsigned long long)hist_s[i]);
}

__global__ void __launch_bounds__(64 * WAVES_PER_WG)
    k_monte_carlo(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, int max_iter, float beta6,
                  float beta3, uint32_t *queue, unsigned long long *counters, unsigned long long *hist, Capture cap) {
    monte_carlo<false>(seed, first_cw, n_cw, mu, sigma, max_iter, beta6, beta3, queue, counters, hist, cap, nullptr);
}

// Random codewords.  The encoder and the tally against the sent bits push
// the register count past 128, and the compiler then schedules the decode
// for 2 waves per SIMD (214 VGPRs, measured 14% slower at 1.5 dB); pinned to
// 4 waves per SIMD the decode loop keeps its registers and only the
// once-per-codeword generation and tally spill (about 35 dwords per lane).
__global__ void __launch_bounds__(64 * WAVES_PER_WG) __attribute__((amdgpu_waves_per_eu(4)))
    k_monte_carlo_cw(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, int max_iter,
                     float beta6, float beta3, uint32_t *queue, unsigned long long *counters,
                     unsigned long long *hist, Capture cap, const uint32_t *__restrict__ gen) {
    monte_carlo<true>(seed, first_cw, n_cw, mu, sigma, max_iter, beta6, beta3, queue, counters, hist, cap, gen);
}

// Channel LLRs of random codewords, one wave per codeword, through the Monte
// Carlo kernel's own generator.
__global__ void __launch_bounds__(64 * WAVES_PER_WG)
    k_gen_channel_cw(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma,
                     uint16_t *__restrict__ out, const uint32_t *__restrict__ gen) {
    __shared__ __attribute__((aligned(16))) uint16_t smem[WAVES_PER_WG][NV + 2 * KINFO / 32];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t c = blockIdx.x * WAVES_PER_WG + wv;
    if (c >= n_cw) return;
    uint16_t *Ls = smem[wv];
    wave_gen_channel<true>(Ls, reinterpret_cast<uint32_t *>(Ls + NV), lane, seed, first_cw + c, mu, sigma, gen);
    wave_lds_fence();
    uint4 *dst = reinterpret_cast<uint4 *>(out + (size_t)c * NV);
#pragma unroll
    for (int i = 0; i < NV / 8 / 64; ++i) dst[i * 64 + lane] = reinterpret_cast<const uint4 *>(Ls)[i * 64 + lane];
}

__global__ void __launch_bounds__(256)
    k_gen_channel(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, uint16_t *__restrict__ out) {
    // one thread per Philox block of 4 symbols
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t blocks_per_cw = NTX / 4;
    if (gid >= (uint64_t)n_cw * blocks_per_cw) return;
    const uint32_t c = (uint32_t)(gid / blocks_per_cw), blk = (uint32_t)(gid % blocks_per_cw);
    const uint64_t w = first_cw + c;
    u32x4 x = philox4x32_10((uint32_t)w, (uint32_t)(w >> 32), blk, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
    float z0, z1, z2, z3;
    box_muller(x.x, x.y, z0, z1);
    box_muller(x.z, x.w, z2, z3);
    uint32_t lo = (llr_bits(mu, sigma, z1) & 0xFFFF0000u) | (llr_bits(mu, sigma, z0) >> 16);
    uint32_t hi = (llr_bits(mu, sigma, z3) & 0xFFFF0000u) | (llr_bits(mu, sigma, z2) >> 16);
    uint16_t *L = out + (size_t)c * NV;
    *reinterpret_cast<uint2 *>(L + 4 * blk) = make_uint2(lo, hi);
    if (blk < (NV - NTX) / 4) *reinterpret_cast<uint2 *>(L + NTX + 4 * blk) = make_uint2(0u, 0u);
}

// Tail counts of the channel's normals: counts[k] += number of z with
// |z| > thr[k] among the normals of codewords first_cw .. first_cw + n_cw - 1
// (Philox blocks 0 .. n_tx/4 - 1 of each, exactly the z the channel of those
// codewords uses).  Grid-stride over (codeword, block) pairs; per-thread
// counts are summed over the wave with ballots and added once per wave.
struct TailThresholds {
    float thr[NOISE_TAIL_MAX];
    int n;
};

__global__ void __launch_bounds__(256)
    k_noise_tails(uint64_t seed, uint64_t first_cw, uint64_t n_cw, uint32_t blocks_per_cw, TailThresholds th,
                  unsigned long long *counts) {
    uint32_t local[NOISE_TAIL_MAX];
#pragma unroll
    for (int k = 0; k < NOISE_TAIL_MAX; ++k) local[k] = 0u;
    const uint64_t total = n_cw * blocks_per_cw;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
        const uint64_t w = first_cw + g / blocks_per_cw;
        const uint32_t blk = (uint32_t)(g % blocks_per_cw);
        const u32x4 x = philox4x32_10((uint32_t)w, (uint32_t)(w >> 32), blk, 0u, (uint32_t)seed,
                                      (uint32_t)(seed >> 32));
        float z[4];
        box_muller(x.x, x.y, z[0], z[1]);
        box_muller(x.z, x.w, z[2], z[3]);
#pragma unroll
        for (int k = 0; k < NOISE_TAIL_MAX; ++k)
            if (k < th.n)
#pragma unroll
                for (int j = 0; j < 4; ++j) local[k] += fabsf(z[j]) > th.thr[k] ? 1u : 0u;
    }
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < NOISE_TAIL_MAX; ++k) {
        if (k >= th.n) break;
        // wave sum of local[k] bit by bit (each lane's count < 2^24 here)
        unsigned long long sum = 0;
        for (int b = 0; b < 24; ++b) sum += (unsigned long long)__popcll(__ballot((local[k] >> b) & 1u)) << b;
        if (lane == 0 && sum) atomicAdd(counts + k, sum);
    }
}

__global__ void k_fill_u64(unsigned long long *p, size_t n, unsigned long long v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers (host side)
// ---------------------------------------------------------------------------
int t12_waves_per_wg() { return WAVES_PER_WG; }

hipError_t t12_occupancy(int *wg_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, k_monte_carlo, 64 * WAVES_PER_WG, 0);
}

hipError_t t12_occupancy_decode(int *wg_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, k_decode_llrs, 64 * WAVES_PER_WG, 0);
}

hipError_t launch_fill_u64(unsigned long long *p, size_t n, unsigned long long v, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t t12_launch_decode(const uint16_t *llr_in, uint16_t *llr_out, uint32_t *iters, uint8_t *ok, uint32_t n_cw,
                             int max_iter, float beta, uint32_t *queue, int grid, hipStream_t st) {
    const float beta3 = beta < 1.0f ? beta : 1.0f;  // stepper:300-304
    hipLaunchKernelGGL(k_decode_llrs, dim3(grid), dim3(64 * WAVES_PER_WG), 0, st, llr_in, llr_out, iters, ok, n_cw,
                       max_iter, beta, beta3, queue);
    return hipGetLastError();
}

hipError_t t12_launch_mc(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, int max_iter,
                         float beta, uint32_t *queue, unsigned long long *counters, unsigned long long *hist, int grid,
                         const Capture &cap, const uint32_t *cw_gen, hipStream_t st) {
    const float beta3 = beta < 1.0f ? beta : 1.0f;
    if (cw_gen)
        hipLaunchKernelGGL(k_monte_carlo_cw, dim3(grid), dim3(64 * WAVES_PER_WG), 0, st, seed, first_cw, n_cw, mu,
                           sigma, max_iter, beta, beta3, queue, counters, hist, cap, cw_gen);
    else
        hipLaunchKernelGGL(k_monte_carlo, dim3(grid), dim3(64 * WAVES_PER_WG), 0, st, seed, first_cw, n_cw, mu,
                           sigma, max_iter, beta, beta3, queue, counters, hist, cap);
    return hipGetLastError();
}

hipError_t t12_launch_gen_cw(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, uint16_t *out,
                             const uint32_t *cw_gen, hipStream_t st) {
    if (!n_cw) return hipSuccess;
    const unsigned grid = (n_cw + WAVES_PER_WG - 1) / WAVES_PER_WG;
    hipLaunchKernelGGL(k_gen_channel_cw, dim3(grid), dim3(64 * WAVES_PER_WG), 0, st, seed, first_cw, n_cw, mu, sigma,
                       out, cw_gen);
    return hipGetLastError();
}

int t12_cw_gen_words() { return CW_GEN_WORDS; }

hipError_t launch_noise_tails(uint64_t seed, uint64_t first_cw, uint64_t n_cw, int n_tx, const float *thr, int n_thr,
                             unsigned long long *counts, int cus, hipStream_t st) {
    if (n_thr < 1 || n_thr > NOISE_TAIL_MAX || n_tx < 4 || (n_tx & 3)) return hipErrorInvalidValue;
    TailThresholds th{};
    for (int k = 0; k < n_thr; ++k) th.thr[k] = thr[k];
    th.n = n_thr;
    const uint64_t total = n_cw * (uint64_t)(n_tx / 4);
    // at most 2^19 Philox blocks (2^21 normals) per thread keeps every count < 2^24
    uint64_t grid = (uint64_t)cus * 32;
    while (grid * 256 * (1ull << 19) < total) grid *= 2;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_noise_tails, dim3((unsigned)grid), dim3(256), 0, st, seed, first_cw, n_cw,
                       (uint32_t)(n_tx / 4), th, counts);
    return hipGetLastError();
}

hipError_t t12_launch_gen(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, uint16_t *out,
                          hipStream_t st) {
    const uint64_t threads = (uint64_t)n_cw * (NTX / 4);
    if (!threads) return hipSuccess;
    hipLaunchKernelGGL(k_gen_channel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, seed, first_cw, n_cw,
                       mu, sigma, out);
    return hipGetLastError();
}

// Compiled-in schedule, for host-side cross-checks: per layer
// [first_row, deg, (base, shift) x deg].
int t12_schedule(int32_t *out, int cap) {
    int n = 0;
    for (int li = 0; li < NLAYERS; ++li) {
        const ar4ja::Layer &L = S12.layer[li];
        if (n + 2 + 2 * L.deg > cap) return -1;
        out[n++] = L.first_row;
        out[n++] = L.deg;
        for (int d = 0; d < L.deg; ++d) {
            out[n++] = L.slot[d].base;
            out[n++] = L.slot[d].shift;
        }
    }
    return n;
}

}  // namespace ldpc
