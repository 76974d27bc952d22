// decode_multi.hip -- Profile T decoder for the nine CCSDS AR4JA codes on
// gfx950: K = 1024, 4096, 16384 at rates 1/2, 2/3, 4/5 (circulant M = 128 ..
// 8192; check degrees 3 and 6, 10 or 18).
//
// Compiled once per code: -DLDPC_KID=0..2 (K = 1024 << 2 KID) and
// -DLDPC_RID=0..2 (rate 1/2, 2/3, 4/5).  Each object instantiates the kernel
// for C = 1 or 2 checks per thread and for two layer orders, and exports its
// variant table (multi_table_k<KID>_r<RID>).
//
// Arithmetic: the golden's Profile T check update (stepper:107-327: q =
// RNE(L - R), tree Approx-Min* with bf16-truncated constants, damping beta,
// copysign, RNE of R' and L').  The degree-3 and degree-6 trees are the
// golden's; even degrees D > 6 (the rate-2/3 and rate-4/5 codes, which the
// reference's Profile T never decodes) pair edge j with edge j + D/2 and
// combine the other pairs through a prefix/suffix pass, which at D = 6 is
// exactly the golden's tree.  Syndrome with hard bit = (bf16 < 0) after every
// iteration; a converged codeword is latched; frame error <=> not converged;
// bit errors = information bits <= 0 of failed codewords (stepper:773-815).
//
// Mapping.  A layer is Z = M/4 checks of one block-row quarter (no two share
// a variable).  A codeword is owned by G = Z / C threads; thread t updates
// checks t and t + G (C = 2) or check t (C = 1) of every layer.
//  * G <= 64 ("wave mode"): a wave holds 64 / G codeword slots in lockstep.
//    LDS requests of one wave complete in issue order, so a compiler fence
//    separates layers.  After each iteration every slot tests its own
//    syndrome; a slot whose codeword converged or ran out of iterations is
//    tallied and refilled from the work queue while the other slots go on.
//  * G > 64 ("group mode"): a workgroup of G threads owns one codeword;
//    workgroup barriers separate layers.
// A codeword's LLRs live in LDS (N bf16); its messages R live in VGPRs, two
// bf16 per dword (the two checks of a thread for C = 2, two edges of the
// check for C = 1).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "ar4ja_schedule.h"
#include "device_math.h"
#include "kernels.h"

#pragma clang fp contract(off)

#if !defined(LDPC_KID) || !defined(LDPC_RID)
#error "compile with -DLDPC_KID=<0|1|2> -DLDPC_RID=<0|1|2>"
#endif

namespace ldpc {
namespace {

constexpr int HIST_LDS = 256;
constexpr int NLAYERS = 12;

__device__ __forceinline__ float ld_bf(const uint16_t *p) { return u2f((uint32_t)(*p) << 16); }
__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }
// Hard decision "bf16 < 0.0f" (stepper:781): sign set and magnitude nonzero.
__device__ __forceinline__ uint32_t hard_neg(uint32_t h) { return h > 0x8000u ? 1u : 0u; }
// Bit error test "bf16 <= 0.0f" (stepper:806).
__device__ __forceinline__ uint32_t nonpos(uint32_t h) { return ((h & 0x8000u) || h == 0u) ? 1u : 0u; }

// Tree min* over the other edges of a check, times beta.
template <int D>
__device__ __forceinline__ void tree(const float (&m)[D], float beta, float (&r)[D]) {
    if constexpr (D == 3) {
        r[0] = __fmul_rn(amin_t(m[1], m[2]), beta);
        r[1] = __fmul_rn(amin_t(m[0], m[2]), beta);
        r[2] = __fmul_rn(amin_t(m[0], m[1]), beta);
    } else {
        static_assert(D % 2 == 0 && D >= 4, "Profile T: degree 3 or even");
        constexpr int H = D / 2;
        float p[H], P[H], S[H];
#pragma unroll
        for (int j = 0; j < H; ++j) p[j] = amin_t(m[j], m[j + H]);
        P[0] = p[0];
#pragma unroll
        for (int j = 1; j < H - 1; ++j) P[j] = amin_t(P[j - 1], p[j]);
        S[H - 1] = p[H - 1];
#pragma unroll
        for (int j = H - 2; j >= 1; --j) S[j] = amin_t(p[j], S[j + 1]);
#pragma unroll
        for (int j = 0; j < H; ++j) {
            const float o = j == 0 ? S[1] : j == H - 1 ? P[H - 2] : amin_t(P[j - 1], S[j + 1]);
            r[j] = __fmul_rn(amin_t(o, m[j + H]), beta);
            r[j + H] = __fmul_rn(amin_t(o, m[j]), beta);
        }
    }
}

template <int C_, int ORD_>
struct Cfg {
    static constexpr int C = C_, ORD = ORD_;
    static constexpr ar4ja::Schedule S = ar4ja::make_schedule(LDPC_KID, LDPC_RID, ORD_);
    static constexpr int Z = S.Z, N = S.N, NP = S.n_punct, NTX = S.N - S.n_punct, K = S.k_info;
    static constexpr int G = Z / C;  // threads per codeword
    static constexpr bool WAVE = G <= 64;
    static constexpr int SLOTS = WAVE ? 64 / G : 1;  // codewords per wave (wave mode)
    static constexpr int WPW = WAVE ? 4 : G / 64;    // waves per workgroup
    static constexpr int THREADS = 64 * WPW;
    static constexpr int CW_PER_WG = WAVE ? SLOTS * WPW : 1;
    static constexpr int rdw(int deg) { return C == 2 ? deg : (deg + 1) / 2; }
    // Messages of some layers live in LDS instead of VGPRs, laid out
    // [dword][thread of the workgroup] (each thread reads only its own
    // column: no bank conflicts, no barrier):
    //  * group mode with 1024 threads and C = 2 (K = 16384 r1/2): 60 message
    //    dwords per thread do not fit a 1024-thread group's 128 VGPRs beside
    //    the working set; the degree-3 layers go to LDS (12 dwords x 1024
    //    threads = 48 KiB beside the 80 KiB of L);
    //  * rate 4/5, C = 1, group mode with at most 256 threads (K = 4096):
    //    80 message dwords and the degree-18 working set take 177 VGPRs,
    //    2 waves per SIMD; the degree-3 layers and the first two degree-18
    //    layers go to LDS (26 dwords per thread): about 148 VGPRs, 3 waves per
    //    SIMD, 6% faster at 3.0 dB;
    //  * rate 4/5, C = 1, group mode with 512 threads (K = 16384): the
    //    degree-3 layers and the first degree-18 layer go to LDS (17 dwords),
    //    the degree-18 layers recompute their store addresses after the tree,
    //    and the kernel is held to 4 waves per SIMD (128 VGPRs, a few dwords
    //    spilled outside the iteration loop): two workgroups per CU.
    // Waves per SIMD the compiler must fit: the K = 16384 r4/5 layout above
    // at 4; rate 4/5 in wave mode with one check per thread (K = 1024 r4/5),
    // which compiles to 171 VGPRs (2 waves per SIMD) unconstrained, at 3,
    // where it fits in 167 with no spill and keeps every message in VGPRs.
    static constexpr bool LDS_R45B = LDPC_RID == 2 && C == 1 && !WAVE && THREADS == 512;
    static constexpr bool RECOMPUTE = LDS_R45B;
    static constexpr int MIN_WAVES = (LDPC_RID == 2 && WAVE && C == 1) ? 3 : LDS_R45B ? 4 : 1;
    static constexpr bool LDS_R12 = !WAVE && THREADS == 1024 && C == 2;
    static constexpr bool LDS_R45 = LDPC_RID == 2 && C == 1 && !WAVE && THREADS <= 256;
    static constexpr int deg18_before(int li) {
        int n = 0;
        for (int i = 0; i < li; ++i) n += S.layer[i].deg == 18 ? 1 : 0;
        return n;
    }
    static constexpr bool in_lds(int li) {
        return (LDS_R12 && S.layer[li].deg == 3) ||
               (LDS_R45 && (S.layer[li].deg == 3 || (S.layer[li].deg == 18 && deg18_before(li) < 2))) ||
               (LDS_R45B && (S.layer[li].deg == 3 || (S.layer[li].deg == 18 && deg18_before(li) < 1)));
    }
    static constexpr int r_off(int li) {  // offset in the VGPR array
        int o = 0;
        for (int i = 0; i < li; ++i) o += in_lds(i) ? 0 : rdw(S.layer[i].deg);
        return o;
    }
    static constexpr int rl_off(int li) {  // offset in the LDS array (dwords per thread)
        int o = 0;
        for (int i = 0; i < li; ++i) o += in_lds(i) ? rdw(S.layer[i].deg) : 0;
        return o;
    }
    static constexpr int NR = r_off(NLAYERS);
    static constexpr int NRL = rl_off(NLAYERS);
    static_assert(G >= 16 && G <= 1024 && (G & (G - 1)) == 0, "threads per codeword");
    static_assert(NTX % 4 == 0 && NP % 4 == 0 && N % 8 == 0, "channel layout");
};

// One layer of the thread's checks; Ls is the codeword's LLR array in LDS.
// Rl: the thread's column of the LDS message array (element k at
// Rl[k * THREADS]); unused when the configuration keeps every message in VGPRs.
template <class CF, int LI>
__device__ __forceinline__ void run_layer(uint16_t *Ls, uint32_t (&R)[CF::NR], uint32_t *Rl, int t, float beta_e,
                                          float beta3) {
    constexpr ar4ja::Layer Ly = CF::S.layer[LI];
    constexpr int D = Ly.deg, R0 = CF::r_off(LI), RL0 = CF::rl_off(LI), Z = CF::Z, G = CF::G;
    const float beta = D == 3 ? beta3 : beta_e;
    // Opaque copy of the thread index: keeps the shifted LDS addresses from
    // being hoisted out of the iteration loop (they would pin VGPRs).
    asm volatile("" : "+v"(t));
    if constexpr (CF::C == 2) {
        int ia[D], ib[D];
        uint32_t qp[D];
        float qa[D], qb[D], ma[D], mb[D];
        uint32_t sp = 0;  // sign parity of both checks: bit 15 (t), bit 31 (t + G)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int base = Ly.slot[d].base, sh = Ly.slot[d].shift;
            const int xa = sh == 0 ? t : ((t + sh) & (Z - 1));
            const int xb = sh == 0 ? t + G : (xa ^ G);  // (t + G + sh) mod Z, G = Z/2
            ia[d] = base + xa;
            ib[d] = base + xb;
            uint32_t rp;
            if constexpr (CF::in_lds(LI))
                rp = Rl[(RL0 + d) * CF::THREADS];
            else
                rp = R[R0 + d];
            qp[d] = cvt_pk_bf16(__fsub_rn(ld_bf(Ls + ia[d]), lo_f(rp)), __fsub_rn(ld_bf(Ls + ib[d]), hi_f(rp)));
            qa[d] = lo_f(qp[d]);
            qb[d] = hi_f(qp[d]);
            ma[d] = fabsf(qa[d]);
            mb[d] = fabsf(qb[d]);
            sp ^= qp[d];
        }
        float ra[D], rb[D];
        tree<D>(ma, beta, ra);
        tree<D>(mb, beta, rb);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const uint32_t rn = cvt_pk_bf16(ra[d], rb[d]) | ((sp ^ qp[d]) & 0x80008000u);
            const uint32_t ln = cvt_pk_bf16(__fadd_rn(qa[d], lo_f(rn)), __fadd_rn(qb[d], hi_f(rn)));
            Ls[ia[d]] = (uint16_t)ln;
            Ls[ib[d]] = (uint16_t)(ln >> 16);
            if constexpr (CF::in_lds(LI))
                Rl[(RL0 + d) * CF::THREADS] = rn;
            else
                R[R0 + d] = rn;
        }
    } else {
        constexpr int DW = (D + 1) / 2;
        int ia[2 * DW];
        uint32_t qw[DW];
        float q[2 * DW], m[D];
        uint32_t sp = 0;  // bit 15: parity of even edges, bit 31: odd edges
#pragma unroll
        for (int w = 0; w < DW; ++w) {
            const int d0 = 2 * w, d1 = 2 * w + 1;
            uint32_t rp;
            if constexpr (CF::in_lds(LI))
                rp = Rl[(RL0 + w) * CF::THREADS];
            else
                rp = R[R0 + w];
            ia[d0] = Ly.slot[d0].base + ((t + Ly.slot[d0].shift) & (Z - 1));
            const float a = __fsub_rn(ld_bf(Ls + ia[d0]), lo_f(rp));
            float b = 0.0f;
            if (d1 < D) {
                ia[d1] = Ly.slot[d1].base + ((t + Ly.slot[d1].shift) & (Z - 1));
                b = __fsub_rn(ld_bf(Ls + ia[d1]), hi_f(rp));
            }
            qw[w] = cvt_pk_bf16(a, b);
            q[d0] = lo_f(qw[w]);
            q[d1] = hi_f(qw[w]);
            sp ^= qw[w];
        }
#pragma unroll
        for (int d = 0; d < D; ++d) m[d] = fabsf(q[d]);
        const uint32_t s1 = (sp ^ (sp >> 16)) & 0x8000u, s2 = s1 | (s1 << 16);
        float r[D];
        tree<D>(m, beta, r);
        // Wide checks recompute their addresses for the stores instead of
        // keeping 2 DW of them live across the tree.
        int tw = t;
        if constexpr (CF::RECOMPUTE && D >= 10) asm volatile("" : "+v"(tw));
#pragma unroll
        for (int w = 0; w < DW; ++w) {
            const int d0 = 2 * w, d1 = 2 * w + 1;
            const uint32_t rn = cvt_pk_bf16(r[d0], d1 < D ? r[d1 < D ? d1 : 0] : 0.0f) | ((s2 ^ qw[w]) & 0x80008000u);
            const uint32_t ln = cvt_pk_bf16(__fadd_rn(q[d0], lo_f(rn)), __fadd_rn(q[d1], hi_f(rn)));
            const int w0 = CF::RECOMPUTE && D >= 10 ? Ly.slot[d0].base + ((tw + Ly.slot[d0].shift) & (Z - 1)) : ia[d0];
            Ls[w0] = (uint16_t)ln;
            if (d1 < D) {
                const int w1 =
                    CF::RECOMPUTE && D >= 10 ? Ly.slot[d1].base + ((tw + Ly.slot[d1].shift) & (Z - 1)) : ia[d1];
                Ls[w1] = (uint16_t)(ln >> 16);
            }
            if constexpr (CF::in_lds(LI))
                Rl[(RL0 + w) * CF::THREADS] = rn;
            else
                R[R0 + w] = rn;
        }
    }
}

template <class CF, int... LI>
__device__ __forceinline__ void run_iteration(uint16_t *Ls, uint32_t (&R)[CF::NR], uint32_t *Rl, int t, float beta_e,
                                              float beta3, std::integer_sequence<int, LI...>) {
    if constexpr (CF::WAVE)
        ((run_layer<CF, LI>(Ls, R, Rl, t, beta_e, beta3), wave_fence()), ...);
    else
        ((run_layer<CF, LI>(Ls, R, Rl, t, beta_e, beta3), __syncthreads()), ...);
}

// 1 if any of the thread's checks of layer LI is unsatisfied.
template <class CF, int LI>
__device__ __forceinline__ bool layer_fails(const uint16_t *Ls, int t) {
    constexpr ar4ja::Layer Ly = CF::S.layer[LI];
    constexpr int D = Ly.deg, Z = CF::Z, G = CF::G;
    asm volatile("" : "+v"(t));  // addresses stay inside the loop (see run_layer)
    uint32_t pa = 0, pb = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int x = (t + Ly.slot[d].shift) & (Z - 1);
        pa ^= hard_neg(Ls[Ly.slot[d].base + x]);
        if constexpr (CF::C == 2) pb ^= hard_neg(Ls[Ly.slot[d].base + (x ^ G)]);
    }
    return (pa | pb) != 0u;
}

// Bit s of the result: some lane of wave slot s (lanes s*G .. s*G+G-1) is set.
template <class CF>
__device__ __forceinline__ uint32_t slot_any(uint64_t b) {
    if constexpr (CF::SLOTS == 1) {
        return b != 0 ? 1u : 0u;
    } else {
        uint32_t r = 0;
#pragma unroll
        for (int s = 0; s < CF::SLOTS; ++s)
            if ((b >> (s * CF::G)) & ((1ull << CF::G) - 1ull)) r |= 1u << s;
        return r;
    }
}

// Wave mode: of the slots in `pass`, those whose every check is satisfied.
// Layers are tested in order and the test stops once every slot has an
// unsatisfied check, so an unconverged iteration usually costs one layer.
template <class CF, int LI>
__device__ __forceinline__ uint32_t syndrome_pass(const uint16_t *Ls, int t, uint32_t pass) {
    if constexpr (LI == NLAYERS) {
        return pass;
    } else {
        asm volatile("" ::: "memory");
        pass &= ~slot_any<CF>(__ballot(layer_fails<CF, LI>(Ls, t)));
        if (pass == 0) return 0;
        return syndrome_pass<CF, LI + 1>(Ls, t, pass);
    }
}

// Group mode: does any check of the codeword fail?  The waves share one LDS
// flag: a wave that finds an unsatisfied check among its threads' checks
// sets it, and a wave stops scanning its layers as soon as it sees it set.
// Near convergence only a few checks fail and most waves would otherwise
// scan all 12 layers before the barrier; the answer is the same either way
// (the flag is set if and only if some check fails).
template <class CF, int LI>
__device__ __forceinline__ void scan_fails_from(const uint16_t *Ls, int t, volatile uint32_t *fail) {
    if constexpr (LI < NLAYERS) {
        asm volatile("" ::: "memory");
        if (__builtin_amdgcn_readfirstlane(*fail)) return;
        if (__ballot(layer_fails<CF, LI>(Ls, t)) != 0) {
            if ((t & 63) == 0) *fail = 1u;
            return;
        }
        scan_fails_from<CF, LI + 1>(Ls, t, fail);
    }
}

// Channel LLRs of codeword w by the G threads of its owner: Philox(ctr =
// (w_lo, w_hi, blk, 0), key = seed) -> 4 normals -> RNE_bf16(mu + sigma z);
// punctured variables = +0.
template <class CF>
__device__ __forceinline__ void gen_channel(uint16_t *Ls, int t, uint64_t seed, uint64_t w, float mu, float sigma) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int b = t; b < CF::NTX / 4; b += CF::G) {
        const u32x4 x = philox4x32_10((uint32_t)w, (uint32_t)(w >> 32), (uint32_t)b, 0u, k0, k1);
        float z0, z1, z2, z3;
        box_muller(x.x, x.y, z0, z1);
        box_muller(x.z, x.w, z2, z3);
        const uint32_t lo = (llr_bits(mu, sigma, z1) & 0xFFFF0000u) | (llr_bits(mu, sigma, z0) >> 16);
        const uint32_t hi = (llr_bits(mu, sigma, z3) & 0xFFFF0000u) | (llr_bits(mu, sigma, z2) >> 16);
        *reinterpret_cast<uint2 *>(Ls + 4 * b) = make_uint2(lo, hi);
    }
    for (int b = t; b < CF::NP / 4; b += CF::G) *reinterpret_cast<uint2 *>(Ls + CF::NTX + 4 * b) = make_uint2(0u, 0u);
}

template <class CF>
__device__ __forceinline__ void copy_in(uint16_t *Ls, const uint16_t *src, int t) {
    for (int i = t; i < CF::N / 4; i += CF::G)
        reinterpret_cast<uint2 *>(Ls)[i] = reinterpret_cast<const uint2 *>(src)[i];
}

template <class CF>
__device__ __forceinline__ void copy_out(uint16_t *dst, const uint16_t *Ls, int t) {
    for (int i = t; i < CF::N / 4; i += CF::G)
        reinterpret_cast<uint2 *>(dst)[i] = reinterpret_cast<const uint2 *>(Ls)[i];
}

template <class CF>
__device__ __forceinline__ uint32_t info_errors(const uint16_t *Ls, int t) {
    uint32_t be = 0;
    for (int i = t; i < CF::K; i += CF::G) be += nonpos(Ls[i]);
    return be;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

struct Run {
    uint64_t seed, first_cw;
    uint32_t n_cw;
    float mu, sigma;
};

__device__ __forceinline__ Run load_run(const MultiParams &P) {
    Run r{P.seed, P.first_cw, P.n_cw, P.mu, P.sigma};
    if (P.dyn) {
        r.seed = P.dyn->seed;
        r.first_cw = P.dyn->first_cw;
        r.n_cw = P.dyn->n_cw;
        r.mu = P.dyn->mu;
        r.sigma = P.dyn->sigma;
    }
    return r;
}

__device__ __forceinline__ void flush_counters(const MultiParams &P, unsigned long long blocks, unsigned long long fe,
                                               unsigned long long be, unsigned long long its) {
    if (blocks) atomicAdd(P.counters + 0, blocks);
    if (fe) atomicAdd(P.counters + 1, fe);
    if (be) atomicAdd(P.counters + 2, be);
    if (its) atomicAdd(P.counters + 3, its);
}

__device__ __forceinline__ void hist_add(uint32_t *hist_s, const MultiParams &P, int bin) {
    if (bin < HIST_LDS)
        atomicAdd(hist_s + bin, 1u);
    else
        atomicAdd(P.hist + bin, 1ull);
}

__device__ __forceinline__ void hist_flush(const uint32_t *hist_s, const MultiParams &P, int threads) {
    const int nb = P.max_iter + 1 < HIST_LDS ? P.max_iter + 1 : HIST_LDS;
    for (int i = threadIdx.x; i < nb; i += threads)
        if (hist_s[i]) atomicAdd(P.hist + i, (unsigned long long)hist_s[i]);
}

// ---------------------------------------------------------------------------
// Wave mode
// ---------------------------------------------------------------------------
template <class CF>
__global__ void __launch_bounds__(CF::THREADS) __attribute__((amdgpu_waves_per_eu(CF::MIN_WAVES)))
k_multi_wave(MultiParams P) {
    __shared__ __attribute__((aligned(16))) uint16_t smem[CF::CW_PER_WG * CF::N];
    __shared__ uint32_t Rl_s[CF::NRL > 0 ? CF::NRL * CF::THREADS : 1];
    __shared__ uint32_t hist_s[HIST_LDS];
    const Run run = load_run(P);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int slot = lane / CF::G, t = lane % CF::G;
    uint16_t *Ls = smem + (wv * CF::SLOTS + slot) * CF::N;
    uint32_t *Rl = Rl_s + threadIdx.x;
    for (int i = threadIdx.x; i < HIST_LDS; i += CF::THREADS) hist_s[i] = 0u;
    __syncthreads();

    // Work distribution: a slot takes its next codeword from the device queue
    // (one atomic per codeword) when it finishes one, so slots that drew easy
    // codewords, or whose codewords fail at max_iter late in the launch, even
    // out.  A round-robin static share of 7/8 of the launch left K=1024 r4/5
    // (4.8% of codewords running 200 iterations) at 0.58x this rate.
    const uint32_t all_slots = (1u << CF::SLOTS) - 1u;
    uint32_t cw = run.n_cw;
    int it = 0;
    uint32_t R[CF::NR];
#pragma unroll
    for (int i = 0; i < CF::NR; ++i) R[i] = 0u;
    unsigned long long blocks = 0, fe = 0, be = 0, its = 0;
    uint32_t done = all_slots, conv = 0, active = 0;
    for (;;) {
        if (done) {
            const bool mine = (done >> slot) & 1u;
            if (mine && cw < run.n_cw) {
                const bool ok = (conv >> slot) & 1u;
                if (!ok) be += info_errors<CF>(Ls, t);
                if (P.llr_in) copy_out<CF>(P.llr_out + (size_t)cw * CF::N, Ls, t);
                if (t == 0) {
                    blocks += 1;
                    its += ok ? (unsigned)it : (unsigned)P.max_iter;
                    fe += ok ? 0u : 1u;
                    hist_add(hist_s, P, ok ? it : 0);
                    if (!ok && !P.llr_in) capture_failure(P.cap, run.first_cw + cw);
                    if (P.llr_in) {
                        P.iters[cw] = ok ? (uint32_t)it : (uint32_t)P.max_iter;
                        P.ok[cw] = ok ? 1 : 0;
                    }
                }
            }
            uint32_t nxt = run.n_cw;
            if (mine && t == 0) {
                nxt = min(atomicAdd(P.queue, 1u), run.n_cw);
            }
            nxt = __shfl(nxt, slot * CF::G);
            if (mine) {
                cw = nxt;
                it = 0;
                if (cw < run.n_cw) {
                    if (P.llr_in)
                        copy_in<CF>(Ls, P.llr_in + (size_t)cw * CF::N, t);
                    else
                        gen_channel<CF>(Ls, t, run.seed, run.first_cw + cw, run.mu, run.sigma);
                }
            }
#pragma unroll
            for (int i = 0; i < CF::NR; ++i) R[i] = mine ? 0u : R[i];  // bf16 +0: iteration 0 has q = L
            if (mine) {
#pragma unroll
                for (int i = 0; i < CF::NRL; ++i) Rl[i * CF::THREADS] = 0u;
            }
            active = slot_any<CF>(__ballot(cw < run.n_cw));
            wave_fence();
        }
        if (!active) break;
        run_iteration<CF>(Ls, R, Rl, t, P.beta, P.beta3, std::make_integer_sequence<int, NLAYERS>{});
        it += 1;
        conv = syndrome_pass<CF, 0>(Ls, t, active);
        done = (conv | slot_any<CF>(__ballot(it >= P.max_iter))) & active;
    }
    blocks = wave_sum(blocks);
    fe = wave_sum(fe);
    be = wave_sum(be);
    its = wave_sum(its);
    if (lane == 0) flush_counters(P, blocks, fe, be, its);
    __syncthreads();
    hist_flush(hist_s, P, CF::THREADS);
}

// ---------------------------------------------------------------------------
// Group mode
// ---------------------------------------------------------------------------
template <class CF>
__global__ void __launch_bounds__(CF::THREADS) __attribute__((amdgpu_waves_per_eu(CF::MIN_WAVES)))
k_multi_group(MultiParams P) {
    __shared__ __attribute__((aligned(16))) uint16_t Ls[CF::N];
    __shared__ uint32_t Rl_s[CF::NRL > 0 ? CF::NRL * CF::THREADS : 1];
    __shared__ uint32_t hist_s[HIST_LDS];
    // fail_s[it & 1]: some check is unsatisfied after iteration it (two
    // flags, so that the one being set is never the one being read or cleared)
    __shared__ uint32_t fail_s[2];
    __shared__ uint32_t cw_s, be_s;
    const Run run = load_run(P);
    const int t = threadIdx.x;
    uint32_t *Rl = Rl_s + t;
    for (int i = t; i < HIST_LDS; i += CF::THREADS) hist_s[i] = 0u;
    unsigned long long blocks = 0, fe = 0, be = 0, its = 0;
    uint32_t R[CF::NR];
    __syncthreads();
    for (;;) {
        if (t == 0) cw_s = min(atomicAdd(P.queue, 1u), run.n_cw);
        __syncthreads();
        const uint32_t cw = cw_s;
        if (cw >= run.n_cw) break;
        if (P.llr_in)
            copy_in<CF>(Ls, P.llr_in + (size_t)cw * CF::N, t);
        else
            gen_channel<CF>(Ls, t, run.seed, run.first_cw + cw, run.mu, run.sigma);
        if (t == 0) {
            be_s = 0u;
            fail_s[0] = fail_s[1] = 0u;
        }
#pragma unroll
        for (int i = 0; i < CF::NR; ++i) R[i] = 0u;
#pragma unroll
        for (int i = 0; i < CF::NRL; ++i) Rl[i * CF::THREADS] = 0u;
        __syncthreads();
        int it = 0;
        bool ok = false;
        while (it < P.max_iter) {
            run_iteration<CF>(Ls, R, Rl, t, P.beta, P.beta3, std::make_integer_sequence<int, NLAYERS>{});
            ++it;
            scan_fails_from<CF, 0>(Ls, t, fail_s + (it & 1));
            __syncthreads();
            const uint32_t any = fail_s[it & 1];
            // every wave read the other flag after the previous iteration's
            // barrier, before this iteration's layer barriers
            if (t == 0) fail_s[(it + 1) & 1] = 0u;
            if (!any) {
                ok = true;
                break;
            }
        }
        if (!ok) {
            const uint32_t e = info_errors<CF>(Ls, t);
            if (e) atomicAdd(&be_s, e);
        }
        if (P.llr_in) copy_out<CF>(P.llr_out + (size_t)cw * CF::N, Ls, t);
        __syncthreads();
        if (t == 0) {
            blocks += 1;
            its += ok ? (unsigned)it : (unsigned)P.max_iter;
            if (!ok) {
                fe += 1;
                be += be_s;
            }
            hist_add(hist_s, P, ok ? it : 0);
            if (!ok && !P.llr_in) capture_failure(P.cap, run.first_cw + cw);
            if (P.llr_in) {
                P.iters[cw] = ok ? (uint32_t)it : (uint32_t)P.max_iter;
                P.ok[cw] = ok ? 1 : 0;
            }
        }
        __syncthreads();
    }
    if (t == 0) flush_counters(P, blocks, fe, be, its);
    __syncthreads();
    hist_flush(hist_s, P, CF::THREADS);
}

template <class CF>
constexpr auto kernel_of() {
    if constexpr (CF::WAVE)
        return &k_multi_wave<CF>;
    else
        return &k_multi_group<CF>;
}

template <class CF>
hipError_t launch(const MultiParams &p, int grid, hipStream_t st) {
    hipLaunchKernelGGL(kernel_of<CF>(), dim3(grid), dim3(CF::THREADS), 0, st, p);
    return hipGetLastError();
}

template <class CF>
hipError_t occupancy(int *wg_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, kernel_of<CF>(), CF::THREADS, 0);
}

template <class CF>
hipError_t attributes(hipFuncAttributes *a) {
    return hipFuncGetAttributes(a, reinterpret_cast<const void *>(kernel_of<CF>()));
}

// ---------------------------------------------------------------------------
// Profile D: the learned layered min-sum decoder (the forward of the
// differentiable decoder, csrc/dmin_torch.hip) with a syndrome early exit.
// fp32 L in LDS, fp32 R in VGPRs, one check per thread; the channel is the
// Profile T stream (bf16 LLRs), so both profiles see the same noise.
//   q = L - R; (min1, idx1), (min2, idx2) first-index; g = min2 on idx1,
//   min1 (or max(min1 - corr(min2 - min1), 0)) elsewhere; z = alpha g - beta;
//   r = copysign(min(max(z, 0), r_max), s ^ sign q); L = q + r; R = r.
//   With three minima (correction 2), g = min*(min2, min3) on idx1,
//   min*(min1, min3) on idx2 and min*(min1, min2) elsewhere, min*(lo, hi) =
//   max(lo - corr(hi - lo), 0).
// Hard decision L < 0, bit errors L <= 0, as in Profile T.
// ---------------------------------------------------------------------------
template <int ORD_>
struct CfgD {
    static constexpr int C = 1, ORD = ORD_;
    static constexpr ar4ja::Schedule S = ar4ja::make_schedule(LDPC_KID, LDPC_RID, ORD_);
    static constexpr int Z = S.Z, N = S.N, NP = S.n_punct, NTX = S.N - S.n_punct, K = S.k_info;
    static constexpr int G = Z;
    static constexpr bool WAVE = G <= 64;
    static constexpr int SLOTS = WAVE ? 64 / G : 1;
    static constexpr int WPW = WAVE ? 4 : G / 64;
    static constexpr int THREADS = 64 * WPW;
    static constexpr int CW_PER_WG = WAVE ? SLOTS * WPW : 1;
    static constexpr int r_off(int li) {
        int o = 0;
        for (int i = 0; i < li; ++i) o += S.layer[i].deg;
        return o;
    }
    static constexpr int NR = r_off(NLAYERS);  // messages per thread = parameters per row
    static_assert(G >= 16 && G <= 1024, "one check per thread");
};

// min*(lo, hi) of two magnitudes lo <= hi, Approx-Min* correction lines
__device__ __forceinline__ float amin_d(float lo, float hi) {
    const float C1 = 0.69140625f, K1 = -0.3984375f, C2 = 0.65625f, K2 = -0.326171875f;
    const float x = __fsub_rn(hi, lo);
    const float l1 = __fadd_rn(__fmul_rn(K1, x), C1);
    const float l2 = __fadd_rn(__fmul_rn(K2, x), C2);
    // max(l1, l2, 0) as one v_max3 (no NaN reaches it; the lines are never -0)
    const float o = __fsub_rn(lo, fmaxf(fmaxf(l1, l2), 0.0f));
    return o > 0.0f ? o : 0.0f;
}

// One layer with the three-minimum correction.  The edge's g is chosen by
// its magnitude, not by the minima's indices: an edge equal to min1 gets
// min*(min2, min3), one equal to min2 (above min1) min*(min1, min3), any
// other min*(min1, min2).  That is the first-index rule of the oracle on
// every input: an edge equal to min1 that is not idx1 means min1 = min2, so
// the pair it gets by index, min*(min1, min3) (or min*(min1, min2) when it is
// past idx2, then min3 = min1 too), equals min*(min2, min3); an edge equal to
// min2 that is not idx2 means min2 = min3, so min*(min1, min2) =
// min*(min1, min3).  No index is tracked.
template <class CF, int LI, bool CAP>
__device__ __forceinline__ void run_layer_d3(float *Ls, float (&R)[CF::NR], int t, const float *al, const float *be,
                                             float rcap) {
    constexpr ar4ja::Layer Ly = CF::S.layer[LI];
    constexpr int D = Ly.deg, R0 = CF::r_off(LI), Z = CF::Z;
    asm volatile("" : "+v"(t));
    typedef const __attribute__((address_space(4))) float cfloat;
    const cfloat *alc = (const cfloat *)al, *bec = (const cfloat *)be;
    int v[D];
    float q[D];
    float m1 = __builtin_inff(), m2 = __builtin_inff(), m3 = __builtin_inff();
    uint32_t s = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        v[d] = Ly.slot[d].base + ((t + Ly.slot[d].shift) & (Z - 1));
        q[d] = __fsub_rn(Ls[v[d]], R[R0 + d]);
        s ^= f2u(q[d]);
        // the three smallest |q| with multiplicity: with m1 <= m2 <= m3 the
        // new m3 is med3(a, m2, m3) and the new m2 med3(a, m1, m2)
        const float a = fabsf(q[d]);
        m3 = __builtin_amdgcn_fmed3f(a, m2, m3);
        m2 = __builtin_amdgcn_fmed3f(a, m1, m2);
        m1 = fminf(a, m1);
    }
    const float o12 = amin_d(m1, m2), o13 = amin_d(m1, m3), o23 = amin_d(m2, m3);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float a = fabsf(q[d]);
        const float g = a == m1 ? o23 : (a == m2 ? o13 : o12);
        const float z = __fsub_rn(__fmul_rn(alc[R0 + d], g), bec[R0 + d]);
        const float mag = CAP ? fminf(z > 0.0f ? z : 0.0f, rcap) : (z > 0.0f ? z : 0.0f);
        const float r = u2f(f2u(mag) | ((s ^ f2u(q[d])) & 0x80000000u));
        Ls[v[d]] = __fadd_rn(q[d], r);
        R[R0 + d] = r;
    }
}

template <class CF, int LI, bool CAP>
__device__ __forceinline__ void run_layer_d(float *Ls, float (&R)[CF::NR], int t, const float *al, const float *be,
                                            bool corr, float rcap) {
    constexpr ar4ja::Layer Ly = CF::S.layer[LI];
    constexpr int D = Ly.deg, R0 = CF::r_off(LI), Z = CF::Z;
    asm volatile("" : "+v"(t));
    // parameter rows through the constant address space: scalar loads where
    // the row is uniform (one codeword per workgroup), no VGPRs held for them
    typedef const __attribute__((address_space(4))) float cfloat;
    const cfloat *alc = (const cfloat *)al, *bec = (const cfloat *)be;
    int v[D];
    float q[D];
    float m1 = __builtin_inff(), m2 = __builtin_inff();
    int i1 = 0;
    uint32_t s = 0;  // bit 31: product of the signs
#pragma unroll
    for (int d = 0; d < D; ++d) {
        v[d] = Ly.slot[d].base + ((t + Ly.slot[d].shift) & (Z - 1));
        q[d] = __fsub_rn(Ls[v[d]], R[R0 + d]);
        s ^= f2u(q[d]);
        // (m1, i1) the first minimum of |q| (first index on ties), m2 the
        // second: "if a < m1: (m2, m1, i1) = (m1, a, d) elif a < m2: m2 = a"
        // without branches; with m1 <= m2 the new m2 is med3(a, m1, m2)
        const float a = fabsf(q[d]);
        i1 = a < m1 ? d : i1;
        m2 = __builtin_amdgcn_fmed3f(a, m1, m2);
        m1 = fminf(a, m1);
    }
    float other = m1;
    if (corr) {
        const float C1 = 0.69140625f, K1 = -0.3984375f, C2 = 0.65625f, K2 = -0.326171875f;
        const float x = __fsub_rn(m2, m1);
        const float l1 = __fadd_rn(__fmul_rn(K1, x), C1);
        const float l2 = __fadd_rn(__fmul_rn(K2, x), C2);
        const float o = __fsub_rn(m1, fmaxf(fmaxf(l1, l2), 0.0f));
        other = o > 0.0f ? o : 0.0f;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float g = d == i1 ? m2 : other;
        const float z = __fsub_rn(__fmul_rn(alc[R0 + d], g), bec[R0 + d]);
        // CAP: r_max > 0; for these non-negative finite magnitudes
        // min(mag, r_max) is "r_max unless mag < r_max"
        const float mag = CAP ? fminf(z > 0.0f ? z : 0.0f, rcap) : (z > 0.0f ? z : 0.0f);
        const float r = u2f(f2u(mag) | ((s ^ f2u(q[d])) & 0x80000000u));
        Ls[v[d]] = __fadd_rn(q[d], r);
        R[R0 + d] = r;
    }
}

// THREE: the three-minimum correction (its own kernel instance: a branch
// between the two forms inside one kernel raised the register count of both,
// 109 -> 162 VGPRs for K=1024 r1/2); otherwise corr selects the two-minimum
// correction or plain min-sum.
template <class CF, bool THREE, bool CAP, int... LI>
__device__ __forceinline__ void run_iteration_d(float *Ls, float (&R)[CF::NR], int t, const float *al, const float *be,
                                                bool corr, float rcap, std::integer_sequence<int, LI...>) {
    if constexpr (THREE) {
        if constexpr (CF::WAVE)
            ((run_layer_d3<CF, LI, CAP>(Ls, R, t, al, be, rcap), wave_fence()), ...);
        else
            ((run_layer_d3<CF, LI, CAP>(Ls, R, t, al, be, rcap), __syncthreads()), ...);
    } else {
        if constexpr (CF::WAVE)
            ((run_layer_d<CF, LI, CAP>(Ls, R, t, al, be, corr, rcap), wave_fence()), ...);
        else
            ((run_layer_d<CF, LI, CAP>(Ls, R, t, al, be, corr, rcap), __syncthreads()), ...);
    }
}

template <class CF, int LI>
__device__ __forceinline__ bool layer_fails_d(const float *Ls, int t) {
    constexpr ar4ja::Layer Ly = CF::S.layer[LI];
    asm volatile("" : "+v"(t));
    uint32_t p = 0;
#pragma unroll
    for (int d = 0; d < Ly.deg; ++d) p ^= Ls[Ly.slot[d].base + ((t + Ly.slot[d].shift) & (CF::Z - 1))] < 0.0f;
    return p != 0u;
}

template <class CF, int LI>
__device__ __forceinline__ uint32_t syndrome_pass_d(const float *Ls, int t, uint32_t pass) {
    if constexpr (LI == NLAYERS) {
        return pass;
    } else {
        asm volatile("" ::: "memory");
        pass &= ~slot_any<CF>(__ballot(layer_fails_d<CF, LI>(Ls, t)));
        if (pass == 0) return 0;
        return syndrome_pass_d<CF, LI + 1>(Ls, t, pass);
    }
}

// The shared-flag scan of scan_fails_from, on fp32 LLRs.
template <class CF, int LI>
__device__ __forceinline__ void scan_fails_from_d(const float *Ls, int t, volatile uint32_t *fail) {
    if constexpr (LI < NLAYERS) {
        asm volatile("" ::: "memory");
        if (__builtin_amdgcn_readfirstlane(*fail)) return;
        if (__ballot(layer_fails_d<CF, LI>(Ls, t)) != 0) {
            if ((t & 63) == 0) *fail = 1u;
            return;
        }
        scan_fails_from_d<CF, LI + 1>(Ls, t, fail);
    }
}

// Profile T channel (bf16 LLRs) as fp32 values.
template <class CF>
__device__ __forceinline__ void gen_channel_d(float *Ls, int t, uint64_t seed, uint64_t w, float mu, float sigma) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int b = t; b < CF::NTX / 4; b += CF::G) {
        const u32x4 x = philox4x32_10((uint32_t)w, (uint32_t)(w >> 32), (uint32_t)b, 0u, k0, k1);
        float z0, z1, z2, z3;
        box_muller(x.x, x.y, z0, z1);
        box_muller(x.z, x.w, z2, z3);
        *reinterpret_cast<float4 *>(Ls + 4 * b) =
            make_float4(u2f(llr_bits(mu, sigma, z0)), u2f(llr_bits(mu, sigma, z1)), u2f(llr_bits(mu, sigma, z2)),
                        u2f(llr_bits(mu, sigma, z3)));
    }
    for (int b = t; b < CF::NP / 4; b += CF::G)
        *reinterpret_cast<float4 *>(Ls + CF::NTX + 4 * b) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

template <class CF>
__device__ __forceinline__ void copy_in_d(float *Ls, const uint16_t *src, int t) {
    for (int i = t; i < CF::N; i += CF::G) Ls[i] = u2f((uint32_t)src[i] << 16);
}

template <class CF>
__device__ __forceinline__ void copy_out_d(uint16_t *dst, const float *Ls, int t) {
    for (int i = t; i < CF::N; i += CF::G) dst[i] = (uint16_t)(rne_bits(Ls[i]) >> 16);
}

template <class CF>
__device__ __forceinline__ uint32_t info_errors_d(const float *Ls, int t) {
    uint32_t be = 0;
    for (int i = t; i < CF::K; i += CF::G) be += Ls[i] <= 0.0f ? 1u : 0u;
    return be;
}

__device__ __forceinline__ int d_row(const MultiParams &P, int it) {
    return P.d_groups <= 1 ? 0 : (it < P.d_groups ? it : P.d_groups - 1);
}

template <class CF, bool THREE, bool CAP>
__global__ void __launch_bounds__(CF::THREADS) k_d_wave(MultiParams P) {
    __shared__ __attribute__((aligned(16))) float smem[CF::CW_PER_WG * CF::N];
    __shared__ uint32_t hist_s[HIST_LDS];
    const Run run = load_run(P);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int slot = lane / CF::G, t = lane % CF::G;
    float *Ls = smem + (wv * CF::SLOTS + slot) * CF::N;
    for (int i = threadIdx.x; i < HIST_LDS; i += CF::THREADS) hist_s[i] = 0u;
    __syncthreads();

    const uint32_t all_slots = (1u << CF::SLOTS) - 1u;
    const bool corr = P.d_corr != 0;
    const float rcap = P.d_rmax;
    uint32_t cw = run.n_cw;
    int it = 0;
    float R[CF::NR];
#pragma unroll
    for (int i = 0; i < CF::NR; ++i) R[i] = 0.0f;
    unsigned long long blocks = 0, fe = 0, be = 0, its = 0;
    uint32_t done = all_slots, conv = 0, active = 0;
    for (;;) {
        if (done) {
            const bool mine = (done >> slot) & 1u;
            if (mine && cw < run.n_cw) {
                const bool ok = (conv >> slot) & 1u;
                if (!ok) be += info_errors_d<CF>(Ls, t);
                if (P.llr_in) copy_out_d<CF>(P.llr_out + (size_t)cw * CF::N, Ls, t);
                if (t == 0) {
                    blocks += 1;
                    its += ok ? (unsigned)it : (unsigned)P.max_iter;
                    fe += ok ? 0u : 1u;
                    hist_add(hist_s, P, ok ? it : 0);
                    if (!ok && !P.llr_in) capture_failure(P.cap, run.first_cw + cw);
                    if (P.llr_in) {
                        P.iters[cw] = ok ? (uint32_t)it : (uint32_t)P.max_iter;
                        P.ok[cw] = ok ? 1 : 0;
                    }
                }
            }
            uint32_t nxt = run.n_cw;
            if (mine && t == 0) {
                nxt = min(atomicAdd(P.queue, 1u), run.n_cw);
            }
            nxt = __shfl(nxt, slot * CF::G);
            if (mine) {
                cw = nxt;
                it = 0;
                if (cw < run.n_cw) {
                    if (P.llr_in)
                        copy_in_d<CF>(Ls, P.llr_in + (size_t)cw * CF::N, t);
                    else
                        gen_channel_d<CF>(Ls, t, run.seed, run.first_cw + cw, run.mu, run.sigma);
                }
            }
#pragma unroll
            for (int i = 0; i < CF::NR; ++i) R[i] = mine ? 0.0f : R[i];
            active = slot_any<CF>(__ballot(cw < run.n_cw));
            wave_fence();
        }
        if (!active) break;
        const int row = d_row(P, it);
        run_iteration_d<CF, THREE, CAP>(Ls, R, t, P.d_alpha + row * CF::NR, P.d_beta + row * CF::NR, corr, rcap,
                            std::make_integer_sequence<int, NLAYERS>{});
        it += 1;
        conv = syndrome_pass_d<CF, 0>(Ls, t, active);
        done = (conv | slot_any<CF>(__ballot(it >= P.max_iter))) & active;
    }
    blocks = wave_sum(blocks);
    fe = wave_sum(fe);
    be = wave_sum(be);
    its = wave_sum(its);
    if (lane == 0) flush_counters(P, blocks, fe, be, its);
    __syncthreads();
    hist_flush(hist_s, P, CF::THREADS);
}

template <class CF, bool THREE, bool CAP>
__global__ void __launch_bounds__(CF::THREADS) k_d_group(MultiParams P) {
    __shared__ __attribute__((aligned(16))) float Ls[CF::N];
    __shared__ uint32_t hist_s[HIST_LDS];
    // fail_s[it & 1]: some check is unsatisfied after iteration it (two
    // flags, so that the one being set is never the one being read or cleared)
    __shared__ uint32_t fail_s[2];
    __shared__ uint32_t cw_s, be_s;
    const Run run = load_run(P);
    const int t = threadIdx.x;
    for (int i = t; i < HIST_LDS; i += CF::THREADS) hist_s[i] = 0u;
    const bool corr = P.d_corr != 0;
    const float rcap = P.d_rmax;
    unsigned long long blocks = 0, fe = 0, be = 0, its = 0;
    float R[CF::NR];
    __syncthreads();
    for (;;) {
        if (t == 0) cw_s = min(atomicAdd(P.queue, 1u), run.n_cw);
        __syncthreads();
        const uint32_t cw = cw_s;
        if (cw >= run.n_cw) break;
        if (P.llr_in)
            copy_in_d<CF>(Ls, P.llr_in + (size_t)cw * CF::N, t);
        else
            gen_channel_d<CF>(Ls, t, run.seed, run.first_cw + cw, run.mu, run.sigma);
        if (t == 0) {
            be_s = 0u;
            fail_s[0] = fail_s[1] = 0u;
        }
#pragma unroll
        for (int i = 0; i < CF::NR; ++i) R[i] = 0.0f;
        __syncthreads();
        int it = 0;
        bool ok = false;
        while (it < P.max_iter) {
            const int row = d_row(P, it);
            run_iteration_d<CF, THREE, CAP>(Ls, R, t, P.d_alpha + row * CF::NR, P.d_beta + row * CF::NR, corr, rcap,
                                std::make_integer_sequence<int, NLAYERS>{});
            ++it;
            scan_fails_from_d<CF, 0>(Ls, t, fail_s + (it & 1));
            __syncthreads();
            const uint32_t any = fail_s[it & 1];
            if (t == 0) fail_s[(it + 1) & 1] = 0u;  // read by every wave before this iteration's layers
            if (!any) {
                ok = true;
                break;
            }
        }
        if (!ok) {
            const uint32_t e = info_errors_d<CF>(Ls, t);
            if (e) atomicAdd(&be_s, e);
        }
        if (P.llr_in) copy_out_d<CF>(P.llr_out + (size_t)cw * CF::N, Ls, t);
        __syncthreads();
        if (t == 0) {
            blocks += 1;
            its += ok ? (unsigned)it : (unsigned)P.max_iter;
            if (!ok) {
                fe += 1;
                be += be_s;
            }
            hist_add(hist_s, P, ok ? it : 0);
            if (!ok && !P.llr_in) capture_failure(P.cap, run.first_cw + cw);
            if (P.llr_in) {
                P.iters[cw] = ok ? (uint32_t)it : (uint32_t)P.max_iter;
                P.ok[cw] = ok ? 1 : 0;
            }
        }
        __syncthreads();
    }
    if (t == 0) flush_counters(P, blocks, fe, be, its);
    __syncthreads();
    hist_flush(hist_s, P, CF::THREADS);
}

#if LDPC_KID == 2 && LDPC_RID == 0
// ---------------------------------------------------------------------------
// Profile D for K = 16384 r1/2 (z = 2048).  One codeword per 1024-thread
// workgroup, thread t updates checks t and t + 1024 of every layer.  The
// codeword's fp32 L (40960 x 4 B = 160 KiB) takes all of the CU's LDS, and
// its fp32 messages (2 x 60 per thread, 480 KiB per codeword) do not fit the
// 128 VGPRs a 1024-thread workgroup gets.  A check of degree above 3 keeps
// its messages as three words, from which they are rebuilt bit for bit
// with the previous iteration's parameter row: g at the first minimum
// (min2), g elsewhere (min1, or its corrected value), and the first
// minimum's index with the message signs; a degree-3 check keeps its three
// messages.  That is six words per thread and layer: six layers' words stay
// in registers, the other six go through a per-workgroup HBM workspace laid
// out [word][thread] (coalesced across the workgroup), 144 KiB read and
// written per codeword-iteration.  With the three-minimum correction a
// compressed check keeps four words: g of the edges equal to min1 (min*(min2,
// min3)), equal to min2 (min*(min1, min3)) and of the others (min*(min1,
// min2)), and the two equality masks with the signs (176 KiB per
// codeword-iteration; eight register words per compressed layer).  A workspace layer's words are loaded one
// layer ahead (each word belongs to one thread and one layer, so the
// prefetch never races a store), and iteration 0 uses R = 0 without
// reading, so neither the registers nor the workspace are cleared.  The workgroup-wide syndrome flag and the queue broadcast
// go through control words at the end of the workspace with workgroup-scope
// atomics (there is no LDS left for them); they are double-buffered by
// iteration parity, so one barrier per iteration suffices.
// ---------------------------------------------------------------------------
constexpr ar4ja::Schedule SDB = ar4ja::make_schedule(2, 0, 0);
constexpr int db_r_off(int li) {
    int o = 0;
    for (int i = 0; i < li; ++i) o += SDB.layer[i].deg;
    return o;
}
// workspace words per check of layer li: the raw messages up to degree 3,
// else the compressed form (three words, four with three minima)
constexpr int db_words(int li, bool three) { return SDB.layer[li].deg <= 3 ? SDB.layer[li].deg : (three ? 4 : 3); }
// Layers whose words go through the workspace; the others keep theirs in
// registers, six per thread and layer (the 128 VGPRs a 1024-thread
// workgroup allows hold six such layers without spilling in the loop;
// seven spill 18 dwords per iteration).  No two workspace layers are
// adjacent, cyclically, so one prefetch buffer serves them all, filled by
// the layer before each.
constexpr uint32_t DB_HBM = 0xAAAu;  // layers 1, 3, 5, 7, 9, 11
constexpr bool db_hbm(int li) { return (DB_HBM >> li) & 1u; }
constexpr int db_w_off(int li, bool three) {
    int o = 0;
    for (int i = 0; i < li; ++i) o += db_hbm(i) ? db_words(i, three) : 0;
    return o;
}
constexpr int db_reg_off(int li, bool three) {
    int o = 0;
    for (int i = 0; i < li; ++i) o += db_hbm(i) ? 0 : 2 * db_words(i, three);
    return o;
}
constexpr bool db_hbm_apart() {
    for (int i = 0; i < NLAYERS; ++i)
        if (db_hbm(i) && db_hbm((i + 1) % NLAYERS)) return false;
    return true;
}
static_assert(db_hbm_apart(), "workspace layers must not be adjacent");
constexpr bool db_words_fit() {
    for (int i = 0; i < NLAYERS; ++i)
        if (db_words(i, false) > 3 || db_words(i, true) > 4 || SDB.layer[i].deg > 6) return false;
    return true;
}
static_assert(db_words_fit(), "three (four) words per check, degrees up to 6");

template <bool THREE>
struct CfgDBT {
    static constexpr int C = 2, ORD = 0;
    static constexpr ar4ja::Schedule S = SDB;
    static constexpr int Z = S.Z, N = S.N, NP = S.n_punct, NTX = S.N - S.n_punct, K = S.k_info;
    static constexpr int THREADS = 1024, G = THREADS, WPW = THREADS / 64, CW_PER_WG = 1;
    static constexpr bool WAVE = false;
    static constexpr int r_off(int li) { return db_r_off(li); }
    static constexpr int w_off(int li) { return db_w_off(li, THREE); }
    static constexpr int reg_off(int li) { return db_reg_off(li, THREE); }
    static constexpr int words(int li) { return db_words(li, THREE); }
    static constexpr int NR = db_r_off(NLAYERS);               // messages per check = parameters per row
    static constexpr int NW = db_w_off(NLAYERS, THREE);        // workspace words per check
    static constexpr int NREGW = db_reg_off(NLAYERS, THREE);   // register-resident words per thread
    static constexpr int CTL = 2 * NW * THREADS;               // control words: flag[2], queue broadcast
    static constexpr int WORK_DWORDS = CTL + 64;               // per workgroup (control padded to 256 B)
    static constexpr int WMAX = THREE ? 4 : 3;  // workspace / register words per check and layer (at most)
    static_assert(Z == 2 * THREADS && N * 4 == 163840, "K = 16384 r1/2 layout");
};
using CfgDB = CfgDBT<false>;
// workspace stride per workgroup, the same for every instance (the larger layout)
constexpr int DB_WORK_STRIDE = CfgDBT<true>::WORK_DWORDS;

// Workgroup barrier that orders LDS only: the messages in HBM are private to
// their thread, so a layer need not wait for its message stores (or the next
// layer's prefetch) to reach memory before the next layer starts.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Control words are shared only by the waves of one workgroup (one CU), so
// their atomics are workgroup-scoped: agent scope would write back and
// invalidate the XCD's L2 on every release / acquire, and the L2 is full of
// dirty message lines.  A control word reads the same in every lane; saying
// so (readfirstlane) keeps the loop state that depends on it scalar.
__device__ __forceinline__ uint32_t ctl_load(const uint32_t *p) {
    return __builtin_amdgcn_readfirstlane(__hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
}
__device__ __forceinline__ void ctl_store(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Workspace words of layer LI for the thread's two checks: word (k, h) at
// Rbase[(2 (W0 + k) + h) * THREADS + t].
template <int LI, bool THREE>
__device__ __forceinline__ void load_r_db(const uint32_t *Rbase, int t, uint32_t (&buf)[2 * CfgDBT<THREE>::WMAX]) {
    constexpr int W = CfgDBT<THREE>::words(LI), W0 = CfgDBT<THREE>::w_off(LI);
    asm volatile("" : "+v"(t));  // addresses are formed here, per layer
    const uint32_t *Rw = Rbase + t;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        buf[2 * k] = Rw[(2 * (W0 + k)) * CfgDB::THREADS];
        buf[2 * k + 1] = Rw[(2 * (W0 + k) + 1) * CfgDB::THREADS];
    }
}

// Stores of layer LI's words from buf (the layout of load_r_db).
template <int LI, bool THREE>
__device__ __forceinline__ void store_r_db(uint32_t *Rbase, int t, const uint32_t (&buf)[2 * CfgDBT<THREE>::WMAX]) {
    constexpr int W = CfgDBT<THREE>::words(LI), W0 = CfgDBT<THREE>::w_off(LI);
    asm volatile("" : "+v"(t));
    uint32_t *Rw = Rbase + t;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        Rw[(2 * (W0 + k)) * CfgDB::THREADS] = buf[2 * k];
        Rw[(2 * (W0 + k) + 1) * CfgDB::THREADS] = buf[2 * k + 1];
    }
}

// Message magnitude from g and the edge's parameters (run_layer_d's
// arithmetic); the same function rebuilds a stored message.  With a cap
// (CAP: r_max > 0), min(mag, r_max) is run_layer_d's "r_max unless
// mag < r_max" for the non-negative finite magnitudes here.
template <bool CAP>
__device__ __forceinline__ float db_mag(float a, float b, float g, float r_max) {
    const float z = __fsub_rn(__fmul_rn(a, g), b);
    const float mag = z > 0.0f ? z : 0.0f;
    return CAP ? fminf(mag, r_max) : mag;
}

// One check of layer LI (index j): the arithmetic of run_layer_d, operation
// for operation.  w[h + 2 k] holds the check's workspace words from the
// previous iteration (ignored when first) and receives this iteration's.
// For a compressed layer, pa / pb are the previous iteration's parameters.
template <int LI, int H, bool CAP>
__device__ __forceinline__ void check_db(float *Ls, int j, uint32_t *w, bool first, const float *pa,
                                         const float *pb, const float *alv, const float *bev, bool corr,
                                         float rcap) {
    constexpr ar4ja::Layer Ly = CfgDB::S.layer[LI];
    constexpr int D = Ly.deg, Z = CfgDB::Z;
    constexpr bool PACK = D > 3;
    asm volatile("" : "+v"(j));  // keeps the shifted LDS addresses inside the layer (not hoisted, no VGPRs pinned)
    float r[D];
    if constexpr (PACK) {
        const float g1 = u2f(w[H]), go = u2f(w[H + 2]);
        const uint32_t pk = w[H + 4];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const float mag = db_mag<CAP>(pa[d], pb[d], (pk & 7u) == (uint32_t)d ? g1 : go, rcap);
            r[d] = first ? 0.0f : u2f(f2u(mag) | (((pk >> (3 + d)) & 1u) << 31));
        }
    } else {
#pragma unroll
        for (int d = 0; d < D; ++d) r[d] = first ? 0.0f : u2f(w[H + 2 * d]);
    }
    int v[D];
    float q[D];
    float m1 = __builtin_inff(), m2 = __builtin_inff();
    int i1 = 0;
    uint32_t s = 0;  // bit 31: product of the signs
#pragma unroll
    for (int d = 0; d < D; ++d) {
        v[d] = Ly.slot[d].base + ((j + Ly.slot[d].shift) & (Z - 1));
        q[d] = __fsub_rn(Ls[v[d]], r[d]);
        s ^= f2u(q[d]);
        // "if a < m1: (m2, m1, i1) = (m1, a, d) elif a < m2: m2 = a" without
        // branches: with m1 <= m2 the new m2 is the median of (a, m1, m2)
        const float a = fabsf(q[d]);
        i1 = a < m1 ? d : i1;
        m2 = __builtin_amdgcn_fmed3f(a, m1, m2);
        m1 = fminf(a, m1);
    }
    float other = m1;
    if (corr) {
        const float C1 = 0.69140625f, K1 = -0.3984375f, C2 = 0.65625f, K2 = -0.326171875f;
        const float x = __fsub_rn(m2, m1);
        const float l1 = __fadd_rn(__fmul_rn(K1, x), C1);
        const float l2 = __fadd_rn(__fmul_rn(K2, x), C2);
        const float lm = l1 >= l2 ? l1 : l2;
        const float o = __fsub_rn(m1, lm > 0.0f ? lm : 0.0f);
        other = o > 0.0f ? o : 0.0f;
    }
    uint32_t sbits = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float mag = db_mag<CAP>(alv[d], bev[d], d == i1 ? m2 : other, rcap);
        const uint32_t sw = s ^ f2u(q[d]);  // bit 31: the message's sign
        const float rr = u2f(f2u(mag) | (sw & 0x80000000u));
        Ls[v[d]] = __fadd_rn(q[d], rr);
        if constexpr (PACK)
            sbits |= (sw >> 31) << (3 + d);
        else
            w[H + 2 * d] = f2u(rr);
    }
    if constexpr (PACK) {
        w[H] = f2u(m2);
        w[H + 2] = f2u(other);
        w[H + 4] = sbits | (uint32_t)i1;
    }
}

// One check of layer LI with the three-minimum correction (run_layer_d3's
// arithmetic).  Compressed words: w[h] g of the edges equal to min1, w[h + 2]
// of those equal to min2, w[h + 4] of the others, w[h + 6] bit d: |q_d| ==
// min1, bit 6 + d: |q_d| == min2, bit 12 + d: the message's sign.  The
// rebuild selects g with the same equalities, so it reproduces the stored
// messages bit for bit.
template <int LI, int H, bool CAP>
__device__ __forceinline__ void check_db3(float *Ls, int j, uint32_t *w, bool first, const float *pa,
                                          const float *pb, const float *alv, const float *bev, float rcap) {
    constexpr ar4ja::Layer Ly = CfgDB::S.layer[LI];
    constexpr int D = Ly.deg, Z = CfgDB::Z;
    constexpr bool PACK = D > 3;
    static_assert(D <= 6, "packed masks hold six edges");
    asm volatile("" : "+v"(j));
    float r[D];
    if constexpr (PACK) {
        const float g1 = u2f(w[H]), g2 = u2f(w[H + 2]), go = u2f(w[H + 4]);
        const uint32_t pk = w[H + 6];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const float g = ((pk >> d) & 1u) ? g1 : (((pk >> (6 + d)) & 1u) ? g2 : go);
            const float mag = db_mag<CAP>(pa[d], pb[d], g, rcap);
            r[d] = first ? 0.0f : u2f(f2u(mag) | (((pk >> (12 + d)) & 1u) << 31));
        }
    } else {
#pragma unroll
        for (int d = 0; d < D; ++d) r[d] = first ? 0.0f : u2f(w[H + 2 * d]);
    }
    int v[D];
    float q[D];
    float m1 = __builtin_inff(), m2 = __builtin_inff(), m3 = __builtin_inff();
    uint32_t s = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        v[d] = Ly.slot[d].base + ((j + Ly.slot[d].shift) & (Z - 1));
        q[d] = __fsub_rn(Ls[v[d]], r[d]);
        s ^= f2u(q[d]);
        const float a = fabsf(q[d]);
        m3 = __builtin_amdgcn_fmed3f(a, m2, m3);
        m2 = __builtin_amdgcn_fmed3f(a, m1, m2);
        m1 = fminf(a, m1);
    }
    const float o12 = amin_d(m1, m2), o13 = amin_d(m1, m3), o23 = amin_d(m2, m3);
    uint32_t pk = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float a = fabsf(q[d]);
        const bool e1 = a == m1, e2 = a == m2;
        const float mag = db_mag<CAP>(alv[d], bev[d], e1 ? o23 : (e2 ? o13 : o12), rcap);
        const uint32_t sw = s ^ f2u(q[d]);  // bit 31: the message's sign
        const float rr = u2f(f2u(mag) | (sw & 0x80000000u));
        Ls[v[d]] = __fadd_rn(q[d], rr);
        if constexpr (PACK) {
            pk |= ((uint32_t)e1 << d) | ((uint32_t)e2 << (6 + d)) | ((sw >> 31) << (12 + d));
            // fold the two equality masks into pk here: left to the scheduler,
            // the twelve 64-bit masks of a check stay live to the end and
            // spill to VGPR lanes (not volatile: it need not stay ordered
            // with the other check's statements)
            asm("" : "+v"(pk));
        } else {
            w[H + 2 * d] = f2u(rr);
        }
    }
    if constexpr (PACK) {
        w[H] = f2u(o23);
        w[H + 2] = f2u(o13);
        w[H + 4] = f2u(o12);
        w[H + 6] = pk;
    }
}

// Layer LI: workspace words arrive in cur (loaded one layer earlier;
// ignored in iteration 0): a workspace layer's in buf, loaded by the layer
// before it, which requests them before its own compute; a register layer's
// in Rr.  al / be are this iteration's parameter row, pal / pbe the previous
// iteration's.
template <int LI, bool CAP, bool THREE>
__device__ __forceinline__ void layer_db(float *Ls, uint32_t *Rbase, int t, uint32_t (&buf)[2 * CfgDBT<THREE>::WMAX],
                                         uint32_t (&Rr)[CfgDBT<THREE>::NREGW], bool first, const float *al,
                                         const float *be, const float *pal, const float *pbe, bool corr,
                                         float rcap) {
    using CF = CfgDBT<THREE>;
    constexpr int D = CF::S.layer[LI].deg, R0 = CF::r_off(LI), W = CF::words(LI);
    constexpr bool PACK = D > 3, HBM = db_hbm(LI);
    constexpr int LN = LI + 1 < NLAYERS ? LI + 1 : 0;
    // The parameter rows are read through the constant address space, i.e.
    // with scalar loads: a vector load here would wait (the vector-memory
    // counter covers stores too, in order) for the previous layer's message
    // stores to be acknowledged, which serialises every layer on a write
    // round trip.
    typedef const __attribute__((address_space(4))) float cfloat;
    const cfloat *alc = (const cfloat *)al, *bec = (const cfloat *)be;
    const cfloat *palc = (const cfloat *)pal, *pbec = (const cfloat *)pbe;
    float alv[D], bev[D], pa[PACK ? D : 1], pb[PACK ? D : 1];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        alv[d] = alc[R0 + d];
        bev[d] = bec[R0 + d];
    }
    if constexpr (PACK) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            pa[d] = palc[R0 + d];
            pb[d] = pbec[R0 + d];
        }
    }
    auto check = [&](int j, uint32_t *w, auto h) {
        constexpr int HH = decltype(h)::value;
        if constexpr (THREE)
            check_db3<LI, HH, CAP>(Ls, j, w, first, pa, pb, alv, bev, rcap);
        else
            check_db<LI, HH, CAP>(Ls, j, w, first, pa, pb, alv, bev, corr, rcap);
    };
    if constexpr (HBM) {
        check(t, buf, std::integral_constant<int, 0>{});
        check(t + CF::THREADS, buf, std::integral_constant<int, 1>{});
        store_r_db<LI, THREE>(Rbase, t, buf);
    } else {
        // unconditional (iteration 0 loads stale words that the next layer
        // ignores, and a branch would only add waits)
        if constexpr (db_hbm(LN)) load_r_db<LN, THREE>(Rbase, t, buf);
        constexpr int O = CF::reg_off(LI);
        uint32_t w[2 * CF::WMAX];
#pragma unroll
        for (int i = 0; i < 2 * W; ++i) w[i] = Rr[O + i];
        check(t, w, std::integral_constant<int, 0>{});
        check(t + CF::THREADS, w, std::integral_constant<int, 1>{});
#pragma unroll
        for (int i = 0; i < 2 * W; ++i) Rr[O + i] = w[i];
    }
    lds_barrier();
}

template <bool CAP, bool THREE, int... LI>
__device__ __forceinline__ void iteration_db(float *Ls, uint32_t *Rbase, int t,
                                             uint32_t (&buf)[2 * CfgDBT<THREE>::WMAX],
                                             uint32_t (&Rr)[CfgDBT<THREE>::NREGW], bool first, const float *al,
                                             const float *be, const float *pal, const float *pbe, bool corr,
                                             float rcap, std::integer_sequence<int, LI...>) {
    (layer_db<LI, CAP, THREE>(Ls, Rbase, t, buf, Rr, first, al, be, pal, pbe, corr, rcap), ...);
}

template <int LI>
__device__ __forceinline__ bool fails_from_db(const float *Ls, int t) {
    if constexpr (LI == NLAYERS) {
        return false;
    } else {
        asm volatile("" ::: "memory");
        const bool f = layer_fails_d<CfgDB, LI>(Ls, t) || layer_fails_d<CfgDB, LI>(Ls, t + CfgDB::THREADS);
        if (__ballot(f) != 0) return true;
        return fails_from_db<LI + 1>(Ls, t);
    }
}

// CAP: r_max > 0; CORR: the two-minimum min* correction of g; THREE: the
// three-minimum correction (all fixed per launch)
template <bool CAP, bool CORR, bool THREE>
__global__ void __launch_bounds__(CfgDB::THREADS) k_d_big(MultiParams P) {
    using CF = CfgDBT<THREE>;
    __shared__ __attribute__((aligned(16))) float Ls[CfgDB::N];  // all 160 KiB of LDS
    const Run run = load_run(P);
    const int t = threadIdx.x, lane = t & 63;
    uint32_t *work = P.work + (size_t)blockIdx.x * DB_WORK_STRIDE;
    uint32_t *Rbase = work;
    uint32_t *ctl = work + CF::CTL;  // [0], [1]: syndrome flags by iteration parity; [2]: queue broadcast
    constexpr bool corr = CORR;
    const float rcap = P.d_rmax;
    unsigned long long blocks = 0, fe = 0, its = 0;
    uint32_t buf[2 * CF::WMAX], Rr[CF::NREGW];
#pragma unroll
    for (int i = 0; i < 2 * CF::WMAX; ++i) buf[i] = 0u;
#pragma unroll
    for (int i = 0; i < CF::NREGW; ++i) Rr[i] = 0u;
    for (;;) {
        if (t == 0) ctl_store(ctl + 2, min(atomicAdd(P.queue, 1u), run.n_cw));
        __syncthreads();
        const uint32_t cw = ctl_load(ctl + 2);
        if (cw >= run.n_cw) break;
        if (P.llr_in)
            copy_in_d<CfgDB>(Ls, P.llr_in + (size_t)cw * CfgDB::N, t);
        else
            gen_channel_d<CfgDB>(Ls, t, run.seed, run.first_cw + cw, run.mu, run.sigma);
        if (t == 0) {
            ctl_store(ctl + 0, 0u);
            ctl_store(ctl + 1, 0u);
        }
        __syncthreads();
        int it = 0;
        bool ok = false;
        while (it < P.max_iter) {
            // the previous iteration's row rebuilds the stored messages
            const int row = d_row(P, it), prow = d_row(P, it > 0 ? it - 1 : 0);
            iteration_db<CAP, THREE>(Ls, Rbase, t, buf, Rr, it == 0, P.d_alpha + row * CfgDB::NR,
                                     P.d_beta + row * CfgDB::NR,
                         P.d_alpha + prow * CfgDB::NR, P.d_beta + prow * CfgDB::NR, corr, rcap,
                         std::make_integer_sequence<int, NLAYERS>{});
            ++it;
            // flag of this iteration; the other parity's was read an iteration
            // (12 barriers) ago and is cleared for the next one
            const bool f = fails_from_db<0>(Ls, t);
            if (lane == 0 && f) __hip_atomic_fetch_or(ctl + (it & 1), 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (t == 0) ctl_store(ctl + ((it + 1) & 1), 0u);
            __syncthreads();
            if (ctl_load(ctl + (it & 1)) == 0u) {
                ok = true;
                break;
            }
        }
        if (!ok) {
            const unsigned long long e = wave_sum((unsigned long long)info_errors_d<CfgDB>(Ls, t));
            if (lane == 0 && e) atomicAdd(P.counters + 2, e);
        }
        if (P.llr_in) copy_out_d<CfgDB>(P.llr_out + (size_t)cw * CfgDB::N, Ls, t);
        if (t == 0) {
            blocks += 1;
            its += ok ? (unsigned)it : (unsigned)P.max_iter;
            fe += ok ? 0u : 1u;
            atomicAdd(P.hist + (ok ? it : 0), 1ull);
            if (!ok && !P.llr_in) capture_failure(P.cap, run.first_cw + cw);
            if (P.llr_in) {
                P.iters[cw] = ok ? (uint32_t)it : (uint32_t)P.max_iter;
                P.ok[cw] = ok ? 1 : 0;
            }
        }
        __syncthreads();  // Ls is rewritten by the next codeword
    }
    if (t == 0) flush_counters(P, blocks, fe, 0ull, its);
}

hipError_t launch_db(const MultiParams &p, int grid, hipStream_t st) {
    const bool cap = p.d_rmax > 0.0f;
    auto k = p.d_corr == 2 ? (cap ? &k_d_big<true, true, true> : &k_d_big<false, true, true>)
             : p.d_corr    ? (cap ? &k_d_big<true, true, false> : &k_d_big<false, true, false>)
                           : (cap ? &k_d_big<true, false, false> : &k_d_big<false, false, false>);
    hipLaunchKernelGGL(k, dim3(grid), dim3(CfgDB::THREADS), 0, st, p);
    return hipGetLastError();
}

// the six instances share the launch shape and the LDS that bounds residency
hipError_t occupancy_db(int *wg_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, k_d_big<true, true, false>, CfgDB::THREADS, 0);
}

hipError_t attributes_db(hipFuncAttributes *a) {
    return hipFuncGetAttributes(a, reinterpret_cast<const void *>(k_d_big<true, true, false>));
}
#endif

template <class CF, bool THREE = false, bool CAP = false>
constexpr auto kernel_of_d() {
    if constexpr (CF::WAVE)
        return &k_d_wave<CF, THREE, CAP>;
    else
        return &k_d_group<CF, THREE, CAP>;
}

// Instances: three-minimum or not, message cap or not.  The grid is sized
// by the residency of the uncapped plain / two-minimum instance; where
// another instance holds fewer workgroups per CU, the ones beyond its
// residency start as others finish and leave at their first queue draw once
// the queue is empty (no workgroup waits on another).
template <class CF>
hipError_t launch_d(const MultiParams &p, int grid, hipStream_t st) {
    const bool cap = p.d_rmax > 0.0f;
    auto k = p.d_corr == 2 ? (cap ? kernel_of_d<CF, true, true>() : kernel_of_d<CF, true, false>())
                           : (cap ? kernel_of_d<CF, false, true>() : kernel_of_d<CF, false, false>());
    hipLaunchKernelGGL(k, dim3(grid), dim3(CF::THREADS), 0, st, p);
    return hipGetLastError();
}

template <class CF>
hipError_t occupancy_d(int *wg_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, kernel_of_d<CF>(), CF::THREADS, 0);
}

template <class CF>
hipError_t attributes_d(hipFuncAttributes *a) {
    return hipFuncGetAttributes(a, reinterpret_cast<const void *>(kernel_of_d<CF>()));
}

template <class CF>
constexpr MultiVariant variant() {
    return MultiVariant{LDPC_KID, LDPC_RID, CF::C, CF::ORD, CF::THREADS, CF::CW_PER_WG,
                        &launch<CF>, &occupancy<CF>, &attributes<CF>, 0};
}

template <class CF>
constexpr MultiVariant variant_d() {
    return MultiVariant{LDPC_KID, LDPC_RID, CF::C, CF::ORD, CF::THREADS, CF::CW_PER_WG,
                        &launch_d<CF>, &occupancy_d<CF>, &attributes_d<CF>, 1};
}

// C = 1 needs Z <= 1024 threads per codeword (all codes but K = 16384 r1/2).
constexpr bool C1_OK = ar4ja::make_schedule(LDPC_KID, LDPC_RID).Z <= 1024;

const MultiVariant TABLE[] = {
    variant<Cfg<2, 0>>(),
    variant<Cfg<2, 1>>(),
#if !(LDPC_KID == 2 && LDPC_RID == 0)
    variant<Cfg<1, 0>>(),
    variant<Cfg<1, 1>>(),
#endif
};

// Profile D: one check per thread, except K = 16384 r1/2 (z = 2048), which
// has its own kernel with the messages in an HBM workspace
#if !(LDPC_KID == 2 && LDPC_RID == 0)
const MultiVariant TABLE_D[] = {variant_d<CfgD<0>>()};
#else
const MultiVariant TABLE_D[] = {MultiVariant{2, 0, CfgDB::C, CfgDB::ORD, CfgDB::THREADS, 1, &launch_db, &occupancy_db,
                                             &attributes_db, 1, sizeof(uint32_t) * (size_t)DB_WORK_STRIDE}};
#endif
static_assert(C1_OK == !(LDPC_KID == 2 && LDPC_RID == 0), "C = 1 availability");

}  // namespace

#define LDPC_CAT3(a, b, c, d) a##b##c##d
#define LDPC_TABLE_NAME(k, r) LDPC_CAT3(multi_table_k, k, _r, r)
#define LDPC_TABLE_D_NAME(k, r) LDPC_CAT3(multi_table_d_k, k, _r, r)

const MultiVariant *LDPC_TABLE_NAME(LDPC_KID, LDPC_RID)(int *n) {
    *n = (int)(sizeof(TABLE) / sizeof(TABLE[0]));
    return TABLE;
}

const MultiVariant *LDPC_TABLE_D_NAME(LDPC_KID, LDPC_RID)() { return &TABLE_D[0]; }

}  // namespace ldpc
