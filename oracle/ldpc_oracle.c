/*
 * ldpc_oracle.c -- CPU oracle for the Monte Carlo LDPC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP product
 * path is compared against; it is loaded only by tests/, by the smoke() entry
 * and by bench.py's cpu_baseline leg.  Product code never links or calls it.
 *
 * It restates, in plain C with every fp32 operation separately rounded (built
 * with -ffp-contract=off, no -ffast-math), the reference's golden layered
 * decoder ("Profile T"):
 *
 *   - tree Approximate-Min* check update:
 *       tests/test_circulant_dma_stepper.cpp:107-261   (degree 6)
 *       tests/test_circulant_dma_stepper.cpp:266-327   (degree 3)
 *       engines/tile_matrix_4cw/sfpi_host_emu.h:566-584 (branchless min*)
 *       engines/tile_matrix_4cw/sfpi_host_emu.h:30-58   (RNE, truncating consts)
 *   - layered schedule, syndrome early exit, convergence latch, tallies:
 *       tests/test_circulant_dma_stepper.cpp:743-815
 *   - dump replay (per-codeword iteration of convergence):
 *       tests/test_circulant_dma_stepper.cpp:861-990
 *
 * and the "Profile R" recursive decoder that run_simulation.py drives
 * (engines/scalar_5w/ldpc_decoder_core.h:155-201, 214-586): unrounded q,
 * min / argmin with the first index winning ties, recursive min* with full
 * fp32 constants (or normalized min-sum, alpha 0.75), optional message and
 * posterior clamps, syndrome on the raw sign bit from the third iteration on,
 * frame error <=> a transmitted sign bit set.
 *
 * The channel generator (Philox4x32-10 counter-based stream, Salmon et al.,
 * SC'11, then Box-Muller 1958 with polynomial log/sin/cos and a Newton
 * reciprocal square root) is this project's own; the reference uses other
 * generators (compute_trisc.cpp:179-294, ldpc_decoder_core.h:128-152,
 * test_circulant_dma_stepper.cpp:735-757), so its parity with the reference
 * is statistical.  It is restated here op for op so that the GPU stream and
 * this one are bit-identical.
 *
 * Pins (see tests/): golden fixtures written by the reference stepper built
 * from source (oracle/_ref), the closed form of min*, Random123's Philox
 * known-answer vectors, and noise moments / tail counts against Q(x).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* bf16 helpers (ldpc_tile_matrix_core.h:432-446; sfpi_host_emu.h:30-40)     */
/* ------------------------------------------------------------------------- */
static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline float bf2f(uint16_t h) { return u2f((uint32_t)h << 16); }
/* Round-to-nearest-even to bf16 by integer bias: u + 0x7FFF + lsb. */
static inline uint16_t f2bf(float f) {
    uint32_t u = f2u(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static inline float rne(float f) { return bf2f(f2bf(f)); }
/* bf16 -> hard decision "value < 0.0f" (stepper:781): sign set and not -0. */
static inline int hard_neg(uint16_t h) { return h > 0x8000u; }
/* bf16 "value <= 0.0f" (stepper:806). */
static inline int nonpos(uint16_t h) { return (h & 0x8000u) || h == 0; }

/* std::min / std::max semantics used by the reference emulator. */
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline float smax(float a, float b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------- */
/* Profile T: tree Approx-Min* with bf16-truncated constants                 */
/* sfpi_host_emu.h:577-583: line1 = sFloat16b(0.69315) + sFloat16b(-0.4)*d,  */
/* line2 = sFloat16b(0.65715) + sFloat16b(-0.328)*d, sFloat16b truncates     */
/* the fp32 pattern to its top 16 bits (sfpi_host_emu.h:49-58).              */
/* ------------------------------------------------------------------------- */
static inline float trunc_bf(float f) { return u2f(f2u(f) & 0xFFFF0000u); }

EXPORT float orc_amin_t(float a, float b) {
    const float c1 = trunc_bf(0.69315f), k1 = trunc_bf(-0.4000f);
    const float c2 = trunc_bf(0.65715f), k2 = trunc_bf(-0.3280f);
    float mn = smin(a, b), mx = smax(a, b);
    float d = mx - mn;
    float p1 = k1 * d;
    float l1 = c1 + p1;
    float p2 = k2 * d;
    float l2 = c2 + p2;
    float corr = smax(smax(l1, l2), 0.0f);
    return smax(mn - corr, 0.0f);
}

/* Tree min* of the other edges for an even degree D = 2h, generalising the
 * golden degree-6 tree (stepper:151-225) to the degree-10 and degree-18
 * checks of the rate-2/3 and rate-4/5 AR4JA codes, which the reference's
 * Profile T never decodes (its golden and silicon paths are rate 1/2 only):
 * edge j is paired with edge j + h; p_j = A(m_j, m_{j+h}); the combination of
 * all pairs but j is formed from a prefix P_j = A(P_{j-1}, p_j) and a suffix
 * S_j = A(p_j, S_{j+1}) as o_j = A(P_{j-1}, S_{j+1}) (o_0 = S_1,
 * o_{h-1} = P_{h-2}); then r_j = A(o_j, m_{j+h}) beta and
 * r_{j+h} = A(o_j, m_j) beta.  For h = 3 these are exactly the golden's
 * o0 = A(p14, p25), o1 = A(p03, p25), o2 = A(p03, p14) (A is symmetric), so
 * degree 6 is the golden tree.                                              */
EXPORT void orc_tree_even(const float *m, int h, float beta, float *r) {
    float p[9], P[9], S[9];
    for (int j = 0; j < h; ++j) p[j] = orc_amin_t(m[j], m[j + h]);
    P[0] = p[0];
    for (int j = 1; j < h - 1; ++j) P[j] = orc_amin_t(P[j - 1], p[j]);
    S[h - 1] = p[h - 1];
    for (int j = h - 2; j >= 1; --j) S[j] = orc_amin_t(p[j], S[j + 1]);
    for (int j = 0; j < h; ++j) {
        float o = j == 0 ? S[1] : j == h - 1 ? P[h - 2] : orc_amin_t(P[j - 1], S[j + 1]);
        r[j] = orc_amin_t(o, m[j + h]) * beta;
        r[j + h] = orc_amin_t(o, m[j]) * beta;
    }
}

/* The golden degree-6 tree as the stepper writes it (stepper:151-225). */
EXPORT void orc_tree6(const float *m, float beta, float *r) {
    float p03 = orc_amin_t(m[0], m[3]), p14 = orc_amin_t(m[1], m[4]), p25 = orc_amin_t(m[2], m[5]);
    float o0 = orc_amin_t(p14, p25);          /* :174 */
    float o1 = orc_amin_t(p03, p25);          /* :199 */
    float o2 = orc_amin_t(p03, p14);          /* :224 */
    r[0] = orc_amin_t(o0, m[3]) * beta;  r[3] = orc_amin_t(o0, m[0]) * beta;
    r[1] = orc_amin_t(o1, m[4]) * beta;  r[4] = orc_amin_t(o1, m[1]) * beta;
    r[2] = orc_amin_t(o2, m[5]) * beta;  r[5] = orc_amin_t(o2, m[2]) * beta;
}

#define ORC_MAX_DEG_T 18

/* One Profile T check update, in place on L (per variable) and R (per edge).
 * deg is 3 or even (4..18); v[] are the check's variables in ascending order;
 * iter0 selects q = L (R treated as 0, stepper:346-348 and :522-531).      */
static void check_update_t(uint16_t *L, uint16_t *R, const int32_t *v, int deg, int iter0, float beta) {
    float q[ORC_MAX_DEG_T], m[ORC_MAX_DEG_T], r[ORC_MAX_DEG_T];
    uint32_t s = 0;
    for (int d = 0; d < deg; ++d) {
        float l = bf2f(L[v[d]]);
        q[d] = iter0 ? l : rne(l - bf2f(R[d]));   /* compute_tile_deg6_sfpi:112-116 */
        s ^= f2u(q[d]);                             /* sign parity, :161 */
        m[d] = fabsf(q[d]);
    }
    if (deg == 6) {
        orc_tree6(m, beta, r);
    } else if (deg == 3) {
        r[0] = orc_amin_t(m[1], m[2]);  r[1] = orc_amin_t(m[0], m[2]);  r[2] = orc_amin_t(m[0], m[1]);
        if (beta < 1.0f) { r[0] *= beta; r[1] *= beta; r[2] *= beta; }   /* :300-304 */
    } else {
        orc_tree_even(m, deg / 2, beta, r);
    }
    for (int d = 0; d < deg; ++d) {
        /* copysgn(r, s ^ q_d) then RNE (sfpi_host_emu.h:466-476; stepper:178-186) */
        uint32_t sign = (s ^ f2u(q[d])) & 0x80000000u;
        float out = rne(u2f((f2u(r[d]) & 0x7FFFFFFFu) | sign));
        R[d] = f2bf(out);
        L[v[d]] = f2bf(q[d] + out);
    }
}

/* One check update on explicit values (unit tests of the tree): L_in/R_in
 * are the check's deg variable and message values; results in L_out/R_out. */
EXPORT int orc_check_update_t(int deg, const uint16_t *L_in, const uint16_t *R_in, int iter0, float beta,
                              uint16_t *L_out, uint16_t *R_out) {
    if (deg != 3 && (deg < 4 || deg > ORC_MAX_DEG_T || (deg & 1))) return -1;
    int32_t v[ORC_MAX_DEG_T];
    for (int d = 0; d < deg; ++d) { v[d] = d; L_out[d] = L_in[d]; R_out[d] = R_in[d]; }
    check_update_t(L_out, R_out, v, deg, iter0, beta);
    return 0;
}

/* Syndrome over all checks with hard bit = (bf16 < 0.0f) (stepper:773-796). */
static int syndrome_ok(const uint16_t *L, int m, const int32_t *rp, const int32_t *col) {
    for (int i = 0; i < m; ++i) {
        int p = 0;
        for (int e = rp[i]; e < rp[i + 1]; ++e) p ^= hard_neg(L[col[e]]);
        if (p) return 0;
    }
    return 1;
}

static int unsat_checks(const uint16_t *L, int m, const int32_t *rp, const int32_t *col) {
    int u = 0;
    for (int i = 0; i < m; ++i) {
        int p = 0;
        for (int e = rp[i]; e < rp[i + 1]; ++e) p ^= hard_neg(L[col[e]]);
        u += p;
    }
    return u;
}

/* Decode one codeword in place.  Returns iterations used (1..max_iter) on
 * convergence, or 0 if the syndrome never cleared.  R has n_edges entries,
 * indexed like col[].  sched lists checks in decoding order.                */
static int decode_one_t(uint16_t *L, uint16_t *R, int m, const int32_t *rp, const int32_t *col,
                        const int32_t *sched, int max_iter, float beta) {
    for (int e = 0; e < rp[m]; ++e) R[e] = 0;
    for (int it = 0; it < max_iter; ++it) {
        for (int s = 0; s < m; ++s) {
            int i = sched[s];
            check_update_t(L, R + rp[i], col + rp[i], rp[i + 1] - rp[i], it == 0, beta);
        }
        if (syndrome_ok(L, m, rp, col)) return it + 1;
    }
    return 0;
}

static int check_degrees_t(int m, const int32_t *rp) {
    for (int i = 0; i < m; ++i) {
        int d = rp[i + 1] - rp[i];
        if (d != 3 && (d < 4 || d > ORC_MAX_DEG_T || (d & 1))) return -1;
    }
    return 0;
}

/* Decode n_cw codewords of n variables each (bf16 LLRs, codeword-major).
 * out_iters[c] = iteration of convergence or max_iter if it never converged
 * (stepper:760,789).  out_ok[c] = 1 if the syndrome cleared.                */
EXPORT int orc_decode_t(int n, int m, const int32_t *rp, const int32_t *col, const int32_t *sched,
                        int n_cw, const uint16_t *llr_in, int max_iter, float beta,
                        uint16_t *llr_out, uint32_t *out_iters, uint8_t *out_ok, int threads) {
    if (check_degrees_t(m, rp)) return -1;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
#endif
    {
        uint16_t *R = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)rp[m]);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int c = 0; c < n_cw; ++c) {
            uint16_t *L = llr_out + (size_t)c * n;
            memcpy(L, llr_in + (size_t)c * n, sizeof(uint16_t) * n);
            int it = decode_one_t(L, R, m, rp, col, sched, max_iter, beta);
            out_iters[c] = it ? (uint32_t)it : (uint32_t)max_iter;
            out_ok[c] = it ? 1 : 0;
        }
        free(R);
    }
    return 0;
}

/* Per-iteration trace of one codeword: unsatisfied checks after each
 * iteration (the "Unsat=" column of stepper:906-938).  Returns iterations run. */
EXPORT int orc_trace_t(int n, int m, const int32_t *rp, const int32_t *col, const int32_t *sched,
                       const uint16_t *llr_in, int max_iter, float beta, uint32_t *unsat_per_iter,
                       uint16_t *llr_out) {
    if (check_degrees_t(m, rp)) return -1;
    uint16_t *R = (uint16_t *)calloc((size_t)rp[m], sizeof(uint16_t));
    memcpy(llr_out, llr_in, sizeof(uint16_t) * n);
    int it;
    for (it = 0; it < max_iter; ++it) {
        for (int s = 0; s < m; ++s) {
            int i = sched[s];
            check_update_t(llr_out, R + rp[i], col + rp[i], rp[i + 1] - rp[i], it == 0, beta);
        }
        unsat_per_iter[it] = (uint32_t)unsat_checks(llr_out, m, rp, col);
        if (unsat_per_iter[it] == 0) { ++it; break; }
    }
    free(R);
    return it;
}

/* ------------------------------------------------------------------------- */
/* Profile R: recursive Approx-Min* (engines/scalar_5w/ldpc_decoder_core.h)  */
/* ------------------------------------------------------------------------- */
/* Correction g(diff) with full fp32 constants (ldpc_decoder_core.h:162,192). */
static inline float corr_r(float diff) {
    return (diff >= 2.0f) ? 0.0f : ((diff < 0.5f) ? (0.69315f - 0.40f * diff) : (0.49315f - 0.328f * (diff - 0.5f)));
}
static inline float absbits(float x) { return u2f(f2u(x) & 0x7FFFFFFFu); }

/* ldpc_decoder_core.h:155-173 */
static void amin_r_deg3(const float *q, uint32_t min_idx, float min_mag, float *r_min, float *r_all) {
    float o0 = (min_idx == 0) ? absbits(q[1]) : absbits(q[0]);
    float o1 = (min_idx == 2) ? absbits(q[1]) : absbits(q[2]);
    float diff = (o0 > o1) ? (o0 - o1) : (o1 - o0);
    float m_min = (o0 < o1) ? o0 : o1;
    float corr = corr_r(diff);
    float m_oth = (m_min > corr) ? (m_min - corr) : 0.0f;
    float diff_all = (m_oth > min_mag) ? (m_oth - min_mag) : (min_mag - m_oth);
    float m_all_min = (m_oth < min_mag) ? m_oth : min_mag;
    float corr_all = corr_r(diff_all);
    *r_min = m_oth;
    *r_all = (m_all_min > corr_all) ? (m_all_min - corr_all) : 0.0f;
}

/* ldpc_decoder_core.h:176-201 */
static void amin_r_general(const float *q, uint32_t deg, uint32_t min_idx, float min_mag, float *r_min, float *r_all) {
    uint32_t first = (min_idx == 0) ? 1u : 0u;
    float m = (deg > 1) ? absbits(q[first]) : min_mag;
    for (uint32_t d = first + 1; d < deg; ++d) {
        if (d == min_idx) continue;
        float a = absbits(q[d]);
        float diff = (m > a) ? (m - a) : (a - m);
        float cur = (m < a) ? m : a;
        float nx = cur - corr_r(diff);
        m = (nx > 0.0f) ? nx : 0.0f;
    }
    float diff_all = (m > min_mag) ? (m - min_mag) : (min_mag - m);
    float cur = (m < min_mag) ? m : min_mag;
    float nx = cur - corr_r(diff_all);
    *r_min = m;
    *r_all = (nx > 0.0f) ? nx : 0.0f;
}

typedef struct {
    int nms;            /* 0: Approx-Min*; 1: normalized min-sum, alpha 0.75 */
    float r_max;        /* check message clamp, 0 = off */
    float l_post_max;   /* posterior clamp, 0 = off */
} orc_cfg_r;

/* One Profile R row update on one codeword (ldpc_decoder_core.h:413-520,
 * NMS :320-411).  v[] in file order, R[] the row's messages. */
static void row_update_r(uint16_t *L, uint16_t *R, const int32_t *v, uint32_t deg, const orc_cfg_r *cfg) {
    float q[32];
    uint32_t sign = 0, idx = 0;
    float r_min, r_all;
    if (cfg->nms) {
        float min1 = 999.0f, min2 = 999.0f;
        for (uint32_t d = 0; d < deg; ++d) {
            q[d] = bf2f(L[v[d]]) - bf2f(R[d]);
            sign ^= f2u(q[d]) >> 31;
            float a = absbits(q[d]);
            if (a < min1) { min2 = min1; min1 = a; idx = d; }
            else if (a < min2) { min2 = a; }
        }
        r_min = 0.75f * min2;
        r_all = 0.75f * min1;
    } else {
        float min_mag = 9999.0f;
        for (uint32_t d = 0; d < deg; ++d) {
            q[d] = bf2f(L[v[d]]) - bf2f(R[d]);
            sign ^= f2u(q[d]) >> 31;
            float a = absbits(q[d]);
            if (a < min_mag) { min_mag = a; idx = d; }
        }
        if (deg == 3) amin_r_deg3(q, idx, min_mag, &r_min, &r_all);
        else amin_r_general(q, deg, idx, min_mag, &r_min, &r_all);
    }
    for (uint32_t d = 0; d < deg; ++d) {
        float mag = (d == idx) ? r_min : r_all;
        if (cfg->r_max > 0.0f && mag > cfg->r_max) mag = cfg->r_max;
        uint32_t msg_sign = sign ^ (f2u(q[d]) >> 31);
        float rn = u2f(f2u(mag) | (msg_sign << 31));
        float ln = q[d] + rn;
        if (cfg->l_post_max > 0.0f) {
            if (ln > cfg->l_post_max) ln = cfg->l_post_max;
            else if (ln < -cfg->l_post_max) ln = -cfg->l_post_max;
        }
        L[v[d]] = f2bf(ln);
        R[d] = f2bf(rn);
    }
}

/* Syndrome with hard bit = bf16 sign bit (ldpc_decoder_core.h:526-553). */
static int syndrome_ok_r(const uint16_t *L, int m, const int32_t *rp, const int32_t *col) {
    for (int i = 0; i < m; ++i) {
        uint32_t p = 0;
        for (int e = rp[i]; e < rp[i + 1]; ++e) p ^= L[col[e]] >> 15;
        if (p & 1u) return 0;
    }
    return 1;
}

/* Decode one codeword; rows in index order.  Returns the iteration at which
 * the syndrome cleared (checked from the third iteration on), or 0. */
static int decode_one_r(uint16_t *L, uint16_t *R, int m, const int32_t *rp, const int32_t *col, int max_iter,
                        const orc_cfg_r *cfg) {
    for (int e = 0; e < rp[m]; ++e) R[e] = 0;
    for (int it = 0; it < max_iter; ++it) {
        for (int i = 0; i < m; ++i) row_update_r(L, R + rp[i], col + rp[i], (uint32_t)(rp[i + 1] - rp[i]), cfg);
        if (it >= 2 && syndrome_ok_r(L, m, rp, col)) return it + 1;
    }
    return 0;
}

static int check_degrees_r(int m, const int32_t *rp) {
    for (int i = 0; i < m; ++i) {
        int d = rp[i + 1] - rp[i];
        if (d < 1 || d > 12) return -1;   /* ldpc_sim_scalar.cpp:34,187 */
    }
    return 0;
}

EXPORT int orc_decode_r(int n, int m, const int32_t *rp, const int32_t *col, int n_cw, const uint16_t *llr_in,
                        int max_iter, int nms, float r_max, float l_post_max, uint16_t *llr_out, uint32_t *out_iters,
                        uint8_t *out_ok, int threads) {
    if (check_degrees_r(m, rp)) return -1;
    orc_cfg_r cfg = {nms, r_max, l_post_max};
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
#endif
    {
        uint16_t *R = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)rp[m]);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int c = 0; c < n_cw; ++c) {
            uint16_t *L = llr_out + (size_t)c * n;
            memcpy(L, llr_in + (size_t)c * n, sizeof(uint16_t) * n);
            int it = decode_one_r(L, R, m, rp, col, max_iter, &cfg);
            out_iters[c] = it ? (uint32_t)it : (uint32_t)max_iter;
            out_ok[c] = it ? 1 : 0;
        }
        free(R);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)                          */
/* ------------------------------------------------------------------------- */
EXPORT void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------------- */
/* Box-Muller with polynomial log / sin / cos and Newton rsqrt, plain fp32   */
/* add/sub/mul only, so any IEEE binary32 machine gives the same bits.       */
/* ------------------------------------------------------------------------- */
/* log1p(f)/f on [sqrt(1/2)-1, sqrt(2)-1], degree 9, Chebyshev least squares */
static const float LOG_Q[10] = {
    0x1.000000p+0f, -0x1.fffff8p-2f, 0x1.555550p-2f, -0x1.000428p-2f, 0x1.99a476p-3f,
    -0x1.5427a8p-3f, 0x1.226874p-3f, -0x1.0f2c26p-3f, 0x1.087b72p-3f, -0x1.38e072p-4f};

EXPORT float orc_log_unit(float u) {          /* natural log of u in (0, 1] */
    uint32_t b = f2u(u);
    int e = (int)(b >> 23) - 127;
    float mant = u2f((b & 0x007FFFFFu) | 0x3F800000u);
    if (mant > 1.41421356f) { mant = mant * 0.5f; e += 1; }
    float f = mant - 1.0f;
    float q = LOG_Q[9];
    for (int k = 8; k >= 0; --k) { q = q * f; q = q + LOG_Q[k]; }
    float p = f * q;
    float el = (float)e * 0x1.62e430p-1f;     /* ln 2 rounded to fp32 */
    return el + p;
}

EXPORT float orc_sqrt_newton(float t) {       /* sqrt(t), t >= 0 */
    uint32_t i = 0x5F3759DFu - (f2u(t) >> 1);
    float y = u2f(i);
    for (int k = 0; k < 3; ++k) {
        float h = 0.5f * t;
        float yy = y * y;
        float g = h * yy;
        float c = 1.5f - g;
        y = y * c;
    }
    return t * y;
}

/* sin/cos of psi in [0, pi/2): Taylor polynomials in psi^2 */
static const float SIN_C[6] = {-0x1.555556p-3f, 0x1.111112p-7f, -0x1.a01a02p-13f,
                               0x1.71de3ap-19f, -0x1.ae6456p-26f, 0x1.612462p-33f};
static const float COS_C[7] = {-0x1.000000p-1f, 0x1.555556p-5f, -0x1.6c16c2p-10f, 0x1.a01a02p-16f,
                               -0x1.27e4fcp-22f, 0x1.1eed8ep-29f, -0x1.93974ap-37f};

EXPORT void orc_sincos_turn(uint32_t x, float *s_out, float *c_out) {
    uint32_t quad = x >> 30;
    float f = (float)(x & 0x3FFFFFFFu) * 0x1p-30f;
    float psi = f * 0x1.921fb6p+0f;           /* pi/2 rounded to fp32 */
    float p2 = psi * psi;
    float sp = SIN_C[5];
    for (int k = 4; k >= 0; --k) { sp = sp * p2; sp = sp + SIN_C[k]; }
    float s = psi * p2; s = s * sp; s = psi + s;
    float cp = COS_C[6];
    for (int k = 5; k >= 0; --k) { cp = cp * p2; cp = cp + COS_C[k]; }
    float c = p2 * cp; c = 1.0f + c;
    float so, co;
    switch (quad) {
        case 0: co = c; so = s; break;
        case 1: co = -s; so = c; break;
        case 2: co = -c; so = -s; break;
        default: co = s; so = -c; break;
    }
    *s_out = so; *c_out = co;
}

/* Two standard normals from two 32-bit words. */
EXPORT void orc_box_muller(uint32_t x0, uint32_t x1, float *z0, float *z1) {
    float u = (float)x0 * 0x1p-32f;
    u = u + 0x1p-33f;                          /* u in (0, 1] */
    float l = orc_log_unit(u);
    float t = -2.0f * l;
    float r = orc_sqrt_newton(t);
    float s, c;
    orc_sincos_turn(x1, &s, &c);
    *z0 = r * c;
    *z1 = r * s;
}

/* Channel LLRs of codewords first_cw .. first_cw+n_cw-1 (all-zero codeword,
 * BPSK +1): symbol i of codeword w uses Philox(ctr = (w_lo, w_hi, i/4, 0),
 * key = (seed_lo, seed_hi)) word pair (i%4)&2 and normal i%4; the LLR is
 * RNE_bf16(mu + sigma * z).  Variables n_tx..n-1 are punctured and set to +0
 * (ldpc_decoder_core.h:295-297; stepper:748-749).                           */
static void gen_one_clamped(const uint32_t key[2], uint64_t w, float mu, float sigma, float l_max, int n_tx, int n,
                            uint16_t *L);
static void gen_one(const uint32_t key[2], uint64_t w, float mu, float sigma, int n_tx, int n, uint16_t *L) {
    gen_one_clamped(key, w, mu, sigma, 0.0f, n_tx, n, L);
}

/* Channel LLRs with the optional magnitude clamp l_max applied before
 * rounding (ldpc_decoder_core.h:266-275). */
static void gen_one_clamped(const uint32_t key[2], uint64_t w, float mu, float sigma, float l_max, int n_tx, int n,
                            uint16_t *L) {
    for (int blk = 0; blk < (n_tx + 3) / 4; ++blk) {
        uint32_t ctr[4] = {(uint32_t)w, (uint32_t)(w >> 32), (uint32_t)blk, 0u}, x[4];
        orc_philox4x32_10(ctr, key, x);
        float z[4];
        orc_box_muller(x[0], x[1], &z[0], &z[1]);
        orc_box_muller(x[2], x[3], &z[2], &z[3]);
        for (int j = 0; j < 4 && 4 * blk + j < n_tx; ++j) {
            float t = sigma * z[j];
            float l = mu + t;
            if (l_max > 0.0f) {
                if (l > l_max) l = l_max;
                else if (l < -l_max) l = -l_max;
            }
            L[4 * blk + j] = f2bf(l);
        }
    }
    for (int v = n_tx; v < n; ++v) L[v] = 0;
}

/* The standard normals behind gen_one (z before mu + sigma z), n_tx per codeword. */
EXPORT void orc_normals(uint64_t seed, uint64_t first_cw, int n_cw, int n_tx, float *out) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int c = 0; c < n_cw; ++c) {
        uint64_t w = first_cw + (uint64_t)c;
        for (int blk = 0; blk < (n_tx + 3) / 4; ++blk) {
            uint32_t ctr[4] = {(uint32_t)w, (uint32_t)(w >> 32), (uint32_t)blk, 0u}, x[4];
            orc_philox4x32_10(ctr, key, x);
            float z[4];
            orc_box_muller(x[0], x[1], &z[0], &z[1]);
            orc_box_muller(x[2], x[3], &z[2], &z[3]);
            for (int j = 0; j < 4 && 4 * blk + j < n_tx; ++j) out[(size_t)c * n_tx + 4 * blk + j] = z[j];
        }
    }
}

EXPORT void orc_gen_channel(uint64_t seed, uint64_t first_cw, int n_cw, float mu, float sigma,
                            int n_tx, int n, uint16_t *out) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int c = 0; c < n_cw; ++c) gen_one(key, first_cw + (uint64_t)c, mu, sigma, n_tx, n, out + (size_t)c * n);
}

/* ------------------------------------------------------------------------- */
/* Random-info-bit codewords.  No reference counterpart: every reference     */
/* engine sends the all-zero codeword (ldpc_decoder_core.h:258-297;          */
/* stepper:744-757).  Info bits of codeword w: Philox4x32-10 with counter    */
/* (w_lo, w_hi, j, 1), j = 0 .. k_info/128 - 1, 128 bits per call, bit b of */
/* output word d = info bit 128 j + 32 d + b.  Parity bits: a dense          */
/* generator P of k_info rows, each row the n - k_info parity bits packed 64 */
/* per word (pw words): parity = XOR of the rows of the set info bits.  The  */
/* codeword is (u, parity); BPSK sends bit c as +mu (c = 0) or -mu (c = 1):  */
/* LLR = RNE(+-mu + sigma z) on the all-zero mode's noise z; punctured = +0. */
/* ------------------------------------------------------------------------- */
static void info_bits(const uint32_t key[2], uint64_t w, int k_info, uint8_t *c) {
    for (int j = 0; j < k_info / 128; ++j) {
        uint32_t ctr[4] = {(uint32_t)w, (uint32_t)(w >> 32), (uint32_t)j, 1u}, x[4];
        orc_philox4x32_10(ctr, key, x);
        for (int b = 0; b < 128; ++b) c[128 * j + b] = (uint8_t)((x[b / 32] >> (b % 32)) & 1u);
    }
}

static void encode_dense(const uint64_t *P, int pw, int k_info, int n, uint8_t *c) {
    uint64_t acc[1024];
    for (int k = 0; k < pw; ++k) acc[k] = 0;
    for (int k = 0; k < k_info; ++k)
        if (c[k])
            for (int q = 0; q < pw; ++q) acc[q] ^= P[(size_t)k * pw + q];
    for (int v = k_info; v < n; ++v) c[v] = (uint8_t)((acc[(v - k_info) / 64] >> ((v - k_info) % 64)) & 1u);
}

static void gen_one_cw(const uint32_t key[2], uint64_t w, float mu, float sigma, int n_tx, int n,
                       const uint8_t *c, uint16_t *L) {
    for (int blk = 0; blk < (n_tx + 3) / 4; ++blk) {
        uint32_t ctr[4] = {(uint32_t)w, (uint32_t)(w >> 32), (uint32_t)blk, 0u}, x[4];
        orc_philox4x32_10(ctr, key, x);
        float z[4];
        orc_box_muller(x[0], x[1], &z[0], &z[1]);
        orc_box_muller(x[2], x[3], &z[2], &z[3]);
        for (int j = 0; j < 4 && 4 * blk + j < n_tx; ++j) {
            float t = sigma * z[j];
            float l = (c[4 * blk + j] ? -mu : mu) + t;
            L[4 * blk + j] = f2bf(l);
        }
    }
    for (int v = n_tx; v < n; ++v) L[v] = 0;
}

/* Tally of a random codeword c: the golden's rule (stepper:799-815) with the
 * errors counted against the sent bits.  A codeword is a frame error if its
 * syndrome never cleared, or if it cleared on a word whose info-bit hard
 * decisions (bf16 < 0 -> 1, stepper:781) differ from the sent bits.  A frame
 * error adds the info bits k < k_info that are not strictly on the sent side
 * (bf16 == +-0, or the sign differs from c_k): for c = 0 that is the
 * golden's bf16 <= 0.0f.  hist / iterations as tally_t (hist[0] counts the
 * codewords that never converged).                                          */
static void tally_cw(const uint16_t *L, const uint8_t *c, int it, int k_info, int max_iter, uint64_t *fe, uint64_t *be,
                     uint64_t *its, uint64_t *h) {
    int mism = 0;
    for (int k = 0; k < k_info; ++k) mism |= hard_neg(L[k]) != c[k];
    if (it) {
        *its += (uint64_t)it;
        h[it] += 1;
    } else {
        *its += (uint64_t)max_iter;
        h[0] += 1;
    }
    if (!it || mism) {
        *fe += 1;
        for (int k = 0; k < k_info; ++k) *be += (uint64_t)(((L[k] & 0x7FFFu) == 0) || ((L[k] >> 15) != c[k]));
    }
}

/* Channel LLRs (and, if bits != NULL, the codeword bits) of random codewords
 * first_cw .. first_cw + n_cw - 1. */
EXPORT int orc_gen_channel_cw(uint64_t seed, uint64_t first_cw, int n_cw, float mu, float sigma, int n_tx, int n,
                              int k_info, const uint64_t *P, int pw, uint16_t *out, uint8_t *bits) {
    if (k_info % 128 || pw > 1024 || 64 * pw < n - k_info) return -1;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        uint8_t *c = (uint8_t *)malloc((size_t)n);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int i = 0; i < n_cw; ++i) {
            info_bits(key, first_cw + (uint64_t)i, k_info, c);
            encode_dense(P, pw, k_info, n, c);
            gen_one_cw(key, first_cw + (uint64_t)i, mu, sigma, n_tx, n, c, out + (size_t)i * n);
            if (bits) memcpy(bits + (size_t)i * n, c, (size_t)n);
        }
        free(c);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Monte Carlo point, Profile T (stepper:708-856)                            */
/* counters: [0] blocks, [1] frame errors, [2] bit errors, [3] iteration sum */
/* hist[i], i = 1..max_iter: codewords that converged at iteration i;        */
/* hist[0]: codewords that never converged (counted as max_iter in [3]).      */
/* Bit errors: info bits k < k_info with bf16 <= 0 of unconverged codewords  */
/* (stepper:801-809).                                                         */
/* ------------------------------------------------------------------------- */
/* Tally of one decoded codeword (stepper:799-815): it = iteration of
 * convergence (1..max_iter) or 0 if the syndrome never cleared.  A codeword
 * still active after max_iter is a frame error, counts max_iter iterations
 * and adds its info bits k < k_info with bf16 <= 0.0f to the bit errors;
 * converged codewords add nothing but their iteration count.               */
static void tally_t(const uint16_t *L, int it, int k_info, int max_iter, uint64_t *fe, uint64_t *be, uint64_t *its,
                    uint64_t *h) {
    if (it) {
        *its += (uint64_t)it;
        h[it] += 1;
    } else {
        *its += (uint64_t)max_iter;
        h[0] += 1;
        *fe += 1;
        for (int k = 0; k < k_info; ++k) *be += (uint64_t)nonpos(L[k]);
    }
}

/* Decode caller-supplied channel LLRs (n_cw x n bf16, codeword-major) and
 * tally them exactly as orc_run_point_t tallies its own channel: the golden
 * Monte Carlo accounting on any channel, e.g. the reference's own draws. */
EXPORT int orc_run_llrs_t(int n, int m, const int32_t *rp, const int32_t *col, const int32_t *sched, int k_info,
                          int n_cw, const uint16_t *llr, int max_iter, float beta, int threads, uint64_t *counters,
                          uint64_t *hist) {
    if (check_degrees_t(m, rp)) return -1;
    uint64_t fe = 0, be = 0, its = 0;
    memset(hist, 0, sizeof(uint64_t) * (size_t)(max_iter + 1));
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads) reduction(+ : fe, be, its)
#endif
    {
        uint16_t *L = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)n);
        uint16_t *R = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)rp[m]);
        uint64_t *h = (uint64_t *)calloc((size_t)max_iter + 1, sizeof(uint64_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int c = 0; c < n_cw; ++c) {
            memcpy(L, llr + (size_t)c * n, sizeof(uint16_t) * n);
            tally_t(L, decode_one_t(L, R, m, rp, col, sched, max_iter, beta), k_info, max_iter, &fe, &be, &its, h);
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        for (int i = 0; i <= max_iter; ++i) hist[i] += h[i];
        free(L); free(R); free(h);
    }
    counters[0] = (uint64_t)n_cw; counters[1] = fe; counters[2] = be; counters[3] = its;
    return 0;
}

EXPORT int orc_run_point_t(int n, int m, const int32_t *rp, const int32_t *col, const int32_t *sched,
                           int n_tx, int k_info, uint64_t seed, uint64_t first_cw, uint64_t n_cw,
                           float mu, float sigma, int max_iter, float beta, int threads,
                           uint64_t *counters, uint64_t *hist) {
    if (check_degrees_t(m, rp)) return -1;
    uint64_t blocks = 0, fe = 0, be = 0, its = 0;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    memset(hist, 0, sizeof(uint64_t) * (size_t)(max_iter + 1));
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads) reduction(+ : blocks, fe, be, its)
#endif
    {
        uint16_t *L = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)n);
        uint16_t *R = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)rp[m]);
        uint64_t *h = (uint64_t *)calloc((size_t)max_iter + 1, sizeof(uint64_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t c = 0; c < (int64_t)n_cw; ++c) {
            gen_one(key, first_cw + (uint64_t)c, mu, sigma, n_tx, n, L);
            blocks += 1;
            tally_t(L, decode_one_t(L, R, m, rp, col, sched, max_iter, beta), k_info, max_iter, &fe, &be, &its, h);
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        for (int i = 0; i <= max_iter; ++i) hist[i] += h[i];
        free(L); free(R); free(h);
    }
    counters[0] = blocks; counters[1] = fe; counters[2] = be; counters[3] = its;
    return 0;
}

/* Monte Carlo point, Profile T, random codewords (tally_cw's rule). */
EXPORT int orc_run_point_t_cw(int n, int m, const int32_t *rp, const int32_t *col, const int32_t *sched, int n_tx,
                              int k_info, const uint64_t *P, int pw, uint64_t seed, uint64_t first_cw, uint64_t n_cw,
                              float mu, float sigma, int max_iter, float beta, int threads, uint64_t *counters,
                              uint64_t *hist) {
    if (check_degrees_t(m, rp) || k_info % 128 || pw > 1024 || 64 * pw < n - k_info) return -1;
    uint64_t blocks = 0, fe = 0, be = 0, its = 0;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    memset(hist, 0, sizeof(uint64_t) * (size_t)(max_iter + 1));
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads) reduction(+ : blocks, fe, be, its)
#endif
    {
        uint16_t *L = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)n);
        uint16_t *R = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)rp[m]);
        uint8_t *c = (uint8_t *)malloc((size_t)n);
        uint64_t *h = (uint64_t *)calloc((size_t)max_iter + 1, sizeof(uint64_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t i = 0; i < (int64_t)n_cw; ++i) {
            const uint64_t w = first_cw + (uint64_t)i;
            info_bits(key, w, k_info, c);
            encode_dense(P, pw, k_info, n, c);
            gen_one_cw(key, w, mu, sigma, n_tx, n, c, L);
            blocks += 1;
            tally_cw(L, c, decode_one_t(L, R, m, rp, col, sched, max_iter, beta), k_info, max_iter, &fe, &be, &its,
                     h);
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        for (int i = 0; i <= max_iter; ++i) hist[i] += h[i];
        free(L); free(R); free(c); free(h);
    }
    counters[0] = blocks; counters[1] = fe; counters[2] = be; counters[3] = its;
    return 0;
}

/* Monte Carlo point, Profile R (ldpc_decoder_core.h:258-577): frame error
 * <=> at least one sign bit set among the n_tx transmitted variables; bit
 * errors count those sign bits.  hist / iteration sum as in Profile T, with
 * "converged" meaning the syndrome cleared (checked from iteration 3). */
EXPORT int orc_run_point_r(int n, int m, const int32_t *rp, const int32_t *col, int n_tx, uint64_t seed,
                           uint64_t first_cw, uint64_t n_cw, float mu, float sigma, float l_max, int max_iter, int nms,
                           float r_max, float l_post_max, int threads, uint64_t *counters, uint64_t *hist) {
    if (check_degrees_r(m, rp)) return -1;
    orc_cfg_r cfg = {nms, r_max, l_post_max};
    uint64_t blocks = 0, fe = 0, be = 0, its = 0;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    memset(hist, 0, sizeof(uint64_t) * (size_t)(max_iter + 1));
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads) reduction(+ : blocks, fe, be, its)
#endif
    {
        uint16_t *L = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)n);
        uint16_t *R = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)rp[m]);
        uint64_t *h = (uint64_t *)calloc((size_t)max_iter + 1, sizeof(uint64_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t c = 0; c < (int64_t)n_cw; ++c) {
            gen_one_clamped(key, first_cw + (uint64_t)c, mu, sigma, l_max, n_tx, n, L);
            int it = decode_one_r(L, R, m, rp, col, max_iter, &cfg);
            blocks += 1;
            its += it ? (uint64_t)it : (uint64_t)max_iter;
            h[it] += 1;
            uint64_t b = 0;
            for (int k = 0; k < n_tx; ++k) b += L[k] >> 15;
            be += b;
            fe += b > 0;
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        for (int i = 0; i <= max_iter; ++i) hist[i] += h[i];
        free(L); free(R); free(h);
    }
    counters[0] = blocks; counters[1] = fe; counters[2] = be; counters[3] = its;
    return 0;
}

EXPORT void orc_gen_channel_clamped(uint64_t seed, uint64_t first_cw, int n_cw, float mu, float sigma, float l_max,
                                    int n_tx, int n, uint16_t *out) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int c = 0; c < n_cw; ++c)
        gen_one_clamped(key, first_cw + (uint64_t)c, mu, sigma, l_max, n_tx, n, out + (size_t)c * n);
}

EXPORT int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
