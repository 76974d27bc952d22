/*
 * ldpc_mi355x.h -- C ABI of libldpc, the MI355X (gfx950) Monte Carlo LDPC
 * engine for the CCSDS AR4JA codes.
 *
 * The reference drives its decoder as a separate process: run_simulation.py
 * builds an argv for the engine binary and reads JSON lines from its stdout
 * (run_simulation.py:165-225; engines/scalar_5w/ldpc_sim_scalar.cpp:271-641).
 * This library is the in-process replacement of that engine: a Python or C
 * host calls these functions through ctypes/dlopen, with plain pointers and
 * fixed-width POD structs.  Each entry point names the reference interface it
 * replaces.
 *
 * Conventions:
 *   - Every function returns LDPC_OK (0) or a negative ldpc_status; none
 *     throws across the ABI.  ldpc_last_error() gives the message.
 *   - Host buffers are caller-owned; device memory, streams and events are
 *     owned by the context.  Calls are synchronous unless named *_async.
 *   - A context is not thread-safe: use one context per host thread.
 *   - LLRs are bfloat16 bit patterns (uint16_t), codeword-major, N per
 *     codeword; punctured variables are the last ones of each codeword.
 */
#ifndef LDPC_MI355X_H
#define LDPC_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDPC_ABI_VERSION 2

typedef enum ldpc_status {
    LDPC_OK = 0,
    LDPC_ERR_ARG = -1,         /* invalid argument */
    LDPC_ERR_UNSUPPORTED = -2, /* code / profile combination not built */
    LDPC_ERR_HIP = -3,         /* HIP runtime error */
    LDPC_ERR_NOMEM = -4,       /* allocation failed */
    LDPC_ERR_NODEV = -5        /* no usable GPU */
} ldpc_status;

/* Decoder arithmetic profiles. */
enum {
    /* Tree Approximate-Min* with bf16 rounding of q, r and L, truncated
     * constants and damping beta: the reference's golden stepper
     * (tests/test_circulant_dma_stepper.cpp:107-327, :743-815).  The nine
     * AR4JA codes; degree-10 and degree-18 checks (rates 2/3, 4/5) use the
     * degree-6 tree generalised by pairs (j, j + d/2) and a prefix/suffix
     * pass over the pairs, which the reference does not define. */
    LDPC_PROFILE_T = 0,
    /* Recursive Approximate-Min* of the scalar_5w engine that
     * run_simulation.py drives (engines/scalar_5w/ldpc_decoder_core.h:
     * 155-201, 413-520): unrounded q, rows in the given order, syndrome on
     * the sign bit from the third iteration, frame error <=> a transmitted
     * sign bit set.  Check degrees 1..12. */
    LDPC_PROFILE_R = 1,
    /* Same engine built with USE_NORMALIZED_MIN_SUM: min-sum, alpha 0.75
     * (ldpc_decoder_core.h:320-411; the engine's -n flag). */
    LDPC_PROFILE_R_NMS = 2,
    /* Learned layered min-sum (no reference counterpart): the forward of the
     * differentiable decoder of ldpc_mi355x.dmin with a syndrome early exit.
     * fp32 LLRs and messages, r = copysign(min(max(alpha_p g - beta_p, 0),
     * r_max), sign) with g the leave-one-out minimum and alpha / beta per
     * (layer, circulant) from ldpc_set_dmin_params; channel as Profile T.
     * All nine AR4JA codes; for K=16384 r1/2 ldpc_create also allocates a
     * message workspace of 176.25 KiB per workgroup of the launch grid (one
     * workgroup per CU). */
    LDPC_PROFILE_D = 3
};

/* Code families. */
enum {
    LDPC_CODE_AR4JA = 1, /* CCSDS AR4JA (matrices/makeH_AR4JA.m) */
    LDPC_CODE_MATRIX = 2 /* explicit parity-check matrix (a parsed .chinn.out) */
};

typedef struct ldpc_code_desc {
    uint32_t family;   /* LDPC_CODE_AR4JA or LDPC_CODE_MATRIX */
    uint32_t k_info;   /* AR4JA: 1024, 4096 or 16384 */
    uint32_t rate_num; /* AR4JA: 1, 2 or 4 */
    uint32_t rate_den; /* AR4JA: 2, 3 or 5 */
    /* LDPC_CODE_MATRIX: m rows in decoding order, each row's variables in
     * order (row_ptr[m+1], col[row_ptr[m]], 0-based; a row may not list a
     * variable twice); the last n_punct variables are punctured.  Copied at
     * ldpc_create. */
    uint32_t n, m, n_punct, reserved;
    const int32_t *row_ptr;
    const int32_t *col;
} ldpc_code_desc;

typedef struct ldpc_decoder_cfg {
    uint32_t profile;  /* LDPC_PROFILE_T, _R or _R_NMS */
    uint32_t max_iter; /* >= 1 (reference default 200; run_simulation.py 16) */
    float beta;        /* Profile T damping of check messages; 0.9375 matches
                          the silicon and the stepper's --load-init default
                          (compute_trisc.cpp:395,566; stepper:1017) */
    /* Profile R clamps, 0 = off (ldpc_sim_scalar.cpp:290-309: -L -R -V -G) */
    float l_max;       /* |channel LLR| <= l_max before rounding */
    float r_max;       /* |check message| <= r_max */
    float l_post_max;  /* |posterior LLR| <= l_post_max */
    float gamma;       /* if > 0: l_max = gamma * mu at each Eb/N0 */
    uint32_t variant;  /* Profile T kernel variant: 0 = library default, or
                          LDPC_VARIANT(checks per thread, layer order) */
} ldpc_decoder_cfg;

/* Profile T kernel variants.  checks_per_thread (1 or 2) sets how many checks
 * of a layer one GPU thread updates (threads per codeword = z / checks);
 * layer_order 0 decodes block rows 1, 2, 0 (the golden's tile order,
 * stepper:765-771), 1 decodes block rows 0, 1, 2.  The checks-per-thread
 * choice never changes results; the layer order does.  0 selects the
 * specialised kernel for AR4JA r1/2 K=1024 and a tuned variant with layer
 * order 0 for the other codes. */
#define LDPC_VARIANT(checks_per_thread, layer_order) \
    (0x100u | ((uint32_t)(checks_per_thread) & 0xFu) | (((uint32_t)(layer_order) & 0xFu) << 4))

typedef struct ldpc_code_info {
    uint32_t n;       /* variables (transmitted + punctured) */
    uint32_t m;       /* checks */
    uint32_t k_info;  /* information bits (the first k_info variables) */
    uint32_t n_tx;    /* transmitted symbols per codeword */
    uint32_t z;       /* layer (circulant) size */
    uint32_t n_layers;
    uint32_t n_edges;
    uint32_t max_iter;
} ldpc_code_info;

/* One Monte Carlo operating point: all-zero codeword, BPSK, AWGN. */
typedef struct ldpc_point_cfg {
    double ebn0_db;
    uint64_t seed;        /* Philox key */
    uint64_t first_cw;    /* codeword id of the first codeword simulated */
    uint64_t max_blocks;  /* stop after this many codewords */
    uint64_t target_fe;   /* stop once this many frame errors (0: none) */
    uint64_t chunk_cw;    /* codewords per launch (0: 1 << 20; at most 1 << 31) */
} ldpc_point_cfg;

typedef struct ldpc_point_result {
    uint64_t blocks;       /* codewords simulated */
    uint64_t frame_errors; /* codewords whose syndrome never cleared */
    uint64_t bit_errors;   /* info bits <= 0 in those codewords */
    uint64_t iter_sum;     /* sum of iterations (max_iter for failures) */
    uint64_t next_cw;      /* first codeword id not yet simulated */
    uint64_t elapsed_ns;   /* host wall time of the call */
    uint64_t kernel_ns;    /* summed decode-kernel time (HIP events) */
    uint64_t launches;     /* decode-kernel launches */
} ldpc_point_result;

typedef struct ldpc_ctx ldpc_ctx;

/* Create a context on a device for one code and decoder configuration.
 * Replaces the engine binary's start-up: matrix parse and device open
 * (ldpc_sim_scalar.cpp:271-330; ldpc_sim_circulant_dma.cpp:98-126). */
int ldpc_create(const ldpc_code_desc *code, const ldpc_decoder_cfg *cfg, int device, ldpc_ctx **out);

void ldpc_destroy(ldpc_ctx *ctx);

/* Last error message of ctx (or, if ctx is NULL, of the last failed call
 * that takes no context: ldpc_create, ldpc_decode_llrs_device_noalloc).  The
 * string lives until the next call on the same context. */
const char *ldpc_last_error(const ldpc_ctx *ctx);

int ldpc_get_code_info(const ldpc_ctx *ctx, ldpc_code_info *info);

/* Layer schedule compiled into the Profile T kernels for an AR4JA code and
 * layer order (0 or 1, see LDPC_VARIANT; needs no GPU): per layer
 * [first_row, deg, (base, shift) * deg], layers in decoding order; variable
 * of check j of a layer = base + (shift + j) mod z.  Returns the number of
 * int32 written or a negative status. */
int ldpc_get_schedule(const ldpc_code_desc *code, uint32_t layer_order, int32_t *out, int32_t cap);

/* Profile T kernel variants built for an AR4JA code (needs no GPU): writes up
 * to cap LDPC_VARIANT values and returns how many exist. */
int ldpc_list_variants(const ldpc_code_desc *code, uint32_t *out, int32_t cap);

/* Level of every row for the Profile R schedule (needs no GPU): level(i) =
 * 1 + max level of the earlier rows sharing a variable with row i.  Rows of
 * one level are decoded in parallel; the result equals the sequential row
 * order.  Writes m levels (1-based) and returns the number of levels. */
int ldpc_row_levels(uint32_t n, uint32_t m, const int32_t *row_ptr, const int32_t *col, int32_t *level);

/* Pass of every row in the Profile R kernel's schedule (needs no GPU; no
 * reference counterpart: it describes how the sequential row order is run
 * in parallel).  A row joins a pass of at most `width` rows once every
 * earlier row sharing a variable with it sits in an earlier pass; among
 * ready rows the longest remaining dependency chain goes first.  The kernel
 * uses width 64.  Writes m pass indices (0-based) and returns the number of
 * passes. */
int ldpc_row_passes(uint32_t n, uint32_t m, const int32_t *row_ptr, const int32_t *col, int32_t width,
                    int32_t *pass);

/* Pass of every row in the two-codeword Profile R kernel's schedule (needs no
 * GPU; no reference counterpart, like ldpc_row_passes).  Same readiness rule,
 * passes of at most `width` rows (the kernel uses 32).  When every row has
 * degree 3 or 6 each pass holds rows of one degree (kind 3 or 6); otherwise
 * every pass has kind 7 (rows of any degree, masked datapath).  Writes m pass
 * indices and, when kind is not NULL and kind_cap covers them, one kind per
 * pass; returns the number of passes, or LDPC_ERR_ARG. */
int ldpc_row_passes_by_degree(uint32_t n, uint32_t m, const int32_t *row_ptr, const int32_t *col, int32_t width,
                              int32_t *pass, uint8_t *kind, int32_t kind_cap);

/* Monte Carlo point: generate, decode and tally codewords first_cw.. until
 * max_blocks or target_fe is reached, in launches of chunk_cw codewords.
 * hist (may be NULL) has max_iter + 1 entries: hist[i] = codewords that
 * converged at iteration i, hist[0] = codewords that did not converge.
 * Results ADD to *res and hist, so a sweep can be resumed chunk by chunk.
 * Replaces the engine's batch loop and stopping rule
 * (ldpc_sim_scalar.cpp:421-641; ldpc_sim_circulant_dma.cpp:295-411) and the
 * golden Monte Carlo (stepper:708-856). */
int ldpc_run_point(ldpc_ctx *ctx, const ldpc_point_cfg *pt, ldpc_point_result *res, uint64_t *hist);

/* Decode caller-supplied channel LLRs (host buffers, n_cw * N bf16 each).
 * out_iters[c] = iteration of convergence, or max_iter if it never converged;
 * out_ok[c] = 1 if the syndrome cleared.  Replaces the stepper's dump replay
 * (stepper:861-990). */
int ldpc_decode_llrs(ldpc_ctx *ctx, const uint16_t *llr, uint32_t n_cw, uint16_t *out_llr, uint32_t *out_iters,
                     uint8_t *out_ok);

/* Same on device buffers, enqueued on the context's stream; returns after
 * the launch (use ldpc_sync).  Buffers must stay valid until then. */
int ldpc_decode_llrs_device_async(ldpc_ctx *ctx, const void *d_llr, uint32_t n_cw, void *d_out_llr,
                                  void *d_out_iters, void *d_out_ok);

/* Channel LLRs of codewords first_cw .. first_cw + n_cw - 1 at ebn0_db
 * (all-zero codeword, BPSK, AWGN from Philox4x32-10 keyed by seed), written
 * to a host buffer of n_cw * N bf16.  Replaces the device channel generator
 * (compute_trisc.cpp:179-294; ldpc_decoder_core.h:128-152). */
int ldpc_gen_channel(ldpc_ctx *ctx, uint64_t seed, uint64_t first_cw, uint32_t n_cw, double ebn0_db,
                     uint16_t *out_llr);

/* BPSK LLR mean and standard deviation at ebn0_db for this code's rate
 * R = k_info / n_tx:  mu = 4 R g, sigma = sqrt(8 R g), g = 10^(ebn0/10)
 * (ldpc_sim_scalar.cpp:352-355). */
int ldpc_channel_params(const ldpc_ctx *ctx, double ebn0_db, float *mu, float *sigma);

/* Systematic encoder of an AR4JA code (needs no GPU; no reference
 * counterpart: every reference engine sends the all-zero codeword,
 * engines/scalar_5w/ldpc_decoder_core.h:258-297).  A codeword is (u, p):
 * the k_info information bits, then 12 parity blocks of z bits, the last 4
 * of them the punctured variables.  With a block's bits read as polynomial
 * coefficients (bit t <-> x^t), p_r = sum_i g_{r,i} u_i mod x^z + 1 over
 * the info blocks u_i.  ldpc_get_generator writes the g_{r,i} as
 * [12][k_info / z][z / 32] uint32 words (bit s of word s / 32 = coefficient
 * of x^s; at most cap words) and returns the total count; ldpc_encode
 * writes the n bits (one byte each, 0 or 1) of each of n_cw codewords from
 * their k_info info bits (one byte each). */
int ldpc_get_generator(const ldpc_code_desc *code, uint32_t *out, int32_t cap);
int ldpc_encode(const ldpc_code_desc *code, const uint8_t *info, uint32_t n_cw, uint8_t *codeword);

/* One variant's measurement by ldpc_autotune. */
typedef struct ldpc_tune_result {
    uint32_t variant;
    uint32_t threads_per_wg;   /* workgroup size */
    uint32_t cw_per_wg;        /* codewords in flight per workgroup */
    uint32_t wg_per_cu;        /* resident workgroups per CU */
    double gsps;               /* transmitted symbols per second of kernel time */
    double mean_iterations;
    uint64_t frame_errors;
    uint64_t kernel_ns;
} ldpc_tune_result;

/* Autotuner: run the same n_cw codewords (seed, codeword ids 0..n_cw-1) at
 * ebn0_db through every Profile T variant of an AR4JA code, best of `reps`
 * timed launches each, and write one ldpc_tune_result per variant (at most
 * cap).  Variants with the same layer order must give identical counters
 * (checked; a mismatch is LDPC_ERR_HIP).  *best receives the fastest variant
 * among those with layer order 0 if any_order is 0, among all otherwise.
 * Returns the number of results.  No reference counterpart: the reference
 * has one hand-tuned kernel per engine. */
int ldpc_autotune(const ldpc_code_desc *code, const ldpc_decoder_cfg *cfg, int device, double ebn0_db, uint64_t seed,
                  uint32_t n_cw, uint32_t reps, int any_order, ldpc_tune_result *out, int32_t cap, uint32_t *best);

/* Mixed-code batch: several AR4JA codes (Profile T), each with its own kernel
 * variant, Eb/N0 and codeword stream, run as one captured HIP graph per
 * round: every code's persistent kernel takes codewords from its own device
 * queue; a round simulates up to pts[i].chunk_cw codewords of every code
 * that has not reached its max_blocks / target_fe.  Inside the graph the
 * codes decoded one codeword per workgroup run in sequence and the codes with
 * several codewords per workgroup run beside them, so kernels may overlap;
 * results do not depend on it.  Replaces running the
 * reference's engine once per code (run_simulation.py:164-256 per matrix). */
typedef struct ldpc_batch ldpc_batch;

/* variants may be NULL (library defaults); cfg->variant is ignored. */
int ldpc_batch_create(const ldpc_code_desc *codes, const uint32_t *variants, uint32_t n_codes,
                      const ldpc_decoder_cfg *cfg, int device, ldpc_batch **out);

/* pts, res: n_codes entries; hist (may be NULL): n_codes * (max_iter + 1),
 * laid out code by code as in ldpc_run_point.  Results ADD to res and hist.
 * res[i].kernel_ns is the graph time of the rounds code i took part in. */
int ldpc_batch_run(ldpc_batch *b, const ldpc_point_cfg *pts, ldpc_point_result *res, uint64_t *hist);

/* Code info of batch member i. */
int ldpc_batch_code_info(const ldpc_batch *b, uint32_t i, ldpc_code_info *info);

const char *ldpc_batch_last_error(const ldpc_batch *b);

void ldpc_batch_destroy(ldpc_batch *b);

/* Trapping-set capture.  After ldpc_set_capture(ctx, max_frames) every
 * later ldpc_run_point records the codeword id of each failing codeword
 * (up to max_frames, in completion order; 0 turns capture off and clears
 * it).  A captured codeword is reproduced exactly from (seed, id): its
 * channel LLRs with ldpc_gen_channel and its final state with
 * ldpc_decode_llrs.  Replaces the circulant engine's first-failure trap and
 * vnode dumps (engines/circulant_dma_4cw/writer_ncrisc.cpp:330-398;
 * ldpc_sim_circulant_dma.cpp:435-472). */
int ldpc_set_capture(ldpc_ctx *ctx, uint32_t max_frames);

/* Captured ids so far, ascending (at most cap written); *n_failed (may be
 * NULL) receives the number of failures seen since ldpc_set_capture, which
 * can exceed what was kept.  Returns the number written. */
int ldpc_get_captures(ldpc_ctx *ctx, uint64_t *cw_ids, uint32_t cap, uint64_t *n_failed);

/* Profile D parameters: alpha and beta are [groups][n_params] fp32 (row t
 * for iteration t, the last row beyond; groups = 1 shares one row), n_params
 * = the sum of the layer degrees (n_edges / z); correction 0 is plain
 * min-sum, 1 applies the Approx-Min* correction to the minimum the other
 * edges get (from the two smallest magnitudes), 2 keeps three minima and
 * gives every edge min* of the two smallest magnitudes among the others;
 * r_max > 0 caps message magnitudes.  Copied to the device; required before Profile D work.  No
 * reference counterpart (the trained decoder's parameters). */
int ldpc_set_dmin_params(ldpc_ctx *ctx, const float *alpha, const float *beta, uint32_t groups, uint32_t n_params,
                         int correction, float r_max);

/* Allocation-free decode.  No counterpart in the reference: this entry point
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It decodes
 * like ldpc_decode_llrs_device_async, Profile T only (any AR4JA code and
 * kernel variant whose kernels need no workspace; cfg as for ldpc_create),
 * but needs no context: the kernels' scratch memory (work-queue counter, tally
 * counters, iteration histogram) is one fixed 64 KiB device array per GPU
 * inside libldpc, held until the process ends, so max_iter is limited to
 * LDPC_NOALLOC_MAX_ITER.  The call allocates nothing and enqueues fill
 * kernels and the decode on `stream` (a hipStream_t; NULL = the null stream)
 * of `device`, then returns: the caller synchronizes.  Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (concurrent
 * streams or host threads) race and must be serialized by the caller.  The
 * first call on a device resolves the array's address (hipGetSymbolAddress);
 * make it outside graph capture.  Errors: ldpc_last_error(NULL). */
#define LDPC_NOALLOC_MAX_ITER 8000
int ldpc_decode_llrs_device_noalloc(const ldpc_code_desc *code, const ldpc_decoder_cfg *cfg, int device,
                                    const void *d_llr, uint32_t n_cw, void *d_out_llr, void *d_out_iters,
                                    void *d_out_ok, void *stream);

/* Statistical check of the channel generator (no reference counterpart):
 * counts[k] = number of the normals z behind the channel of codewords
 * first_cw .. first_cw + n_cw - 1 (n_tx per codeword, the Philox4x32-10 +
 * Box-Muller stream of ldpc_gen_channel) with |z| > thresholds[k],
 * 1 <= n_thr <= 16, n_cw <= 2^40.  The stream's u = x 2^-32 + 2^-33 bounds
 * |z| by sqrt(66 ln 2) = 6.764. */
int ldpc_noise_tails(ldpc_ctx *ctx, uint64_t seed, uint64_t first_cw, uint64_t n_cw, const float *thresholds,
                     uint32_t n_thr, uint64_t *counts);

/* Codeword mode of ldpc_run_point and ldpc_gen_channel.  No reference
 * counterpart: every reference engine sends the all-zero codeword
 * (ldpc_decoder_core.h:258-297; stepper:744-757), which is the default here
 * too.  LDPC_CODEWORD_RANDOM sends random information bits, systematically
 * encoded (ldpc_encode): the info bits of codeword w are Philox4x32-10 with
 * counter (w_lo, w_hi, j, 1) for j = 0 .. k_info/128 - 1 and key seed, 128
 * bits per call (bit b of output word d = info bit 128 j + 32 d + b); BPSK
 * sends bit 1 as -mu on the same noise as the all-zero mode.  The tally
 * then counts against the sent bits: a frame error is a codeword whose
 * syndrome never cleared or cleared with an info-bit hard decision
 * (bf16 < 0 -> 1) other than the sent bit; its bit errors are the info bits
 * not strictly on the sent side (bf16 +-0 or the wrong sign), which for an
 * all-zero word is the golden's bf16 <= 0.0f.  hist[0] still counts only the
 * codewords that never converged.  Built for the Profile T kernel of AR4JA
 * r1/2 K=1024 (variant 0); other contexts return LDPC_ERR_UNSUPPORTED. */
enum { LDPC_CODEWORD_ZERO = 0, LDPC_CODEWORD_RANDOM = 1 };
int ldpc_set_codeword_mode(ldpc_ctx *ctx, uint32_t mode);

/* Wait for all work on the context's stream. */
int ldpc_sync(ldpc_ctx *ctx);

/* The context's HIP stream (hipStream_t), for callers that time or order
 * work on it. */
void *ldpc_stream(ldpc_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* LDPC_MI355X_H */
