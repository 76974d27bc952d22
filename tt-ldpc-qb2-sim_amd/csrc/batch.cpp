// batch.cpp -- Profile T kernel-variant autotuner and the mixed-code batch
// (several AR4JA codes per captured HIP graph) of libldpc.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/ldpc_mi355x.h"
#include "kernels.h"
#include "runtime_internal.h"

namespace {

std::string g_batch_error;

int set_error(std::string *where, int code, const std::string &msg) {
    (where ? *where : g_batch_error) = msg;
    return code;
}

bool same_counters(const ldpc_point_result &a, const ldpc_point_result &b) {
    return a.blocks == b.blocks && a.frame_errors == b.frame_errors && a.bit_errors == b.bit_errors &&
           a.iter_sum == b.iter_sum;
}

}  // namespace

struct ldpc_batch {
    int device = 0;
    uint32_t n_codes = 0, nh = 0;
    ldpc_decoder_cfg cfg{};
    std::vector<const ldpc::MultiVariant *> mv;
    std::vector<ldpc_code_info> info;
    std::vector<int> grid;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // The codes' kernels run on side streams forked from `stream` inside the
    // graph: one stream for the codes decoded one codeword per workgroup
    // (group mode, whole-CU workgroups), one for the codes with several
    // codewords per workgroup (wave mode), so the short wave-mode kernels and
    // their tails (codewords still iterating) overlap the long ones.
    std::vector<hipStream_t> side;
    std::vector<int> lane;  // side stream of each code
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    uint32_t *d_queue = nullptr;               // [n_codes]
    unsigned long long *d_counters = nullptr;  // [4 n_codes]
    unsigned long long *d_hist = nullptr;      // [n_codes nh]
    ldpc::MultiDyn *d_dyn = nullptr;           // [n_codes]
    ldpc::MultiDyn *h_dyn = nullptr;           // pinned [n_codes]
    unsigned long long *h_counters = nullptr;  // pinned [4 n_codes]
    unsigned long long *h_hist = nullptr;      // pinned [n_codes nh]
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string err;
};

#define B_TRY(b, expr)                                                                                         \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if (e_ != hipSuccess)                                                                                  \
            return set_error((b) ? &(b)->err : nullptr, LDPC_ERR_HIP,                                          \
                             std::string(#expr ": ") + hipGetErrorString(e_));                                 \
    } while (0)

namespace {

// One round of the batch as a graph: parameters in, counters cleared, every
// code's persistent kernel, counters and histograms out.  Counters are
// cleared with fill kernels, not memset nodes.  The kernels are independent
// (own queue, counters, histogram and codeword ids), so they hang off a
// fork/join of side streams and may run concurrently.
int capture_graph(ldpc_batch *b) {
    const uint32_t n = b->n_codes;
    B_TRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    auto step = [&]() -> hipError_t {
        hipError_t e;
        if ((e = hipMemcpyAsync(b->d_dyn, b->h_dyn, n * sizeof(ldpc::MultiDyn), hipMemcpyHostToDevice, b->stream)))
            return e;
        if ((e = ldpc::launch_fill_u32(b->d_queue, n, 0u, b->stream))) return e;
        if ((e = ldpc::launch_fill_u64(b->d_counters, 4 * (size_t)n, 0ull, b->stream))) return e;
        if ((e = ldpc::launch_fill_u64(b->d_hist, (size_t)n * b->nh, 0ull, b->stream))) return e;
        if ((e = hipEventRecord(b->fork, b->stream))) return e;
        for (hipStream_t s : b->side)
            if ((e = hipStreamWaitEvent(s, b->fork, 0))) return e;
        for (uint32_t i = 0; i < n; ++i) {
            ldpc::MultiParams p{};
            p.queue = b->d_queue + i;
            p.counters = b->d_counters + 4 * i;
            p.hist = b->d_hist + (size_t)i * b->nh;
            p.dyn = b->d_dyn + i;
            p.max_iter = (int)b->cfg.max_iter;
            p.beta = b->cfg.beta;
            p.beta3 = b->cfg.beta < 1.0f ? b->cfg.beta : 1.0f;  // stepper:300-304
            if ((e = b->mv[i]->launch(p, b->grid[i], b->side[b->lane[i]]))) return e;
        }
        for (size_t k = 0; k < b->side.size(); ++k) {
            if ((e = hipEventRecord(b->join[k], b->side[k]))) return e;
            if ((e = hipStreamWaitEvent(b->stream, b->join[k], 0))) return e;
        }
        if ((e = hipMemcpyAsync(b->h_counters, b->d_counters, 4 * n * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, b->stream)))
            return e;
        return hipMemcpyAsync(b->h_hist, b->d_hist, (size_t)n * b->nh * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, b->stream);
    };
    const hipError_t e1 = step();
    const hipError_t e2 = hipStreamEndCapture(b->stream, &b->graph);
    B_TRY(b, e1);
    B_TRY(b, e2);
    B_TRY(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return LDPC_OK;
}

}  // namespace

extern "C" {

int ldpc_autotune(const ldpc_code_desc *code, const ldpc_decoder_cfg *cfg, int device, double ebn0_db, uint64_t seed,
                  uint32_t n_cw, uint32_t reps, int any_order, ldpc_tune_result *out, int32_t cap, uint32_t *best) {
    if (!code || !cfg || !best || (!out && cap > 0) || cap < 0 || n_cw == 0 || reps == 0 || !std::isfinite(ebn0_db))
        return set_error(nullptr, LDPC_ERR_ARG, "autotune: bad argument");
    if (cfg->profile != LDPC_PROFILE_T) return set_error(nullptr, LDPC_ERR_ARG, "autotune: Profile T only");
    std::vector<uint32_t> vars(16);
    const int nv = ldpc_list_variants(code, vars.data(), (int32_t)vars.size());
    if (nv < 0) return set_error(nullptr, nv, "autotune: not an AR4JA code");
    int k_id, rate_id;
    ldpc::ar4ja_ids(*code, k_id, rate_id);
    int cus = 0;
    B_TRY((ldpc_batch *)nullptr, hipSetDevice(device));
    B_TRY((ldpc_batch *)nullptr, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    ldpc_point_result ref[2]{};
    bool have_ref[2] = {false, false};
    double best_gsps = -1.0;
    *best = 0;
    for (int i = 0; i < nv && i < cap; ++i) {
        ldpc_decoder_cfg c2 = *cfg;
        c2.variant = vars[i];
        ldpc_ctx *ctx = nullptr;
        int rc = ldpc_create(code, &c2, device, &ctx);
        if (rc) return set_error(nullptr, rc, std::string("autotune: ") + ldpc_last_error(nullptr));
        ldpc_code_info info{};
        ldpc_get_code_info(ctx, &info);
        const ldpc_point_cfg pt{ebn0_db, seed, 0, n_cw, 0, n_cw};
        ldpc_point_result warm{};
        rc = ldpc_run_point(ctx, &pt, &warm, nullptr);  // warm-up; also the counters to compare
        uint64_t best_ns = UINT64_MAX;
        for (uint32_t r = 0; r < reps && !rc; ++r) {
            ldpc_point_result res{};
            rc = ldpc_run_point(ctx, &pt, &res, nullptr);
            if (!rc && !same_counters(res, warm)) rc = LDPC_ERR_HIP;
            best_ns = std::min<uint64_t>(best_ns, res.kernel_ns);
        }
        std::string msg = rc ? ldpc_last_error(ctx) : "";
        ldpc_destroy(ctx);
        if (rc) return set_error(nullptr, rc, "autotune: variant run failed or was not repeatable: " + msg);
        const ldpc::MultiVariant *mv = ldpc::find_variant(k_id, rate_id, vars[i]);
        const int order = mv->order;
        if (have_ref[order] && !same_counters(ref[order], warm))
            return set_error(nullptr, LDPC_ERR_HIP, "autotune: variants of one layer order disagree");
        ref[order] = warm;
        have_ref[order] = true;
        int occ = 0;
        B_TRY((ldpc_batch *)nullptr, mv->occupancy(&occ));
        ldpc_tune_result &t = out[i];
        t.variant = vars[i];
        t.threads_per_wg = (uint32_t)mv->threads;
        t.cw_per_wg = (uint32_t)mv->cw_per_wg;
        t.wg_per_cu = (uint32_t)occ;
        t.kernel_ns = best_ns;
        t.gsps = best_ns ? (double)n_cw * info.n_tx / (double)best_ns : 0.0;
        t.mean_iterations = (double)warm.iter_sum / (double)warm.blocks;
        t.frame_errors = warm.frame_errors;
        if ((any_order || order == 0) && t.gsps > best_gsps) {
            best_gsps = t.gsps;
            *best = vars[i];
        }
    }
    return nv < cap ? nv : cap;
}

const char *ldpc_batch_last_error(const ldpc_batch *b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

void ldpc_batch_destroy(ldpc_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void *p : {(void *)b->d_queue, (void *)b->d_counters, (void *)b->d_hist, (void *)b->d_dyn})
        if (p) (void)hipFree(p);
    for (void *p : {(void *)b->h_dyn, (void *)b->h_counters, (void *)b->h_hist})
        if (p) (void)hipHostFree(p);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t s : b->side)
        if (s) (void)hipStreamDestroy(s);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int ldpc_batch_create(const ldpc_code_desc *codes, const uint32_t *variants, uint32_t n_codes,
                      const ldpc_decoder_cfg *cfg, int device, ldpc_batch **out) {
    if (!codes || !cfg || !out || n_codes == 0 || n_codes > 64) return set_error(nullptr, LDPC_ERR_ARG, "bad argument");
    *out = nullptr;
    if (cfg->profile != LDPC_PROFILE_T) return set_error(nullptr, LDPC_ERR_UNSUPPORTED, "batches run Profile T");
    if (cfg->max_iter < 1 || cfg->max_iter > 100000) return set_error(nullptr, LDPC_ERR_ARG, "max_iter out of range");
    if (!(cfg->beta > 0.0f) || !std::isfinite(cfg->beta)) return set_error(nullptr, LDPC_ERR_ARG, "beta must be > 0");
    if (cfg->l_max != 0.0f || cfg->r_max != 0.0f || cfg->l_post_max != 0.0f || cfg->gamma != 0.0f)
        return set_error(nullptr, LDPC_ERR_ARG, "clamps belong to Profile R");
    std::vector<const ldpc::MultiVariant *> mv(n_codes);
    std::vector<ldpc_code_info> info(n_codes);
    for (uint32_t i = 0; i < n_codes; ++i) {
        int k_id, rate_id;
        if (!ldpc::ar4ja_ids(codes[i], k_id, rate_id))
            return set_error(nullptr, LDPC_ERR_ARG, "batch member " + std::to_string(i) + " is not an AR4JA code");
        const uint32_t v = variants && variants[i] ? variants[i] : ldpc::default_variant(k_id, rate_id);
        mv[i] = ldpc::find_variant(k_id, rate_id, v);
        if (!mv[i])
            return set_error(nullptr, LDPC_ERR_UNSUPPORTED, "batch member " + std::to_string(i) + ": no such variant");
        ldpc::fill_code_info(k_id, rate_id, info[i]);
        info[i].max_iter = cfg->max_iter;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return set_error(nullptr, LDPC_ERR_NODEV, "no HIP device");
    if (device < 0 || device >= ndev) return set_error(nullptr, LDPC_ERR_ARG, "device index out of range");

    ldpc_batch *b = new ldpc_batch();
    b->device = device;
    b->n_codes = n_codes;
    b->nh = cfg->max_iter + 1;
    b->cfg = *cfg;
    b->mv = mv;
    b->info = info;
    auto bail = [&](int rc) {
        g_batch_error = b->err;
        ldpc_batch_destroy(b);
        return rc;
    };
    int rc;
#define B_STEP(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) {                                                                       \
            b->err = std::string(#expr ": ") + hipGetErrorString(e_);                                 \
            return bail(LDPC_ERR_HIP);                                                                \
        }                                                                                             \
    } while (0)
    B_STEP(hipSetDevice(device));
    int cus = 0;
    B_STEP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    b->grid.resize(n_codes);
    for (uint32_t i = 0; i < n_codes; ++i) {
        int occ = 0;
        B_STEP(mv[i]->occupancy(&occ));
        b->grid[i] = cus * (occ > 0 ? occ : 1);
    }
    B_STEP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    B_STEP(hipEventCreate(&b->ev0));
    B_STEP(hipEventCreate(&b->ev1));
    B_STEP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    b->lane.resize(n_codes);
    bool any_group = false, any_wave = false;
    for (uint32_t i = 0; i < n_codes; ++i) (mv[i]->cw_per_wg > 1 ? any_wave : any_group) = true;
    for (uint32_t i = 0; i < n_codes; ++i) b->lane[i] = (any_group && any_wave && mv[i]->cw_per_wg > 1) ? 1 : 0;
    b->side.assign(any_group && any_wave ? 2 : 1, nullptr);  // default priority, like every stream here
    b->join.assign(b->side.size(), nullptr);
    for (size_t k = 0; k < b->side.size(); ++k) {
        B_STEP(hipStreamCreateWithFlags(&b->side[k], hipStreamNonBlocking));
        B_STEP(hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming));
    }
    B_STEP(hipMalloc(&b->d_queue, n_codes * sizeof(uint32_t)));
    B_STEP(hipMalloc(&b->d_counters, 4 * n_codes * sizeof(unsigned long long)));
    B_STEP(hipMalloc(&b->d_hist, (size_t)n_codes * b->nh * sizeof(unsigned long long)));
    B_STEP(hipMalloc(&b->d_dyn, n_codes * sizeof(ldpc::MultiDyn)));
    B_STEP(hipHostMalloc(&b->h_dyn, n_codes * sizeof(ldpc::MultiDyn), hipHostMallocDefault));
    B_STEP(hipHostMalloc(&b->h_counters, 4 * n_codes * sizeof(unsigned long long), hipHostMallocDefault));
    B_STEP(hipHostMalloc(&b->h_hist, (size_t)n_codes * b->nh * sizeof(unsigned long long), hipHostMallocDefault));
#undef B_STEP
    for (uint32_t i = 0; i < n_codes; ++i) b->h_dyn[i] = ldpc::MultiDyn{};
    if ((rc = capture_graph(b))) return bail(rc);
    *out = b;
    return LDPC_OK;
}

int ldpc_batch_code_info(const ldpc_batch *b, uint32_t i, ldpc_code_info *info) {
    if (!b || !info || i >= b->n_codes) return LDPC_ERR_ARG;
    *info = b->info[i];
    return LDPC_OK;
}

int ldpc_batch_run(ldpc_batch *b, const ldpc_point_cfg *pts, ldpc_point_result *res, uint64_t *hist) {
    if (!b) return set_error(nullptr, LDPC_ERR_ARG, "null batch");
    if (!pts || !res) return set_error(&b->err, LDPC_ERR_ARG, "null argument");
    const uint32_t n = b->n_codes;
    std::vector<uint64_t> chunk(n), done(n, 0), fe(n, 0), next(n);
    std::vector<float> mu(n), sigma(n);
    std::vector<bool> finished(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (!std::isfinite(pts[i].ebn0_db)) return set_error(&b->err, LDPC_ERR_ARG, "ebn0_db must be finite");
        chunk[i] = pts[i].chunk_cw ? pts[i].chunk_cw : (1ull << 20);
        if (chunk[i] > (1ull << 31)) return set_error(&b->err, LDPC_ERR_ARG, "chunk_cw must be <= 2^31");
        ldpc::channel_params(b->info[i], pts[i].ebn0_db, &mu[i], &sigma[i]);
        next[i] = pts[i].first_cw;
        finished[i] = pts[i].max_blocks == 0;
    }
    B_TRY(b, hipSetDevice(b->device));
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<uint64_t> ns(n, 0);
    while (std::find(finished.begin(), finished.end(), false) != finished.end()) {
        std::vector<uint32_t> cnt(n, 0);
        for (uint32_t i = 0; i < n; ++i) {
            if (!finished[i]) cnt[i] = (uint32_t)std::min<uint64_t>(chunk[i], pts[i].max_blocks - done[i]);
            b->h_dyn[i] = ldpc::MultiDyn{pts[i].seed, next[i], cnt[i], mu[i], sigma[i], 0u};
        }
        B_TRY(b, hipEventRecord(b->ev0, b->stream));
        B_TRY(b, hipGraphLaunch(b->exec, b->stream));
        B_TRY(b, hipEventRecord(b->ev1, b->stream));
        B_TRY(b, hipStreamSynchronize(b->stream));
        float ms = 0.0f;
        B_TRY(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
        for (uint32_t i = 0; i < n; ++i) {
            if (!cnt[i]) continue;
            const unsigned long long *c = b->h_counters + 4 * i;
            if (c[0] != cnt[i])
                return set_error(&b->err, LDPC_ERR_HIP, "batch kernel processed an unexpected number of codewords");
            res[i].blocks += c[0];
            res[i].frame_errors += c[1];
            res[i].bit_errors += c[2];
            res[i].iter_sum += c[3];
            res[i].kernel_ns += (uint64_t)((double)ms * 1e6);
            res[i].launches += 1;
            if (hist)
                for (uint32_t k = 0; k < b->nh; ++k) hist[(size_t)i * b->nh + k] += b->h_hist[(size_t)i * b->nh + k];
            done[i] += cnt[i];
            fe[i] += c[1];
            next[i] += cnt[i];
            finished[i] = done[i] >= pts[i].max_blocks || (pts[i].target_fe && fe[i] >= pts[i].target_fe);
        }
    }
    const uint64_t el = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                            std::chrono::steady_clock::now() - t0)
                            .count();
    for (uint32_t i = 0; i < n; ++i) {
        res[i].next_cw = next[i];
        res[i].elapsed_ns += el;
    }
    return LDPC_OK;
}

}  // extern "C"
