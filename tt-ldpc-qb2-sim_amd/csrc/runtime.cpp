// runtime.cpp -- the C ABI of libldpc (include/ldpc_mi355x.h): contexts,
// device buffers, the chunked Monte Carlo driver and its stopping rule.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ldpc_mi355x.h"
#include "kernels.h"

namespace {

std::string g_create_error;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct ldpc_ctx {
    int device = 0;
    ldpc_code_desc code{};
    ldpc_decoder_cfg cfg{};
    ldpc_code_info info{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int grid_mc = 0, grid_dec = 0;
    uint32_t *d_queue = nullptr;
    unsigned long long *d_counters = nullptr;  // [4]
    unsigned long long *d_hist = nullptr;      // [max_iter + 1]
    DevBuf in, out, iters, ok;
    std::string err;
};

namespace {

int fail(ldpc_ctx *c, int code, const std::string &msg) {
    if (c)
        c->err = msg;
    else
        g_create_error = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return fail((ctx), LDPC_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_));         \
    } while (0)

int ensure(ldpc_ctx *c, DevBuf &b, size_t bytes) {
    if (b.bytes >= bytes) return LDPC_OK;
    if (b.p) HIP_TRY(c, hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
    HIP_TRY(c, hipMalloc(&b.p, bytes));
    b.bytes = bytes;
    return LDPC_OK;
}

bool is_r12_k1024(const ldpc_code_desc &d) {
    return d.family == LDPC_CODE_AR4JA && d.k_info == 1024 && d.rate_num == 1 && d.rate_den == 2;
}

}  // namespace

extern "C" {

const char *ldpc_last_error(const ldpc_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int ldpc_create(const ldpc_code_desc *code, const ldpc_decoder_cfg *cfg, int device, ldpc_ctx **out) {
    if (!code || !cfg || !out) return fail(nullptr, LDPC_ERR_ARG, "null argument");
    *out = nullptr;
    if (code->family != LDPC_CODE_AR4JA) return fail(nullptr, LDPC_ERR_ARG, "unknown code family");
    const bool known_k = code->k_info == 1024 || code->k_info == 4096 || code->k_info == 16384;
    const bool known_r = (code->rate_num == 1 && code->rate_den == 2) || (code->rate_num == 2 && code->rate_den == 3) ||
                         (code->rate_num == 4 && code->rate_den == 5);
    if (!known_k || !known_r) return fail(nullptr, LDPC_ERR_ARG, "no AR4JA code with this (k_info, rate)");
    if (cfg->profile != LDPC_PROFILE_T) return fail(nullptr, LDPC_ERR_UNSUPPORTED, "unknown decoder profile");
    if (cfg->max_iter < 1 || cfg->max_iter > 100000) return fail(nullptr, LDPC_ERR_ARG, "max_iter out of range");
    if (!(cfg->beta > 0.0f) || !std::isfinite(cfg->beta)) return fail(nullptr, LDPC_ERR_ARG, "beta must be > 0");
    if (!is_r12_k1024(*code))
        return fail(nullptr, LDPC_ERR_UNSUPPORTED, "Profile T kernels are built for AR4JA r1/2 K=1024 only");

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(nullptr, LDPC_ERR_NODEV, "no HIP device");
    if (device < 0 || device >= ndev) return fail(nullptr, LDPC_ERR_ARG, "device index out of range");

    ldpc_ctx *c = new ldpc_ctx();
    c->device = device;
    c->code = *code;
    c->cfg = *cfg;
    c->info.n = 2560;
    c->info.m = 1536;
    c->info.k_info = 1024;
    c->info.n_tx = 2048;
    c->info.z = 128;
    c->info.n_layers = 12;
    c->info.n_edges = 7680;
    c->info.max_iter = cfg->max_iter;

    auto bail = [&](hipError_t e, const char *what) {
        g_create_error = std::string(what) + ": " + hipGetErrorString(e);
        ldpc_destroy(c);
        return LDPC_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return bail(e, "hipSetDevice");
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return bail(e, "hipStreamCreate");
    if ((e = hipEventCreate(&c->ev0)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreate(&c->ev1)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipMalloc(&c->d_queue, sizeof(uint32_t))) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = hipMalloc(&c->d_counters, 4 * sizeof(unsigned long long))) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = hipMalloc(&c->d_hist, (cfg->max_iter + 1) * sizeof(unsigned long long))) != hipSuccess)
        return bail(e, "hipMalloc");
    int cus = 0, occ_mc = 0, occ_dec = 0;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
        return bail(e, "hipDeviceGetAttribute");
    if ((e = ldpc::t12_occupancy(&occ_mc)) != hipSuccess) return bail(e, "occupancy");
    if ((e = ldpc::t12_occupancy_decode(&occ_dec)) != hipSuccess) return bail(e, "occupancy");
    c->grid_mc = cus * (occ_mc > 0 ? occ_mc : 1);
    c->grid_dec = cus * (occ_dec > 0 ? occ_dec : 1);
    *out = c;
    return LDPC_OK;
}

void ldpc_destroy(ldpc_ctx *c) {
    if (!c) return;
    // Teardown is best effort: errors here have no caller to report to.
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (DevBuf *b : {&c->in, &c->out, &c->iters, &c->ok})
        if (b->p) (void)hipFree(b->p);
    if (c->d_queue) (void)hipFree(c->d_queue);
    if (c->d_counters) (void)hipFree(c->d_counters);
    if (c->d_hist) (void)hipFree(c->d_hist);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int ldpc_get_code_info(const ldpc_ctx *c, ldpc_code_info *info) {
    if (!c || !info) return LDPC_ERR_ARG;
    *info = c->info;
    return LDPC_OK;
}

int ldpc_get_schedule(const ldpc_code_desc *code, int32_t *out, int32_t cap) {
    if (!code || !out || cap <= 0) return LDPC_ERR_ARG;
    if (!is_r12_k1024(*code)) return LDPC_ERR_UNSUPPORTED;
    int n = ldpc::t12_schedule(out, cap);
    return n < 0 ? LDPC_ERR_ARG : n;
}

int ldpc_channel_params(const ldpc_ctx *c, double ebn0_db, float *mu, float *sigma) {
    if (!c || !mu || !sigma || !std::isfinite(ebn0_db)) return LDPC_ERR_ARG;
    const double rate = (double)c->info.k_info / (double)c->info.n_tx;
    const double g = std::pow(10.0, ebn0_db / 10.0);
    *mu = (float)(4.0 * rate * g);
    *sigma = (float)std::sqrt(8.0 * rate * g);
    return LDPC_OK;
}

int ldpc_sync(ldpc_ctx *c) {
    if (!c) return LDPC_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LDPC_OK;
}

void *ldpc_stream(ldpc_ctx *c) { return c ? (void *)c->stream : nullptr; }

int ldpc_run_point(ldpc_ctx *c, const ldpc_point_cfg *pt, ldpc_point_result *res, uint64_t *hist) {
    if (!c || !pt || !res) return fail(c, LDPC_ERR_ARG, "null argument");
    if (!std::isfinite(pt->ebn0_db)) return fail(c, LDPC_ERR_ARG, "ebn0_db must be finite");
    const uint64_t chunk = pt->chunk_cw ? pt->chunk_cw : (1ull << 20);
    // the device queue counter is 32-bit and also counts one exit ticket per wave
    if (chunk > (1ull << 31)) return fail(c, LDPC_ERR_ARG, "chunk_cw must be <= 2^31");
    HIP_TRY(c, hipSetDevice(c->device));
    float mu, sigma;
    ldpc_channel_params(c, pt->ebn0_db, &mu, &sigma);
    const uint32_t nh = c->cfg.max_iter + 1;
    std::vector<unsigned long long> h(nh);
    unsigned long long cnt[4];
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t done = 0, fe = 0;
    uint64_t next = pt->first_cw;
    while (done < pt->max_blocks && (pt->target_fe == 0 || fe < pt->target_fe)) {
        const uint64_t n = std::min<uint64_t>(chunk, pt->max_blocks - done);
        HIP_TRY(c, ldpc::launch_fill_u32(c->d_queue, 1, 0u, c->stream));
        HIP_TRY(c, ldpc::launch_fill_u64(c->d_counters, 4, 0ull, c->stream));
        HIP_TRY(c, ldpc::launch_fill_u64(c->d_hist, nh, 0ull, c->stream));
        HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
        HIP_TRY(c, ldpc::t12_launch_mc(pt->seed, next, (uint32_t)n, mu, sigma, (int)c->cfg.max_iter, c->cfg.beta,
                                       c->d_queue, c->d_counters, c->d_hist, c->grid_mc, c->stream));
        HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
        HIP_TRY(c, hipMemcpyAsync(cnt, c->d_counters, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(h.data(), c->d_hist, nh * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                  c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        float ms = 0.0f;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        if (cnt[0] != n) return fail(c, LDPC_ERR_HIP, "kernel processed an unexpected number of codewords");
        res->blocks += cnt[0];
        res->frame_errors += cnt[1];
        res->bit_errors += cnt[2];
        res->iter_sum += cnt[3];
        res->kernel_ns += (uint64_t)((double)ms * 1e6);
        res->launches += 1;
        if (hist)
            for (uint32_t i = 0; i < nh; ++i) hist[i] += h[i];
        done += n;
        fe += cnt[1];
        next += n;
    }
    res->next_cw = next;
    res->elapsed_ns += (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                           std::chrono::steady_clock::now() - t0)
                           .count();
    return LDPC_OK;
}

int ldpc_decode_llrs_device_async(ldpc_ctx *c, const void *d_llr, uint32_t n_cw, void *d_out_llr, void *d_out_iters,
                                  void *d_out_ok) {
    if (!c) return LDPC_ERR_ARG;
    if (n_cw == 0) return LDPC_OK;
    if (!d_llr || !d_out_llr || !d_out_iters || !d_out_ok) return fail(c, LDPC_ERR_ARG, "null buffer");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, ldpc::launch_fill_u32(c->d_queue, 1, 0u, c->stream));
    HIP_TRY(c, ldpc::t12_launch_decode((const uint16_t *)d_llr, (uint16_t *)d_out_llr, (uint32_t *)d_out_iters,
                                       (uint8_t *)d_out_ok, n_cw, (int)c->cfg.max_iter, c->cfg.beta, c->d_queue,
                                       c->grid_dec, c->stream));
    return LDPC_OK;
}

int ldpc_decode_llrs(ldpc_ctx *c, const uint16_t *llr, uint32_t n_cw, uint16_t *out_llr, uint32_t *out_iters,
                     uint8_t *out_ok) {
    if (!c) return LDPC_ERR_ARG;
    if (n_cw == 0) return LDPC_OK;
    if (!llr || !out_llr || !out_iters || !out_ok) return fail(c, LDPC_ERR_ARG, "null buffer");
    const size_t nb = (size_t)n_cw * c->info.n * sizeof(uint16_t);
    int rc;
    if ((rc = ensure(c, c->in, nb)) || (rc = ensure(c, c->out, nb)) ||
        (rc = ensure(c, c->iters, (size_t)n_cw * sizeof(uint32_t))) || (rc = ensure(c, c->ok, (size_t)n_cw)))
        return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(c->in.p, llr, nb, hipMemcpyHostToDevice, c->stream));
    if ((rc = ldpc_decode_llrs_device_async(c, c->in.p, n_cw, c->out.p, c->iters.p, c->ok.p))) return rc;
    HIP_TRY(c, hipMemcpyAsync(out_llr, c->out.p, nb, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(out_iters, c->iters.p, (size_t)n_cw * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(out_ok, c->ok.p, (size_t)n_cw, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LDPC_OK;
}

int ldpc_gen_channel(ldpc_ctx *c, uint64_t seed, uint64_t first_cw, uint32_t n_cw, double ebn0_db, uint16_t *out_llr) {
    if (!c) return LDPC_ERR_ARG;
    if (n_cw == 0) return LDPC_OK;
    if (!out_llr) return fail(c, LDPC_ERR_ARG, "null buffer");
    float mu, sigma;
    if (ldpc_channel_params(c, ebn0_db, &mu, &sigma)) return fail(c, LDPC_ERR_ARG, "ebn0_db must be finite");
    const size_t nb = (size_t)n_cw * c->info.n * sizeof(uint16_t);
    int rc;
    if ((rc = ensure(c, c->out, nb))) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, ldpc::t12_launch_gen(seed, first_cw, n_cw, mu, sigma, (uint16_t *)c->out.p, c->stream));
    HIP_TRY(c, hipMemcpyAsync(out_llr, c->out.p, nb, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LDPC_OK;
}

}  // extern "C"
