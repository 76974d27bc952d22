// noalloc.hip -- ldpc_decode_llrs_device_noalloc: the Profile T decode of
// device buffers without a context and without any device allocation, for
// callers that cannot allocate (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// The decode kernels need scratch memory beyond their inputs and outputs: the
// work-queue counter and, for the nine-code kernels, the tally counters and
// the iteration histogram.  ldpc_decode_llrs_device_async takes them from its
// context; here they live in one file-scope device array per GPU, `scratch`.
// Its device address is resolved once per device with hipGetSymbolAddress
// (the array's name is only a host-side stand-in) and that pointer is what
// the kernels receive.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/ldpc_mi355x.h"
#include "kernels.h"
#include "runtime_internal.h"

namespace {

// Layout: [0, 4) queue counter; [256, 288) four u64 tally counters;
// [512, 512 + 8 (max_iter + 1)) the u64 iteration histogram.
constexpr size_t kQueueOff = 0, kCountersOff = 256, kHistOff = 512;
constexpr size_t kScratchBytes = kHistOff + 8 * ((size_t)LDPC_NOALLOC_MAX_ITER + 1);
static_assert(kScratchBytes <= 65536, "scratch stays a modest fixed size");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];  // device address of `scratch` on each device (null: not resolved yet)
int g_cus[kMaxDevices];

int status(hipError_t e, const char *what) {
    if (e == hipSuccess) return LDPC_OK;
    ldpc::set_context_free_error(std::string(what) + ": " + hipGetErrorString(e));
    return LDPC_ERR_HIP;
}

int arg_error(int code, const char *msg) {
    ldpc::set_context_free_error(msg);
    return code;
}

// Resolve the array's address and the CU count on `device` (current).
int device_setup(int device, unsigned char **ptr, int *cus) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void *p = nullptr;
        int rc = status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        if (rc) return rc;
        int n = 0;
        if ((rc = status(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device),
                         "hipDeviceGetAttribute")))
            return rc;
        g_cus[device] = n;
        g_ptr[device] = p;
    }
    *ptr = static_cast<unsigned char *>(g_ptr[device]);
    *cus = g_cus[device];
    return LDPC_OK;
}

}  // namespace

extern "C" int ldpc_decode_llrs_device_noalloc(const ldpc_code_desc *code, const ldpc_decoder_cfg *cfg, int device,
                                               const void *d_llr, uint32_t n_cw, void *d_out_llr, void *d_out_iters,
                                               void *d_out_ok, void *stream) {
    if (!code || !cfg) return arg_error(LDPC_ERR_ARG, "null argument");
    int k_id = -1, rate_id = -1;
    if (!ldpc::ar4ja_ids(*code, k_id, rate_id)) return arg_error(LDPC_ERR_ARG, "no AR4JA code with this (k_info, rate)");
    if (cfg->profile != LDPC_PROFILE_T)
        return arg_error(LDPC_ERR_UNSUPPORTED, "the allocation-free decode runs Profile T only");
    if (cfg->max_iter < 1 || cfg->max_iter > LDPC_NOALLOC_MAX_ITER)
        return arg_error(LDPC_ERR_ARG, "max_iter must be 1..LDPC_NOALLOC_MAX_ITER (8000): the iteration "
                                       "histogram lives in a fixed 64 KiB device array");
    if (!(cfg->beta > 0.0f) || !std::isfinite(cfg->beta)) return arg_error(LDPC_ERR_ARG, "beta must be > 0");
    if (cfg->l_max != 0.0f || cfg->r_max != 0.0f || cfg->l_post_max != 0.0f || cfg->gamma != 0.0f)
        return arg_error(LDPC_ERR_ARG, "clamps belong to Profile R; the golden Profile T has none");
    const bool specialised = k_id == 0 && rate_id == 0 && cfg->variant == 0;
    const ldpc::MultiVariant *mv = nullptr;
    if (!specialised) {
        mv = ldpc::find_variant(k_id, rate_id, cfg->variant ? cfg->variant : ldpc::default_variant(k_id, rate_id));
        if (!mv) return arg_error(LDPC_ERR_UNSUPPORTED, "no such Profile T kernel variant for this code");
        if (mv->work_per_wg) return arg_error(LDPC_ERR_UNSUPPORTED, "this kernel variant needs a workspace");
    }
    if (device < 0 || device >= kMaxDevices) return arg_error(LDPC_ERR_ARG, "device index out of range");
    if (n_cw == 0) return LDPC_OK;
    if (!d_llr || !d_out_llr || !d_out_iters || !d_out_ok) return arg_error(LDPC_ERR_ARG, "null buffer");
    int rc;
    if ((rc = status(hipSetDevice(device), "hipSetDevice"))) return rc;
    unsigned char *base = nullptr;
    int cus = 0;
    if ((rc = device_setup(device, &base, &cus))) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    uint32_t *queue = reinterpret_cast<uint32_t *>(base + kQueueOff);
    auto *counters = reinterpret_cast<unsigned long long *>(base + kCountersOff);
    auto *hist = reinterpret_cast<unsigned long long *>(base + kHistOff);
    // zero what the kernels count from, with fill kernels (no memset nodes
    // if the caller captures this into a graph)
    if ((rc = status(ldpc::launch_fill_u32(queue, 1, 0u, st), "fill"))) return rc;
    int occ = 0;
    if (specialised) {
        if ((rc = status(ldpc::t12_occupancy_decode(&occ), "occupancy"))) return rc;
        return status(ldpc::t12_launch_decode((const uint16_t *)d_llr, (uint16_t *)d_out_llr, (uint32_t *)d_out_iters,
                                              (uint8_t *)d_out_ok, n_cw, (int)cfg->max_iter, cfg->beta, queue,
                                              cus * (occ > 0 ? occ : 1), st),
                      "launch");
    }
    if ((rc = status(ldpc::launch_fill_u64(counters, 4, 0ull, st), "fill"))) return rc;
    if ((rc = status(ldpc::launch_fill_u64(hist, cfg->max_iter + 1, 0ull, st), "fill"))) return rc;
    if ((rc = status(mv->occupancy(&occ), "occupancy"))) return rc;
    ldpc::MultiParams p{};
    p.llr_in = (const uint16_t *)d_llr;
    p.llr_out = (uint16_t *)d_out_llr;
    p.iters = (uint32_t *)d_out_iters;
    p.ok = (uint8_t *)d_out_ok;
    p.queue = queue;
    p.counters = counters;
    p.hist = hist;
    p.n_cw = n_cw;
    p.max_iter = (int)cfg->max_iter;
    p.beta = cfg->beta;
    p.beta3 = cfg->beta < 1.0f ? cfg->beta : 1.0f;  // stepper:300-304
    return status(mv->launch(p, cus * (occ > 0 ? occ : 1), st), "launch");
}
