// kernels.h -- host-side launchers of the HIP kernels (internal to libldpc).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ldpc {

// Profile T, AR4JA r1/2 K=1024 (decode_t.hip)
int t12_waves_per_wg();
hipError_t t12_occupancy(int *wg_per_cu);
hipError_t t12_occupancy_decode(int *wg_per_cu);
hipError_t t12_launch_decode(const uint16_t *llr_in, uint16_t *llr_out, uint32_t *iters, uint8_t *ok, uint32_t n_cw,
                             int max_iter, float beta, uint32_t *queue, int grid, hipStream_t st);
hipError_t t12_launch_mc(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, int max_iter,
                         float beta, uint32_t *queue, unsigned long long *counters, unsigned long long *hist, int grid,
                         hipStream_t st);
hipError_t t12_launch_gen(uint64_t seed, uint64_t first_cw, uint32_t n_cw, float mu, float sigma, uint16_t *out,
                          hipStream_t st);
int t12_schedule(int32_t *out, int cap);

hipError_t launch_fill_u64(unsigned long long *p, size_t n, unsigned long long v, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t st);

}  // namespace ldpc
