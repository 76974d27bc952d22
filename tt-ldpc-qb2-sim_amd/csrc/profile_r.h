// profile_r.h -- arithmetic of the Profile R row update (the scalar_5w
// engine), shared by its two kernels (decode_r.hip: one codeword per wave,
// messages in LDS; decode_r2.hip: two codewords per wave, messages in VGPRs).
//
// engines/scalar_5w/ldpc_decoder_core.h:
//   recursive Approx-Min* with fp32 constants                    (:176-201)
//     (the degree-3 variant :155-173 is the same function: for finite
//      floats a > b <=> a - b > 0, so "(m > c) ? m - c : 0" and
//      "max(m - c, 0)" agree)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.h"

#pragma clang fp contract(off)

namespace ldpc {
namespace pr {

__device__ __forceinline__ float bf(uint16_t h) { return u2f((uint32_t)h << 16); }
__device__ __forceinline__ float absbits(float x) { return u2f(f2u(x) & 0x7FFFFFFFu); }

// Correction of the recursive min* (ldpc_decoder_core.h:162):
// 0 for diff >= 2; 0.69315 - 0.40 diff below 0.5; 0.49315 - 0.328 (diff - 0.5)
__device__ __forceinline__ float corr_r(float diff) {
    const float l1 = __fsub_rn(0.69315f, __fmul_rn(0.40f, diff));
    const float l2 = __fsub_rn(0.49315f, __fmul_rn(0.328f, __fsub_rn(diff, 0.5f)));
    const float c = diff < 0.5f ? l1 : l2;
    return diff >= 2.0f ? 0.0f : c;
}

// One fold step: max(min(m, a) - g(|m - a|), 0).  m, a >= +0 and the
// difference is never -0, so fmaxf agrees with the reference's select.
// For non-negative finite floats the bit patterns order like the values, so
// min(m, a) is an unsigned integer min (no canonicalisation of the inputs).
// The result is >= +0 with the sign bit clear.
__device__ __forceinline__ float fold_r(float m, float a) {
    const float diff = fabsf(__fsub_rn(m, a));
    const float cur = u2f(min(f2u(m), f2u(a)));
    return fmaxf(__fsub_rn(cur, corr_r(diff)), 0.0f);
}

struct RCfg {
    float r_max, l_post_max;
};

}  // namespace pr
}  // namespace ldpc
