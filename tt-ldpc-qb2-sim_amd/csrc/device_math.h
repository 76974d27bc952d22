// device_math.h -- bit-exact fp32/bf16 arithmetic shared by the HIP kernels.
//
// Every floating-point operation here must be a single IEEE binary32 add,
// sub or mul (round to nearest even), never fused: the file is compiled with
// -ffp-contract=off and additionally pins contraction off with the pragma.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace ldpc {

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }

// Round-to-nearest-even to bf16, result kept as an fp32 bit pattern with the
// low 16 bits cleared:  (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
// (engines/tile_matrix_4cw/sfpi_host_emu.h:30-40; ldpc_tile_matrix_core.h:438-446).
// For finite u the carry never reaches bit 31, so rounding commutes with the
// sign bit: rne(sign | m) == sign | rne(m).
__device__ __forceinline__ uint32_t rne_bits(float x) {
    uint32_t u = f2u(x);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
}

// Approximate min* of two magnitudes (a, b >= +0), Profile T
// (sfpi_host_emu.h:566-584): two correction lines with bf16-truncated
// constants, corr = max(line1, line2, 0), result max(min - corr, 0).
//   C1 = trunc_bf16(0.69315)  = 0.69140625    K1 = trunc_bf16(-0.4)   = -0.3984375
//   C2 = trunc_bf16(0.65715)  = 0.65625       K2 = trunc_bf16(-0.328) = -0.326171875
// |a - b| equals max(a,b) - min(a,b) exactly (IEEE subtraction is sign
// symmetric).  No -0 can reach the min/max ops (inputs are fabs values or
// earlier results >= +0, and x - x = +0), so v_min/v_max agree with the
// reference's std::min/std::max on every input.
__device__ __forceinline__ float amin_t(float a, float b) {
    const float C1 = 0.69140625f, K1 = -0.3984375f;
    const float C2 = 0.65625f, K2 = -0.326171875f;
    float mn = fminf(a, b);
    float d = fabsf(a - b);
    float l1 = __fmul_rn(K1, d);
    l1 = __fadd_rn(l1, C1);
    float l2 = __fmul_rn(K2, d);
    l2 = __fadd_rn(l2, C2);
    float corr = fmaxf(fmaxf(l1, l2), 0.0f);
    return fmaxf(__fsub_rn(mn, corr), 0.0f);
}

// Two fp32 values -> packed bf16 pair (lo = a, hi = b), round to nearest
// even.  v_cvt_pk_bf16_f32 was checked exhaustively against the integer
// formula (u + 0x7FFF + lsb) >> 16 on every non-NaN fp32 pattern (2^32 - 2^24
// + 2 inputs, 0 mismatches, MI355X); NaN never reaches it here.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t cvt_pk_bf16(float a, float b) {
    const bf16x2_t r = __builtin_convertvector((f32x2_t){a, b}, bf16x2_t);
    return __builtin_bit_cast(uint32_t, r);
}
__device__ __forceinline__ float lo_f(uint32_t p) { return u2f(p << 16); }
__device__ __forceinline__ float hi_f(uint32_t p) { return u2f(p & 0xFFFF0000u); }

// ---------------------------------------------------------------------------
// Channel noise: Philox4x32-10 (Salmon et al., SC'11) + Box-Muller with
// polynomial log / sin / cos and a Newton reciprocal square root, every step
// a plain fp32 add/sub/mul so the CPU oracle reproduces the same bits.
// ---------------------------------------------------------------------------
struct u32x4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {c0, c1, c2, c3};
}

__device__ __forceinline__ float log_unit(float u) {  // ln u, u in (0, 1]
    const float Q0 = 0x1.000000p+0f, Q1 = -0x1.fffff8p-2f, Q2 = 0x1.555550p-2f, Q3 = -0x1.000428p-2f,
                Q4 = 0x1.99a476p-3f, Q5 = -0x1.5427a8p-3f, Q6 = 0x1.226874p-3f, Q7 = -0x1.0f2c26p-3f,
                Q8 = 0x1.087b72p-3f, Q9 = -0x1.38e072p-4f;
    uint32_t b = f2u(u);
    int e = (int)(b >> 23) - 127;
    float m = u2f((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) {
        m = __fmul_rn(m, 0.5f);
        e += 1;
    }
    float f = __fsub_rn(m, 1.0f);
    float q = Q9;
    q = __fadd_rn(__fmul_rn(q, f), Q8);
    q = __fadd_rn(__fmul_rn(q, f), Q7);
    q = __fadd_rn(__fmul_rn(q, f), Q6);
    q = __fadd_rn(__fmul_rn(q, f), Q5);
    q = __fadd_rn(__fmul_rn(q, f), Q4);
    q = __fadd_rn(__fmul_rn(q, f), Q3);
    q = __fadd_rn(__fmul_rn(q, f), Q2);
    q = __fadd_rn(__fmul_rn(q, f), Q1);
    q = __fadd_rn(__fmul_rn(q, f), Q0);
    float p = __fmul_rn(f, q);
    float el = __fmul_rn((float)e, 0x1.62e430p-1f);
    return __fadd_rn(el, p);
}

__device__ __forceinline__ float sqrt_newton(float t) {  // t >= 0
    float y = u2f(0x5F3759DFu - (f2u(t) >> 1));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float h = __fmul_rn(0.5f, t);
        float yy = __fmul_rn(y, y);
        float g = __fmul_rn(h, yy);
        float c = __fsub_rn(1.5f, g);
        y = __fmul_rn(y, c);
    }
    return __fmul_rn(t, y);
}

__device__ __forceinline__ void sincos_turn(uint32_t x, float &so, float &co) {
    const float S1 = -0x1.555556p-3f, S2 = 0x1.111112p-7f, S3 = -0x1.a01a02p-13f, S4 = 0x1.71de3ap-19f,
                S5 = -0x1.ae6456p-26f, S6 = 0x1.612462p-33f;
    const float C1 = -0x1.000000p-1f, C2 = 0x1.555556p-5f, C3 = -0x1.6c16c2p-10f, C4 = 0x1.a01a02p-16f,
                C5 = -0x1.27e4fcp-22f, C6 = 0x1.1eed8ep-29f, C7 = -0x1.93974ap-37f;
    uint32_t quad = x >> 30;
    float f = __fmul_rn((float)(x & 0x3FFFFFFFu), 0x1p-30f);
    float psi = __fmul_rn(f, 0x1.921fb6p+0f);
    float p2 = __fmul_rn(psi, psi);
    float sp = S6;
    sp = __fadd_rn(__fmul_rn(sp, p2), S5);
    sp = __fadd_rn(__fmul_rn(sp, p2), S4);
    sp = __fadd_rn(__fmul_rn(sp, p2), S3);
    sp = __fadd_rn(__fmul_rn(sp, p2), S2);
    sp = __fadd_rn(__fmul_rn(sp, p2), S1);
    float s = __fmul_rn(psi, p2);
    s = __fmul_rn(s, sp);
    s = __fadd_rn(psi, s);
    float cp = C7;
    cp = __fadd_rn(__fmul_rn(cp, p2), C6);
    cp = __fadd_rn(__fmul_rn(cp, p2), C5);
    cp = __fadd_rn(__fmul_rn(cp, p2), C4);
    cp = __fadd_rn(__fmul_rn(cp, p2), C3);
    cp = __fadd_rn(__fmul_rn(cp, p2), C2);
    cp = __fadd_rn(__fmul_rn(cp, p2), C1);
    float c = __fmul_rn(p2, cp);
    c = __fadd_rn(1.0f, c);
    // quadrant rotation: (c, s), (-s, c), (-c, -s), (s, -c)
    float a0 = (quad & 1u) ? s : c;
    float a1 = (quad & 1u) ? c : s;
    co = ((quad + 1u) & 2u) ? -a0 : a0;
    so = (quad & 2u) ? -a1 : a1;
}

__device__ __forceinline__ void box_muller(uint32_t x0, uint32_t x1, float &z0, float &z1) {
    float u = __fmul_rn((float)x0, 0x1p-32f);
    u = __fadd_rn(u, 0x1p-33f);
    float l = log_unit(u);
    float t = __fmul_rn(-2.0f, l);
    float r = sqrt_newton(t);
    float s, c;
    sincos_turn(x1, s, c);
    z0 = __fmul_rn(r, c);
    z1 = __fmul_rn(r, s);
}

// LLR of one transmitted symbol of the all-zero codeword: RNE_bf16(mu + sigma*z).
__device__ __forceinline__ uint32_t llr_bits(float mu, float sigma, float z) {
    return rne_bits(__fadd_rn(mu, __fmul_rn(sigma, z)));
}

}  // namespace ldpc
