// dmin_torch.hip -- differentiable layered min-sum decoder for PyTorch on
// gfx950: forward and hand-written backward kernels plus the checkpointed
// driver, bound as the Python module ldpc_dmin.
//
// Decoder (AR4JA layered schedule, no early exit, T iterations, fp32):
//   q_d = L_v(d) - R_d;  a_d = |q_d|;  (min1, idx1), (min2, idx2) first-index
//   g_d = min2 on idx1, min1 elsewhere (min-sum, corr 0), or with the
//         Approx-Min* correction min*(a, b) = max(a - corr(b - a), 0),
//         corr(x) = max(C1 + K1 x, C2 + K2 x, 0) (bf16-truncated constants of
//         the reference, sfpi_host_emu.h:566-584):
//         corr 1: min2 on idx1, min*(min1, min2) elsewhere;
//         corr 2 (three minima): min*(min2, min3) on idx1, min*(min1, min3)
//         on idx2, min*(min1, min2) elsewhere
//   z_d = alpha_p * g_d - beta_p;  r_d = copysign(max(z_d, 0), s ^ sign q_d),
//   the magnitude capped at r_max if r_max > 0
//   L_v(d) = q_d + r_d;  R_d = r_d
// alpha / beta are learnable per layer slot p (one circulant of the layer),
// optionally per iteration.  The reference has no counterpart.
//
// Layout: one GPU thread per (codeword, check of the layer); L [B, N] and R
// [B, E] fp32 in HBM, R edge e = layer_offset + d * Z + j so that the threads
// of a layer read and write consecutive addresses (as do the cyclic-shift
// variable reads).  One launch per layer step.
//
// Backward: per layer step, given the layer's q (a tape rebuilt from the
// nearest (L, R) checkpoint), the adjoint of L after the layer and of the R
// it wrote, it produces the adjoint of L before the layer and of the R it
// read, and per-workgroup partial sums of the alpha / beta gradients
// (summed afterwards in a fixed order: no float atomics).  Subgradients:
// sign(q) is piecewise constant; d|q|/dq = sign(q) with sign(0) = 0; the
// relu, the cap, the correction's max and the clamp at zero pass the gradient
// only through the selected branch (0 < z < r_max; line 1 if l1 >= l2;
// corrected min > 0);
// the minima route it to their edges (idx1 / idx2 / idx3).
#include <c10/hip/HIPStream.h>
#include <hip/hip_runtime.h>
#include <torch/extension.h>

#include <vector>

#pragma clang fp contract(off)

namespace {

constexpr float C1 = 0.69140625f, K1 = -0.3984375f, C2 = 0.65625f, K2 = -0.326171875f;
constexpr int MAXD = 18;
constexpr int TPB = 256;

struct LayerArgs {
    int z, n, e, e_off, p_off, deg;
    int base[MAXD], shift[MAXD];
};

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }

// min*(lo, hi) of two magnitudes lo <= hi with the Approx-Min* correction:
// max(lo - corr(hi - lo), 0).  cp is d corr / dx of the selected line (0 if
// corr == 0); pos records lo - corr > 0 (the gradient passes only then).
struct Pair {
    float o, cp;
    bool pos;
};
__device__ __forceinline__ Pair amin_pair(float lo, float hi) {
    const float x = __fsub_rn(hi, lo);
    const float l1 = __fadd_rn(__fmul_rn(K1, x), C1);
    const float l2 = __fadd_rn(__fmul_rn(K2, x), C2);
    const float lm = l1 >= l2 ? l1 : l2;
    const float c = lm > 0.0f ? lm : 0.0f;
    Pair p;
    p.cp = lm > 0.0f ? (l1 >= l2 ? K1 : K2) : 0.0f;
    const float o = __fsub_rn(lo, c);
    p.pos = o > 0.0f;
    p.o = p.pos ? o : 0.0f;
    return p;
}

// Check-node core shared by forward and backward.  CM selects the
// magnitude g_d: 0 min-sum (min2 on idx1, min1 elsewhere); 1 the two-minimum
// correction (min2 on idx1, min*(min1, min2) elsewhere); 2 the
// three-minimum correction, each edge's pair taken from the others:
// min*(min2, min3) on idx1, min*(min1, min3) on idx2, min*(min1, min2)
// elsewhere.
template <int D, int CM>
struct Check {
    float q[D];
    int v[D];
    float m1, m2, m3;
    int i1, i2, i3;
    uint32_t s;  // XOR of the sign bits
    Pair p12, p13, p23;

    __device__ __forceinline__ void minima() {
        m1 = __builtin_inff();
        m2 = __builtin_inff();
        m3 = __builtin_inff();
        i1 = 0;
        i2 = 0;
        i3 = 0;
        s = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            s ^= f2u(q[d]) >> 31;
            const float a = fabsf(q[d]);
            if (a < m1) {
                if constexpr (CM == 2) {
                    m3 = m2;
                    i3 = i2;
                }
                m2 = m1;
                i2 = i1;
                m1 = a;
                i1 = d;
            } else if (a < m2) {
                if constexpr (CM == 2) {
                    m3 = m2;
                    i3 = i2;
                }
                m2 = a;
                i2 = d;
            } else if (CM == 2 && a < m3) {
                m3 = a;
                i3 = d;
            }
        }
        if constexpr (CM == 0) {
            p12 = Pair{m1, 0.0f, true};
        } else {
            p12 = amin_pair(m1, m2);
        }
        if constexpr (CM == 2) {
            p13 = amin_pair(m1, m3);
            p23 = amin_pair(m2, m3);
        }
    }
    __device__ __forceinline__ float g(int d) const {
        if constexpr (CM == 2) return d == i1 ? p23.o : (d == i2 ? p13.o : p12.o);
        return d == i1 ? m2 : p12.o;
    }
};

template <int D, int CM>
__global__ void __launch_bounds__(TPB) k_fwd(float *__restrict__ L, float *__restrict__ R, float *__restrict__ tape,
                                             const float *__restrict__ alpha, const float *__restrict__ beta,
                                             LayerArgs la, float r_max, int64_t n_threads) {
    const int64_t gid = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (gid >= n_threads) return;
    const int64_t cw = gid / la.z;
    const int j = (int)(gid % la.z);
    float *Lc = L + cw * la.n;
    float *Rc = R + cw * la.e + la.e_off;
    Check<D, CM> c;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        c.v[d] = la.base[d] + ((j + la.shift[d]) & (la.z - 1));
        c.q[d] = __fsub_rn(Lc[c.v[d]], Rc[d * la.z + j]);
        if (tape) tape[cw * (D * la.z) + d * la.z + j] = c.q[d];
    }
    c.minima();
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float z = __fsub_rn(__fmul_rn(alpha[la.p_off + d], c.g(d)), beta[la.p_off + d]);
        float mag = z > 0.0f ? z : 0.0f;
        if (r_max > 0.0f && !(mag < r_max)) mag = r_max;
        const uint32_t sb = (c.s ^ (f2u(c.q[d]) >> 31)) & 1u;
        const float r = u2f(f2u(mag) | (sb << 31));
        Lc[c.v[d]] = __fadd_rn(c.q[d], r);
        Rc[d * la.z + j] = r;
    }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

template <int D, int CM>
__global__ void __launch_bounds__(TPB) k_bwd(float *__restrict__ gL, float *__restrict__ gR,
                                             const float *__restrict__ tape, const float *__restrict__ alpha,
                                             const float *__restrict__ beta, float *__restrict__ partial,
                                             LayerArgs la, float r_max, int64_t n_threads) {
    __shared__ float red[TPB / 64][2 * D];
    const int64_t gid = (int64_t)blockIdx.x * TPB + threadIdx.x;
    const bool valid = gid < n_threads;
    float pa[D], pb[D];
#pragma unroll
    for (int d = 0; d < D; ++d) pa[d] = pb[d] = 0.0f;
    if (valid) {
        const int64_t cw = gid / la.z;
        const int j = (int)(gid % la.z);
        float *gLc = gL + cw * la.n;
        float *gRc = gR + cw * la.e + la.e_off;
        Check<D, CM> c;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            c.v[d] = la.base[d] + ((j + la.shift[d]) & (la.z - 1));
            c.q[d] = tape[cw * (D * la.z) + d * la.z + j];
        }
        c.minima();
        float dq[D], ga[D];
#pragma unroll
        for (int d = 0; d < D; ++d) ga[d] = 0.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const float gd = c.g(d);
            const float z = __fsub_rn(__fmul_rn(alpha[la.p_off + d], gd), beta[la.p_off + d]);
            const uint32_t sb = (c.s ^ (f2u(c.q[d]) >> 31)) & 1u;
            const float gln = gLc[c.v[d]];
            const float gr = gln + gRc[d * la.z + j];
            const bool pass = z > 0.0f && (r_max <= 0.0f || z < r_max);
            const float gz = pass ? (sb ? -gr : gr) : 0.0f;
            pa[d] = gz * gd;
            pb[d] = -gz;
            const float gg = gz * alpha[la.p_off + d];
            dq[d] = gln;  // L' = q + r
            if constexpr (CM == 2) {
                // the pair of the edge's g: (lo, hi) minima and their correction
                const bool e1 = d == c.i1, e2 = d == c.i2;
                const int lo = e1 ? c.i2 : c.i1, hi = (e1 || e2) ? c.i3 : c.i2;
                const Pair &pr = e1 ? c.p23 : (e2 ? c.p13 : c.p12);
                if (pr.pos) {
                    ga[lo] += gg * (1.0f + pr.cp);
                    ga[hi] += gg * -pr.cp;
                }
            } else if (d == c.i1) {
                ga[c.i2] += gg;
            } else if (c.p12.pos) {
                ga[c.i1] += gg * (1.0f + c.p12.cp);
                ga[c.i2] += gg * -c.p12.cp;
            }
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const float sq = c.q[d] > 0.0f ? 1.0f : (c.q[d] < 0.0f ? -1.0f : 0.0f);
            const float t = dq[d] + ga[d] * sq;
            gLc[c.v[d]] = t;
            gRc[d * la.z + j] = -t;
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float a = wave_sum(pa[d]), b = wave_sum(pb[d]);
        if (lane == 0) {
            red[wv][d] = a;
            red[wv][D + d] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * D) {
        float s = 0.0f;
#pragma unroll
        for (int w = 0; w < TPB / 64; ++w) s += red[w][threadIdx.x];
        partial[(int64_t)blockIdx.x * 2 * D + threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct Sched {
    std::vector<LayerArgs> layers;
    int n = 0, e = 0, p = 0;
};

// tab: int32 CPU tensor [n_layers, 2 + 2 * MAXD] = (deg, z, base0, shift0, ...)
Sched make_sched(const at::Tensor &tab, int64_t n) {
    TORCH_CHECK(tab.device().is_cpu() && tab.scalar_type() == at::kInt && tab.dim() == 2 &&
                    tab.size(1) == 2 + 2 * MAXD,
                "schedule table: int32 CPU tensor [n_layers, 38]");
    auto t = tab.contiguous();
    const int32_t *p = t.data_ptr<int32_t>();
    Sched s;
    s.n = (int)n;
    for (int64_t i = 0; i < t.size(0); ++i) {
        const int32_t *r = p + i * (2 + 2 * MAXD);
        LayerArgs la{};
        la.deg = r[0];
        la.z = r[1];
        TORCH_CHECK(la.deg == 3 || la.deg == 6 || la.deg == 10 || la.deg == 18, "layer degree must be 3, 6, 10, 18");
        TORCH_CHECK(la.z > 0 && (la.z & (la.z - 1)) == 0, "circulant size must be a power of two");
        for (int d = 0; d < la.deg; ++d) {
            la.base[d] = r[2 + 2 * d];
            la.shift[d] = r[3 + 2 * d];
            TORCH_CHECK(la.base[d] >= 0 && la.base[d] + la.z <= n && la.shift[d] >= 0 && la.shift[d] < la.z,
                        "schedule slot out of range");
        }
        la.e_off = s.e;
        la.p_off = s.p;
        la.n = (int)n;
        s.e += la.deg * la.z;
        s.p += la.deg;
        s.layers.push_back(la);
    }
    for (auto &la : s.layers) la.e = s.e;
    return s;
}

template <int D, int CM>
void fwd_launch(float *L, float *R, float *tape, const float *al, const float *be, const LayerArgs &la, float r_max,
                int64_t B, hipStream_t st) {
    const int64_t nt = B * la.z;
    hipLaunchKernelGGL((k_fwd<D, CM>), dim3((unsigned)((nt + TPB - 1) / TPB)), dim3(TPB), 0, st, L, R, tape, al, be,
                       la, r_max, nt);
}

template <int D, int CM>
void bwd_launch(float *gL, float *gR, const float *tape, const float *al, const float *be, float *partial,
                const LayerArgs &la, float r_max, int64_t B, hipStream_t st) {
    const int64_t nt = B * la.z;
    hipLaunchKernelGGL((k_bwd<D, CM>), dim3((unsigned)((nt + TPB - 1) / TPB)), dim3(TPB), 0, st, gL, gR, tape, al,
                       be, partial, la, r_max, nt);
}

#define LDPC_DISPATCH_DEG(deg, cm, F, ...)              \
    do {                                               \
        switch (deg) {                                 \
            case 3: F<3, cm>(__VA_ARGS__); break;      \
            case 6: F<6, cm>(__VA_ARGS__); break;      \
            case 10: F<10, cm>(__VA_ARGS__); break;    \
            default: F<18, cm>(__VA_ARGS__); break;    \
        }                                              \
    } while (0)
#define LDPC_DISPATCH(deg, corr, F, ...)                               \
    do {                                                               \
        if (corr == 2)                                                 \
            LDPC_DISPATCH_DEG(deg, 2, F, __VA_ARGS__);                 \
        else if (corr == 1)                                            \
            LDPC_DISPATCH_DEG(deg, 1, F, __VA_ARGS__);                 \
        else                                                           \
            LDPC_DISPATCH_DEG(deg, 0, F, __VA_ARGS__);                 \
    } while (0)

void check_inputs(const at::Tensor &llr, const at::Tensor &alpha, const at::Tensor &beta, const Sched &s,
                  int64_t T) {
    TORCH_CHECK(llr.is_cuda() && llr.scalar_type() == at::kFloat && llr.dim() == 2 && llr.size(1) == s.n &&
                    llr.is_contiguous(),
                "llr: contiguous float32 GPU tensor [B, N]");
    for (const at::Tensor *t : {&alpha, &beta})
        TORCH_CHECK(t->is_cuda() && t->scalar_type() == at::kFloat && t->dim() == 2 && t->size(1) == s.p &&
                        (t->size(0) == 1 || t->size(0) == T) && t->is_contiguous() &&
                        t->device() == llr.device(),
                    "alpha / beta: contiguous float32 GPU tensors [1 or T, P] on the LLRs' device");
    TORCH_CHECK(T >= 1, "T >= 1");
}

void check_corr(int64_t corr) { TORCH_CHECK(corr >= 0 && corr <= 2, "corr: 0 (min-sum), 1 or 2 (minima corrected)"); }

// Returns [out, ckL_0, ckR_0, ckL_1, ckR_1, ...] (checkpoints only if keep);
// out is L after the last iteration [B, N], or after every iteration
// [T, B, N] if all_iters.
std::vector<at::Tensor> dmin_forward(at::Tensor llr, at::Tensor alpha, at::Tensor beta, at::Tensor tab, int64_t T,
                                     int64_t ckpt_every, int64_t corr, bool keep, double r_max, bool all_iters) {
    const Sched s = make_sched(tab, llr.size(1));
    check_inputs(llr, alpha, beta, s, T);
    check_corr(corr);
    TORCH_CHECK(ckpt_every >= 1, "ckpt_every >= 1");
    const int64_t B = llr.size(0);
    hipStream_t st = c10::hip::getCurrentHIPStream(llr.device().index()).stream();
    at::Tensor L = llr.clone();
    at::Tensor R = at::zeros({B, s.e}, llr.options());
    at::Tensor all = all_iters ? at::empty({T, B, (int64_t)s.n}, llr.options()) : at::Tensor();
    std::vector<at::Tensor> out{all_iters ? all : L};
    const int nl = (int)s.layers.size();
    for (int64_t step = 0; step < T * nl; ++step) {
        if (keep && step % ckpt_every == 0) {
            out.push_back(L.clone());
            out.push_back(R.clone());
        }
        const int64_t g = alpha.size(0) == T ? step / nl : 0;
        const LayerArgs &la = s.layers[step % nl];
        LDPC_DISPATCH(la.deg, corr, fwd_launch, L.data_ptr<float>(), R.data_ptr<float>(), nullptr,
                      alpha.data_ptr<float>() + g * s.p, beta.data_ptr<float>() + g * s.p, la, (float)r_max, B, st);
        if (all_iters && step % nl == nl - 1) all[step / nl].copy_(L);
    }
    TORCH_CHECK(hipGetLastError() == hipSuccess, "dmin forward launch failed");
    return out;
}

// grad_out: [B, N] (adjoint of the last iteration's L) or [T, B, N] (of every
// iteration's L).  trunc > 0 truncates backpropagation through iterations:
// the adjoints of L and R are cut at every iteration boundary that is a
// multiple of trunc (the state there is treated as a constant), so each
// window of trunc iterations only sees the losses of its own iterations.
// Returns [grad_llr, grad_alpha, grad_beta].
std::vector<at::Tensor> dmin_backward(at::Tensor grad_out, std::vector<at::Tensor> ckpts, at::Tensor alpha,
                                      at::Tensor beta, at::Tensor tab, int64_t T, int64_t ckpt_every, int64_t corr,
                                      double r_max, int64_t trunc) {
    const bool all_iters = grad_out.dim() == 3;
    TORCH_CHECK(!all_iters || grad_out.size(0) == T, "per-iteration gradient must be [T, B, N]");
    const at::Tensor g_last = all_iters ? grad_out[T - 1] : grad_out;
    const Sched s = make_sched(tab, g_last.size(1));
    check_inputs(g_last.contiguous(), alpha, beta, s, T);
    check_corr(corr);
    const int nl = (int)s.layers.size();
    const int64_t n_steps = T * nl, n_seg = (n_steps + ckpt_every - 1) / ckpt_every;
    TORCH_CHECK((int64_t)ckpts.size() == 2 * n_seg, "checkpoint list does not match T and ckpt_every");
    const int64_t B = g_last.size(0);
    hipStream_t st = c10::hip::getCurrentHIPStream(grad_out.device().index()).stream();
    at::Tensor gL = g_last.contiguous().clone();
    at::Tensor gR = at::zeros({B, s.e}, grad_out.options());
    at::Tensor ga = at::zeros_like(alpha), gb = at::zeros_like(beta);
    for (int64_t seg = n_seg - 1; seg >= 0; --seg) {
        const int64_t s0 = seg * ckpt_every, s1 = std::min(n_steps, s0 + ckpt_every);
        at::Tensor L = ckpts[2 * seg].clone(), R = ckpts[2 * seg + 1].clone();
        std::vector<int64_t> off;
        int64_t tot = 0;
        for (int64_t step = s0; step < s1; ++step) {
            off.push_back(tot);
            tot += B * s.layers[step % nl].deg * s.layers[step % nl].z;
        }
        at::Tensor tape = at::empty({tot}, grad_out.options());
        for (int64_t step = s0; step < s1; ++step) {
            const int64_t g = alpha.size(0) == T ? step / nl : 0;
            const LayerArgs &la = s.layers[step % nl];
            LDPC_DISPATCH(la.deg, corr, fwd_launch, L.data_ptr<float>(), R.data_ptr<float>(),
                          tape.data_ptr<float>() + off[step - s0], alpha.data_ptr<float>() + g * s.p,
                          beta.data_ptr<float>() + g * s.p, la, (float)r_max, B, st);
        }
        for (int64_t step = s1 - 1; step >= s0; --step) {
            const int64_t g = alpha.size(0) == T ? step / nl : 0;
            const LayerArgs &la = s.layers[step % nl];
            // entering iteration t from its end: cut at a truncation boundary,
            // then add the adjoint of the iteration's output
            if (step % nl == nl - 1 && step / nl < T - 1) {
                if (trunc > 0 && (step / nl + 1) % trunc == 0) {
                    gL.zero_();
                    gR.zero_();
                }
                if (all_iters) gL.add_(grad_out[step / nl]);
            }
            const int64_t nblk = (B * la.z + TPB - 1) / TPB;
            at::Tensor partial = at::empty({nblk, 2, (int64_t)la.deg}, grad_out.options());
            LDPC_DISPATCH(la.deg, corr, bwd_launch, gL.data_ptr<float>(), gR.data_ptr<float>(),
                          tape.data_ptr<float>() + off[step - s0], alpha.data_ptr<float>() + g * s.p,
                          beta.data_ptr<float>() + g * s.p, partial.data_ptr<float>(), la, (float)r_max, B, st);
            const at::Tensor sums = partial.sum(0);
            ga[g].narrow(0, la.p_off, la.deg).add_(sums[0]);
            gb[g].narrow(0, la.p_off, la.deg).add_(sums[1]);
        }
    }
    TORCH_CHECK(hipGetLastError() == hipSuccess, "dmin backward launch failed");
    return {gL, ga, gb};
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "Differentiable layered min-sum decoder (HIP, gfx950)";
    m.def("forward", &dmin_forward, "forward: [L_out, checkpoints...]");
    m.def("backward", &dmin_backward, "backward: [grad_llr, grad_alpha, grad_beta]");
    m.attr("MAXD") = MAXD;
}
