// encoder.cpp -- systematic encoder of the CCSDS AR4JA codes (host side).
//
// The reference sends only the all-zero codeword (engines/scalar_5w/
// ldpc_decoder_core.h:258-297; tests/test_circulant_dma_stepper.cpp:744-757),
// so this encoder has no counterpart there; it serves the random-info-bit
// Monte Carlo mode and ldpc_encode.
//
// Every AR4JA parity-check matrix is quasi-cyclic with circulant size
// Z = M/4: its Z x Z blocks are sums of cyclic shifts (the permutations pi_k
// of makeH_AR4JA.m:88-91 map quarter to quarter with a shift inside).  With a
// block's bits read as the coefficients of a polynomial (bit t <-> x^t), the
// block "check j sees variable (j + s) mod Z" acts as multiplication by
// x^-s in R = GF(2)[x] / (x^Z + 1).  H = [H_u | H_p] becomes a 12 x (K/Z + 12)
// matrix over R (12 = 3M/Z block rows; the parity part is the last 3M
// variables, the punctured M among them), and the parity blocks follow from
// H_p p = H_u u as p = G u with G = H_p^-1 H_u, a 12 x (K/Z) matrix over R.
//
// Z is a power of two, so x^Z + 1 = (x + 1)^Z and R is a local ring: an
// element is a unit iff it has an odd number of terms, and then a^Z = 1
// (Frobenius: (1 + y)^Z = 1 + y^Z), so a^-1 = a^(Z-1) = a a^2 a^4 ... a^(Z/2),
// where squaring spreads coefficient i to 2i mod Z.  Gauss-Jordan elimination
// with unit pivots inverts H_p whenever it is invertible (the residue matrix
// of term-count parities is then invertible over GF(2)).
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ldpc_mi355x.h"
#include "ar4ja_schedule.h"
#include "runtime_internal.h"

namespace ldpc {
namespace {

// A polynomial of R as Z bits in 32-bit words (Z >= 32, a power of two).
struct Ring {
    int Z, W;
    explicit Ring(int z) : Z(z), W(z / 32) {}
    using Poly = std::vector<uint32_t>;
    Poly zero() const { return Poly(W, 0u); }
    Poly monomial(int e) const {
        Poly p = zero();
        e = ((e % Z) + Z) % Z;
        p[e / 32] |= 1u << (e % 32);
        return p;
    }
    static bool bit(const Poly &p, int i) { return (p[i / 32] >> (i % 32)) & 1u; }
    bool is_zero(const Poly &p) const {
        for (uint32_t w : p)
            if (w) return false;
        return true;
    }
    bool is_unit(const Poly &p) const {
        int n = 0;
        for (uint32_t w : p) n += __builtin_popcount(w);
        return n & 1;
    }
    void add_to(Poly &a, const Poly &b) const {
        for (int k = 0; k < W; ++k) a[k] ^= b[k];
    }
    // b * x^s
    Poly shifted(const Poly &b, int s) const {
        Poly out = zero();
        const int ws = s / 32, bs = s % 32;
        for (int k = 0; k < W; ++k) {
            const uint32_t lo = b[((k - ws) % W + W) % W];
            const uint32_t lo1 = b[((k - ws - 1) % W + W) % W];
            out[k] = bs ? (lo << bs) | (lo1 >> (32 - bs)) : lo;
        }
        return out;
    }
    Poly mul(const Poly &a, const Poly &b) const {
        Poly out = zero();
        for (int s = 0; s < Z; ++s)
            if (bit(a, s)) add_to(out, shifted(b, s));
        return out;
    }
    Poly square(const Poly &a) const {
        Poly out = zero();
        for (int s = 0; s < Z; ++s)
            if (bit(a, s)) {
                const int e = (2 * s) % Z;
                out[e / 32] ^= 1u << (e % 32);
            }
        return out;
    }
    Poly inverse(const Poly &a) const {  // a a unit
        Poly acc = a, p = a;
        for (int k = 1; (1 << k) < Z; ++k) {
            p = square(p);
            acc = mul(acc, p);
        }
        return acc;
    }
};

struct Generator {
    int Z = 0, nu = 0;
    std::vector<uint32_t> words;  // [12][nu][Z / 32]
};

int build_generator(int k_id, int rate_id, Generator &out, std::string &why) {
    const ar4ja::Schedule s = ar4ja::make_schedule(k_id, rate_id, 1);
    const int Z = s.Z, nb = s.N / Z, nu = s.k_info / Z, np = nb - nu;  // np = 12
    const Ring R(Z);
    // H over R: row blocks of Z checks (the 12 layers), column blocks of Z variables
    std::vector<std::vector<Ring::Poly>> H(np, std::vector<Ring::Poly>(nb, R.zero()));
    for (int li = 0; li < s.n_layers; ++li) {
        const ar4ja::Layer &L = s.layer[li];
        const int rb = L.first_row / Z;
        for (int d = 0; d < L.deg; ++d) R.add_to(H[rb][L.slot[d].base / Z], R.monomial(-L.slot[d].shift));
    }
    // Gauss-Jordan on [H_p | H_u]
    std::vector<std::vector<Ring::Poly>> A(np);
    for (int r = 0; r < np; ++r) {
        for (int c = 0; c < np; ++c) A[r].push_back(H[r][nu + c]);
        for (int c = 0; c < nu; ++c) A[r].push_back(H[r][c]);
    }
    for (int c = 0; c < np; ++c) {
        int piv = -1;
        for (int r = c; r < np && piv < 0; ++r)
            if (R.is_unit(A[r][c])) piv = r;
        if (piv < 0) {
            why = "AR4JA parity part is not invertible over GF(2)[x]/(x^Z+1)";
            return LDPC_ERR_UNSUPPORTED;
        }
        std::swap(A[c], A[piv]);
        const Ring::Poly inv = R.inverse(A[c][c]);
        for (auto &e : A[c]) e = R.mul(inv, e);
        for (int r = 0; r < np; ++r) {
            if (r == c || R.is_zero(A[r][c])) continue;
            const Ring::Poly f = A[r][c];
            for (int k = 0; k < np + nu; ++k) R.add_to(A[r][k], R.mul(f, A[c][k]));
        }
    }
    out.Z = Z;
    out.nu = nu;
    out.words.assign((size_t)np * nu * (Z / 32), 0u);
    for (int r = 0; r < np; ++r)
        for (int i = 0; i < nu; ++i) {
            const Ring::Poly &g = A[r][np + i];
            std::memcpy(&out.words[((size_t)r * nu + i) * (Z / 32)], g.data(), sizeof(uint32_t) * (size_t)(Z / 32));
        }
    return LDPC_OK;
}

std::mutex g_mu;
std::map<std::pair<int, int>, Generator> g_cache;

}  // namespace

int qc_generator(int k_id, int rate_id, const uint32_t **words, int *z, int *nu, std::string &why) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find({k_id, rate_id});
    if (it == g_cache.end()) {
        Generator g;
        const int rc = build_generator(k_id, rate_id, g, why);
        if (rc) return rc;
        it = g_cache.emplace(std::make_pair(k_id, rate_id), std::move(g)).first;
    }
    *words = it->second.words.data();
    *z = it->second.Z;
    *nu = it->second.nu;
    return LDPC_OK;
}

}  // namespace ldpc

extern "C" {

int ldpc_get_generator(const ldpc_code_desc *code, uint32_t *out, int32_t cap) {
    if (!code || cap < 0 || (!out && cap > 0)) return LDPC_ERR_ARG;
    int k_id, rate_id;
    if (!ldpc::ar4ja_ids(*code, k_id, rate_id)) return LDPC_ERR_UNSUPPORTED;
    const uint32_t *w;
    int Z, nu;
    std::string why;
    const int rc = ldpc::qc_generator(k_id, rate_id, &w, &Z, &nu, why);
    if (rc) return rc;
    const int n = 12 * nu * (Z / 32);
    if (cap > 0) std::memcpy(out, w, sizeof(uint32_t) * (size_t)(cap < n ? cap : n));
    return n;
}

int ldpc_encode(const ldpc_code_desc *code, const uint8_t *info, uint32_t n_cw, uint8_t *codeword) {
    if (!code || (n_cw && (!info || !codeword))) return LDPC_ERR_ARG;
    int k_id, rate_id;
    if (!ldpc::ar4ja_ids(*code, k_id, rate_id)) return LDPC_ERR_UNSUPPORTED;
    const uint32_t *w;
    int Z, nu;
    std::string why;
    const int rc = ldpc::qc_generator(k_id, rate_id, &w, &Z, &nu, why);
    if (rc) return rc;
    const int K = nu * Z, N = (nu + 12) * Z;
    std::vector<uint8_t> g((size_t)12 * nu * Z);  // coefficients of every g_{r,i}
    for (size_t k = 0; k < g.size(); ++k) g[k] = (w[k / 32] >> (k % 32)) & 1u;
    for (size_t k = 0; k < (size_t)n_cw * K; ++k)
        if (info[k] > 1) return LDPC_ERR_ARG;
    for (uint32_t c = 0; c < n_cw; ++c) {
        const uint8_t *u = info + (size_t)c * K;
        uint8_t *x = codeword + (size_t)c * N;
        std::memcpy(x, u, (size_t)K);
        // parity block r, bit t: sum over i, s of g_{r,i}[s] u_i[(t - s) mod Z]
        for (int r = 0; r < 12; ++r)
            for (int t = 0; t < Z; ++t) {
                uint8_t p = 0;
                for (int i = 0; i < nu; ++i) {
                    const uint8_t *gi = &g[((size_t)r * nu + i) * Z];
                    const uint8_t *ui = u + (size_t)i * Z;
                    for (int s = 0; s < Z; ++s) p ^= gi[s] & ui[(t - s) & (Z - 1)];
                }
                x[K + r * Z + t] = p;
            }
    }
    return LDPC_OK;
}

}  // extern "C"
