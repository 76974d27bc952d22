// runtime.cpp -- the C ABI of libldpc (include/ldpc_mi355x.h): contexts,
// device buffers, code set-up for both decoder profiles, the chunked Monte
// Carlo driver and its stopping rule.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <queue>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ldpc_mi355x.h"
#include "ar4ja_schedule.h"
#include "kernels.h"
#include "runtime_internal.h"

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
    // Profile T on the generic nine-code kernels (null: the specialised
    // AR4JA r1/2 K=1024 kernel), or the Profile D kernel
    const ldpc::MultiVariant *mv = nullptr;
    // trapping-set capture (ldpc_set_capture)
    uint32_t *d_cap_count = nullptr;
    unsigned long long *d_cap_ids = nullptr;
    void *d_work = nullptr;  // per-workgroup message workspace (Profile D, K=16384 r1/2)
    uint32_t cap_max = 0;
    // Profile D parameters on the device
    float *d_alpha = nullptr, *d_beta = nullptr;
    int d_groups = 0, d_corr = 0;
    float d_rmax = 0.0f;
    // random-info-bit codewords (ldpc_set_codeword_mode): generator words of
    // the transmitted parity blocks on the device; null = the all-zero codeword
    uint32_t *d_cw_gen = nullptr;
    bool cw_random = false;
    // Profile R: pass table on the device (and the pass kinds of the
    // two-codeword kernel)
    void *d_table = nullptr;
    uint32_t *d_kinds = nullptr;
    uint16_t *d_slot = nullptr;
    ldpc::RLaunch rl{};
    DevBuf in, out, iters, ok, tails;
    int cus = 0;
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

// Parity-check rows of an AR4JA code in the natural (makeH) row order, each
// row's variables ascending.
void ar4ja_rows(int k_id, int rate_id, std::vector<int32_t> &rp, std::vector<int32_t> &col, uint32_t &n,
                uint32_t &m, uint32_t &np) {
    const ar4ja::Schedule s = ar4ja::make_schedule(k_id, rate_id);
    n = (uint32_t)s.N;
    m = (uint32_t)s.n_checks;
    np = (uint32_t)s.n_punct;
    std::vector<const ar4ja::Layer *> by_first(s.n_checks / s.Z);
    for (int li = 0; li < s.n_layers; ++li) by_first[s.layer[li].first_row / s.Z] = &s.layer[li];
    rp.assign(1, 0);
    col.clear();
    for (int r = 0; r < s.n_checks; ++r) {
        const ar4ja::Layer &L = *by_first[r / s.Z];
        const int j = r % s.Z;
        for (int d = 0; d < L.deg; ++d) col.push_back(L.slot[d].base + (L.slot[d].shift + j) % s.Z);
        rp.push_back((int32_t)col.size());
    }
}

int levelize(uint32_t n, uint32_t m, const int32_t *rp, const int32_t *col, int32_t *level) {
    std::vector<int32_t> last(n, 0);
    int n_levels = 0;
    for (uint32_t i = 0; i < m; ++i) {
        int l = 0;
        for (int32_t e = rp[i]; e < rp[i + 1]; ++e) l = std::max(l, last[col[e]]);
        level[i] = l + 1;
        for (int32_t e = rp[i]; e < rp[i + 1]; ++e) last[col[e]] = l + 1;
        n_levels = std::max(n_levels, l + 1);
    }
    return n_levels;
}

// Rows in passes of at most `width`, for the Profile R kernel.  A row is
// ready once every earlier row (file order) sharing a variable with it sits
// in an earlier pass, so the rows of one pass touch disjoint variables and
// each reads exactly the LLRs the sequential order reads.  Among ready rows
// the one with the longest chain of dependent rows after it goes first (then
// file order).  For the r1/2 K=1024 matrix, passes of 64 take 37 steps, the
// length of the chain (cutting its 37 ASAP levels into passes of 64 takes
// 38); passes of 32 would take 48, all full.  Writes the pass of every row
// (0-based) and returns the number of passes.
int schedule_passes(uint32_t n, uint32_t m, const int32_t *rp, const int32_t *col, int width, int32_t *pass) {
    std::vector<int32_t> last(n, -1);
    std::vector<std::vector<uint32_t>> succ(m);
    std::vector<int32_t> indeg(m, 0);
    for (uint32_t i = 0; i < m; ++i)
        for (int32_t e = rp[i]; e < rp[i + 1]; ++e) {
            const int32_t p = last[col[e]];
            if (p >= 0 && p != (int32_t)i) {  // a row listing a variable twice has no edge to itself
                succ[p].push_back(i);
                ++indeg[i];
            }
            last[col[e]] = (int32_t)i;
        }
    std::vector<int32_t> height(m);
    for (uint32_t k = m; k-- > 0;) {
        int32_t x = 0;
        for (uint32_t s : succ[k]) x = std::max(x, height[s]);
        height[k] = x + 1;
    }
    auto later = [&](uint32_t a, uint32_t b) { return height[a] != height[b] ? height[a] < height[b] : a > b; };
    std::priority_queue<uint32_t, std::vector<uint32_t>, decltype(later)> ready(later);
    for (uint32_t i = 0; i < m; ++i)
        if (!indeg[i]) ready.push(i);
    int n_pass = 0;
    std::vector<uint32_t> cur;
    while (!ready.empty()) {
        cur.clear();
        while (!ready.empty() && (int)cur.size() < width) {
            cur.push_back(ready.top());
            ready.pop();
        }
        for (uint32_t r : cur) {
            pass[r] = n_pass;
            for (uint32_t s : succ[r])
                if (--indeg[s] == 0) ready.push(s);
        }
        ++n_pass;
    }
    for (uint32_t i = 0; i < m; ++i)
        if (indeg[i]) return -1;  // unreachable for a dependency graph built in file order
    return n_pass;
}

// Rows in degree-homogeneous passes of at most `width`, for the two-codeword
// Profile R kernel.  Same readiness rule as schedule_passes.  When every row
// has degree 3 or 6, each pass takes ready rows of one degree only (kind
// R2_PK_EXACT3 / R2_PK_EXACT6), and a degree-3 pass costs about 0.42 of a
// degree-6 one (its datapath's instruction count).  The degree of each pass
// is chosen by rollout: for each degree with ready rows, the rest of the
// schedule is completed greedily (each pass takes the degree of the ready
// row with the longest chain after it) and the cheaper total wins; for the
// r1/2 K=1024 matrix with passes of 32 this gives 32 degree-6 passes (the
// minimum) and 18 degree-3 ones.  Other matrices get plain list-scheduled
// passes of kind R2_PK_MASKED.  Writes the pass of every row and the kind of
// every pass; returns the number of passes (-1 on a scheduling failure).
int schedule_passes_by_degree(uint32_t n, uint32_t m, const int32_t *rp, const int32_t *col, int width,
                              int32_t *pass, std::vector<uint8_t> &kind) {
    bool exact = true;
    for (uint32_t i = 0; i < m; ++i) {
        const int d = rp[i + 1] - rp[i];
        exact = exact && (d == 3 || d == 6);
    }
    if (!exact) {
        const int np = schedule_passes(n, m, rp, col, width, pass);
        kind.assign(np > 0 ? np : 0, (uint8_t)ldpc::R2_PK_MASKED);
        return np;
    }
    std::vector<int32_t> last(n, -1);
    std::vector<std::vector<uint32_t>> succ(m);
    std::vector<int32_t> indeg0(m, 0);
    for (uint32_t i = 0; i < m; ++i)
        for (int32_t e = rp[i]; e < rp[i + 1]; ++e) {
            const int32_t p = last[col[e]];
            if (p >= 0 && p != (int32_t)i) {
                succ[p].push_back(i);
                ++indeg0[i];
            }
            last[col[e]] = (int32_t)i;
        }
    std::vector<int32_t> height(m);
    for (uint32_t k = m; k-- > 0;) {
        int32_t x = 0;
        for (uint32_t s : succ[k]) x = std::max(x, height[s]);
        height[k] = x + 1;
    }
    auto later = [&](uint32_t a, uint32_t b) { return height[a] != height[b] ? height[a] < height[b] : a > b; };
    using Heap = std::priority_queue<uint32_t, std::vector<uint32_t>, decltype(later)>;
    auto cls_of = [&](uint32_t r) { return rp[r + 1] - rp[r] == 3 ? 0 : 1; };
    const double cost[2] = {0.42, 1.0};
    struct State {
        std::vector<int32_t> indeg;
        std::vector<Heap> ready;
        uint32_t left;
    };
    State st{indeg0, {Heap(later), Heap(later)}, m};
    for (uint32_t i = 0; i < m; ++i)
        if (!indeg0[i]) st.ready[cls_of(i)].push(i);
    // one pass of class c from state s; rows go to `out` when given
    auto step = [&](State &s, int c, std::vector<uint32_t> *out) {
        std::vector<uint32_t> cur;
        while (!s.ready[c].empty() && (int)cur.size() < width) {
            cur.push_back(s.ready[c].top());
            s.ready[c].pop();
        }
        for (uint32_t r : cur)
            for (uint32_t x : succ[r])
                if (--s.indeg[x] == 0) s.ready[cls_of(x)].push(x);
        s.left -= (uint32_t)cur.size();
        if (out) *out = cur;
    };
    // class of the ready row with the longest chain after it (-1: none ready)
    auto top_cls = [&](const State &s) {
        if (s.ready[0].empty()) return s.ready[1].empty() ? -1 : 1;
        if (s.ready[1].empty()) return 0;
        return later(s.ready[0].top(), s.ready[1].top()) ? 1 : 0;
    };
    auto greedy_cost = [&](State s) {
        double c = 0.0;
        for (int k; s.left && (k = top_cls(s)) >= 0;) {
            step(s, k, nullptr);
            c += cost[k];
        }
        return s.left ? 1e30 : c;
    };
    // the rollout costs O(passes^2 * m); larger matrices (which the
    // two-codeword kernel cannot hold anyway) take the greedy choice
    const bool rollout = m <= (uint32_t)(ldpc::R2_MAX_PASSES * width);
    kind.clear();
    std::vector<uint32_t> rows;
    while (st.left) {
        int best = top_cls(st);
        if (best < 0) return -1;
        if (rollout && !st.ready[0].empty() && !st.ready[1].empty()) {
            double bc = 1e31;
            for (int k : {best, 1 - best}) {
                State t = st;
                step(t, k, nullptr);
                const double c = cost[k] + greedy_cost(t);
                if (c < bc) {
                    bc = c;
                    best = k;
                }
            }
        }
        step(st, best, &rows);
        for (uint32_t r : rows) pass[r] = (int32_t)kind.size();
        kind.push_back(best ? (uint8_t)ldpc::R2_PK_EXACT6 : (uint8_t)ldpc::R2_PK_EXACT3);
    }
    return (int)kind.size();
}

// LDS slot of every variable for the two-codeword Profile R kernel.  Each
// (pass, edge position) is one gather: a ds_read_u16 whose lanes read one
// variable each (slot 0 on idle lanes and past a row's degree).  The
// LDS serves a half-wave's reads in one cycle when they fall in distinct
// 4-byte banks (bank = (slot / 2) mod 32) and in k cycles when k share one.
// The kernel reaches the variables only through the pass table, the channel
// generator and the LLR copies, which all go through this map, so the slots
// are free: a local search swaps the slots of two variables while that lowers
// the sum over gathers of the squared reads per bank (simulated annealing,
// fixed seed, so the map is deterministic).  Transmitted variables keep slots
// below n_tx and punctured ones slots in [n_tx, n): the tally of the sign bits
// of the transmitted variables still reads slots 0 .. n_tx - 1.
std::vector<uint16_t> place_for_banks(uint32_t n, uint32_t n_tx, uint32_t m, const int32_t *rp, const int32_t *col,
                                      const int32_t *pass, int np) {
    constexpr uint32_t B = 32;
    std::vector<std::vector<uint32_t>> rows_of(np);
    for (uint32_t r = 0; r < m; ++r) rows_of[pass[r]].push_back(r);
    std::vector<std::vector<uint32_t>> gathers;
    for (int p = 0; p < np; ++p)
        for (int d = 0; d < 6; ++d) {
            std::vector<uint32_t> g;
            bool zero = rows_of[p].size() < (size_t)ldpc::R2_PASS_ROWS;
            for (uint32_t r : rows_of[p]) {
                if (rp[r + 1] - rp[r] > d)
                    g.push_back((uint32_t)col[rp[r] + d]);
                else
                    zero = true;
            }
            if (g.empty()) continue;
            if (zero) g.push_back(n);
            gathers.push_back(std::move(g));
        }
    std::vector<uint32_t> slot(n + 1);
    for (uint32_t v = 0; v < n; ++v) slot[v] = v;
    slot[n] = 0;  // the reads of idle lanes and of slots past a row's degree
    std::vector<std::vector<uint32_t>> gof(n + 1);
    for (uint32_t g = 0; g < gathers.size(); ++g)
        for (uint32_t v : gathers[g]) gof[v].push_back(g);
    std::vector<int32_t> load(gathers.size() * B, 0);
    auto bank = [](uint32_t s) { return (s >> 1) & (B - 1); };
    for (uint32_t g = 0; g < gathers.size(); ++g)
        for (uint32_t v : gathers[g]) ++load[g * B + bank(slot[v])];
    uint64_t x = 0x9E3779B97F4A7C15ull;
    auto rnd = [&x]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return x;
    };
    // move v's read from bank from to bank to in each of its gathers; returns the change of the objective
    auto shift = [&](uint32_t v, uint32_t from, uint32_t to) {
        int32_t d = 0;
        for (uint32_t g : gof[v]) {
            int32_t &lf = load[g * B + from], &lt = load[g * B + to];
            d += 2 * (lt - lf) + 2;
            --lf;
            ++lt;
        }
        return d;
    };
    const uint64_t iters = std::min<uint64_t>(4000000, 400 * (uint64_t)n);
    for (uint64_t it = 0; it < iters; ++it) {
        const bool tx = (rnd() % n) < n_tx || n_tx == n;
        const uint32_t lo = tx ? 0 : n_tx, span = tx ? n_tx : n - n_tx;
        if (span < 2) continue;
        const uint32_t a = lo + (uint32_t)(rnd() % span), b = lo + (uint32_t)(rnd() % span);
        const uint32_t ba = bank(slot[a]), bb = bank(slot[b]);
        if (ba == bb) continue;
        const int32_t d = shift(a, ba, bb) + shift(b, bb, ba);
        const double t = 2.0 * (1.0 - (double)it / (0.7 * (double)iters));
        const bool take = d <= 0 || (t > 0.0 && (double)(rnd() >> 11) * 0x1.0p-53 < std::exp(-(double)d / t));
        if (take) {
            std::swap(slot[a], slot[b]);
        } else {
            shift(b, ba, bb);
            shift(a, bb, ba);
        }
    }
    return std::vector<uint16_t>(slot.begin(), slot.end() - 1);
}

int validate_rows(uint32_t n, uint32_t m, const int32_t *rp, const int32_t *col, std::string &why) {
    if (!rp || !col || n == 0 || m == 0) {
        why = "matrix: empty or null arrays";
        return LDPC_ERR_ARG;
    }
    if (rp[0] != 0) {
        why = "matrix: row_ptr[0] must be 0";
        return LDPC_ERR_ARG;
    }
    for (uint32_t i = 0; i < m; ++i) {
        const int d = rp[i + 1] - rp[i];
        if (d < 1 || d > 12) {
            why = "matrix: check degree must be 1..12 (ldpc_sim_scalar.cpp:34)";
            return LDPC_ERR_ARG;
        }
        for (int32_t e = rp[i]; e < rp[i + 1]; ++e) {
            if (col[e] < 0 || (uint32_t)col[e] >= n) {
                why = "matrix: variable index out of range";
                return LDPC_ERR_ARG;
            }
            // A repeated (row, column) entry has no single meaning (over GF(2) it
            // cancels; as two edges it double-counts a message), so it is refused.
            for (int32_t f = rp[i]; f < e; ++f)
                if (col[f] == col[e]) {
                    why = "matrix: row " + std::to_string(i + 1) + " lists variable " + std::to_string(col[e] + 1) +
                          " twice";
                    return LDPC_ERR_ARG;
                }
        }
    }
    return LDPC_OK;
}

// Profile R set-up: rows -> passes of <= R_PASS_ROWS rows -> device table.
int setup_profile_r(ldpc_ctx *c, const std::vector<int32_t> &rp, const std::vector<int32_t> &col, uint32_t n,
                    uint32_t m, uint32_t np, uint32_t k_info) {
    std::string why;
    int rc = validate_rows(n, m, rp.data(), col.data(), why);
    if (rc) return fail(nullptr, rc, why);
    if (n > 65535) return fail(nullptr, LDPC_ERR_UNSUPPORTED, "Profile R: more than 65535 variables");
    const uint32_t n_edges = (uint32_t)rp[m];
    int maxd = 0;
    for (uint32_t i = 0; i < m; ++i) maxd = std::max(maxd, rp[i + 1] - rp[i]);
    std::vector<int32_t> level(m);
    const int n_levels = levelize(n, m, rp.data(), col.data(), level.data());
    // The two-codeword kernel (decode_r2.hip) where the code fits it: degree
    // <= 6, at most R2_MAX_PASSES passes of 32 rows, both codewords' LLRs in
    // 64 KiB of LDS.  LDPC_R_ONE_CW=1 in the environment keeps the
    // one-codeword kernel (for comparisons).
    const char *one_cw = std::getenv("LDPC_R_ONE_CW");
    // bytes of LDS per half: the LLRs (2560 variables: 5 KiB, 16 workgroups
    // of one wave per CU)
    const uint32_t hs = (2 * n + 15) & ~15u;
    if (maxd <= 6 && 2 * hs <= 65536 && !(one_cw && one_cw[0] == '1')) {
        std::vector<int32_t> pass(m);
        std::vector<uint8_t> kind;
        const int np2 = schedule_passes_by_degree(n, m, rp.data(), col.data(), ldpc::R2_PASS_ROWS, pass.data(), kind);
        if (np2 > 0 && np2 <= ldpc::R2_MAX_PASSES) {
            const int n_groups = (np2 + 3) / 4;
            const size_t n_entries = (size_t)(4 * n_groups + 1) * 64;
            // entry (p, lane): lane k < 32 runs the k-th row of pass p on the
            // first codeword, lane k + 32 on the second (offsets + hs).  Word 0:
            // degree (0 on idle lanes); words 1..3: LDS byte offsets of the
            // row's variables, 16 bits each, the half's slot 0 past the degree
            std::vector<uint32_t> table(n_entries * 4, 0u);
            const std::vector<uint16_t> slot = place_for_banks(n, n - np, m, rp.data(), col.data(), pass.data(), np2);
            const uint32_t zero = 0;
            for (size_t e = 0; e < n_entries; ++e) {
                const uint32_t z = zero + (e % 64 >= 32 ? hs : 0);
                for (int w = 1; w < 4; ++w) table[e * 4 + w] = z | (z << 16);
            }
            std::vector<int> fill(np2, 0);
            for (uint32_t r = 0; r < m; ++r) {
                const int k = fill[pass[r]]++;
                const uint32_t deg = (uint32_t)(rp[r + 1] - rp[r]);
                for (int h = 0; h < 2; ++h) {
                    uint32_t *e = &table[((size_t)pass[r] * 64 + k + 32 * h) * 4];
                    e[0] = deg;
                    for (uint32_t d = 0; d < deg; ++d) {
                        const uint32_t o = 2 * (uint32_t)slot[col[rp[r] + d]] + (h ? hs : 0);
                        uint32_t &wd = e[1 + d / 2];
                        wd = (d & 1) ? ((wd & 0xFFFFu) | (o << 16)) : ((wd & 0xFFFF0000u) | o);
                    }
                }
            }
            std::vector<uint32_t> kinds(n_groups, 0u);
            for (int p = 0; p < np2; ++p) kinds[p / 4] |= (uint32_t)kind[p] << (8 * (p % 4));
            HIP_TRY(c, hipMalloc(&c->d_table, table.size() * sizeof(uint32_t)));
            HIP_TRY(c, hipMemcpy(c->d_table, table.data(), table.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
            HIP_TRY(c, hipMalloc(&c->d_kinds, kinds.size() * sizeof(uint32_t)));
            HIP_TRY(c, hipMemcpy(c->d_kinds, kinds.data(), kinds.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
            // slot map, padded to a multiple of 4 entries (the generator reads 4 at a time)
            std::vector<uint16_t> slot4((n + 3) & ~3u, 0);
            std::copy(slot.begin(), slot.end(), slot4.begin());
            HIP_TRY(c, hipMalloc(&c->d_slot, slot4.size() * sizeof(uint16_t)));
            HIP_TRY(c, hipMemcpy(c->d_slot, slot4.data(), slot4.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
            c->rl.r2_mode = kind[0] == ldpc::R2_PK_MASKED ? 1 : 0;
            c->rl.table2 = c->d_table;
            c->rl.kinds2 = c->d_kinds;
            c->rl.slot2 = c->d_slot;
            c->rl.n_groups = n_groups;
            c->rl.hs = (int)hs;
            c->rl.n_pass = np2;
            c->rl.n_edges = (int)n_edges;
        }
    }
    if (c->rl.r2_mode < 0) {
        const int kmaxd = maxd <= 6 ? 6 : 12;  // the kernel's slot count
        // Message layout in LDS (private to the kernel, zeroed on entry): rows in
        // file order, except that the last row of the largest degree is moved to
        // the end, plus kmaxd - maxd spare slots.  Then es + kmaxd - 1 is a valid
        // slot for every row, and the kernel reads all kmaxd slots unmasked.
        const uint32_t r_slots = n_edges + (uint32_t)(kmaxd - maxd);
        if (n_edges >= (1u << 28) || ldpc::r_lds_bytes((int)n, (int)r_slots) > 65536)
            return fail(nullptr, LDPC_ERR_UNSUPPORTED,
                        "Profile R keeps a codeword's LLRs and messages in LDS: needs n + edges <= 16384");
        std::vector<uint32_t> r_off(m);
        {
            uint32_t last = 0;
            for (uint32_t i = 0; i < m; ++i)
                if (rp[i + 1] - rp[i] == maxd) last = i;
            uint32_t off = 0;
            for (uint32_t i = 0; i < m; ++i)
                if (i != last) {
                    r_off[i] = off;
                    off += (uint32_t)(rp[i + 1] - rp[i]);
                }
            r_off[last] = off;
        }
        std::vector<int32_t> pass(m);
        const int n_pass = schedule_passes(n, m, rp.data(), col.data(), ldpc::R_PASS_ROWS, pass.data());
        if (n_pass <= 0) return fail(nullptr, LDPC_ERR_ARG, "Profile R: rows could not be scheduled");
        const int words = kmaxd == 6 ? 4 : 8;  // uint32 per row entry
        // entry (p, k): the k-th row of pass p (file order inside a pass); unused
        // entries stay zero (degree 0)
        std::vector<uint32_t> table((size_t)n_pass * ldpc::R_PASS_ROWS * words, 0u);
        std::vector<int> fill(n_pass, 0);
        for (uint32_t r = 0; r < m; ++r) {
            uint32_t *e = &table[((size_t)pass[r] * ldpc::R_PASS_ROWS + fill[pass[r]]++) * words];
            const uint32_t deg = (uint32_t)(rp[r + 1] - rp[r]);
            e[0] = deg | (r_off[r] << 4);
            // variables 0..5 in words 1..3, 6..11 in words 4..6
            for (uint32_t d = 0; d < deg; ++d)
                e[1 + d / 2] |= ((uint32_t)col[rp[r] + d] & 0xFFFFu) << (16 * (d & 1));
        }
        HIP_TRY(c, hipMalloc(&c->d_table, table.size() * sizeof(uint32_t)));
        HIP_TRY(c, hipMemcpy(c->d_table, table.data(), table.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        c->rl.table = c->d_table;
        c->rl.n_pass = n_pass;
        c->rl.n_edges = (int)r_slots;
    }
    c->rl.n = (int)n;
    c->rl.n_tx = (int)(n - np);
    c->rl.max_iter = (int)c->cfg.max_iter;
    c->rl.r_max = c->cfg.r_max;
    c->rl.l_post_max = c->cfg.l_post_max;
    c->rl.l_max = c->cfg.l_max;
    c->rl.maxd = maxd;
    c->rl.nms = c->cfg.profile == LDPC_PROFILE_R_NMS;
    c->info.n = n;
    c->info.m = m;
    c->info.k_info = k_info;
    c->info.n_tx = n - np;
    c->info.z = 0;
    c->info.n_layers = (uint32_t)n_levels;
    c->info.n_edges = n_edges;
    return LDPC_OK;
}

}  // namespace

namespace ldpc {

bool ar4ja_ids(const ldpc_code_desc &d, int &k_id, int &rate_id) {
    k_id = d.family != LDPC_CODE_AR4JA ? -1
           : d.k_info == 1024          ? 0
           : d.k_info == 4096          ? 1
           : d.k_info == 16384         ? 2
                                       : -1;
    rate_id = (d.rate_num == 1 && d.rate_den == 2)   ? 0
              : (d.rate_num == 2 && d.rate_den == 3) ? 1
              : (d.rate_num == 4 && d.rate_den == 5) ? 2
                                                     : -1;
    return k_id >= 0 && rate_id >= 0;
}

const MultiVariant *multi_table(int k_id, int rate_id, int *n) {
    static const MultiVariant *(*const T[3][3])(int *) = {
        {multi_table_k0_r0, multi_table_k0_r1, multi_table_k0_r2},
        {multi_table_k1_r0, multi_table_k1_r1, multi_table_k1_r2},
        {multi_table_k2_r0, multi_table_k2_r1, multi_table_k2_r2}};
    return T[k_id][rate_id](n);
}

uint32_t variant_code(const MultiVariant &v) { return LDPC_VARIANT(v.c, v.order); }

const MultiVariant *find_variant(int k_id, int rate_id, uint32_t variant) {
    if (!(variant & 0x100u) || (variant & ~0x1FFu)) return nullptr;
    int n = 0;
    const MultiVariant *t = multi_table(k_id, rate_id, &n);
    for (int i = 0; i < n; ++i)
        if (variant_code(t[i]) == variant) return &t[i];
    return nullptr;
}

// Default variant per code (layer order 0): the checks per thread that
// ldpc_autotune found fastest on an MI355X in each code's waterfall
// (K = 1024 / 4096 / 16384 rows; rates 1/2, 2/3, 4/5 columns).  One check per
// thread wins wherever it exists except K = 4096 r1/2; K = 16384 r1/2
// (z = 2048) has two-check threads only.
const MultiVariant *profile_d_kernel(int k_id, int rate_id) {
    static const MultiVariant *(*const T[3][3])() = {
        {multi_table_d_k0_r0, multi_table_d_k0_r1, multi_table_d_k0_r2},
        {multi_table_d_k1_r0, multi_table_d_k1_r1, multi_table_d_k1_r2},
        {multi_table_d_k2_r0, multi_table_d_k2_r1, multi_table_d_k2_r2}};
    return T[k_id][rate_id]();
}

uint32_t default_variant(int k_id, int rate_id) {
    static const int C[3][3] = {{1, 1, 1}, {2, 1, 1}, {2, 1, 1}};
    return LDPC_VARIANT(C[k_id][rate_id], 0);
}

void fill_code_info(int k_id, int rate_id, ldpc_code_info &info) {
    const ar4ja::Schedule s = ar4ja::make_schedule(k_id, rate_id);
    info.n = (uint32_t)s.N;
    info.m = (uint32_t)s.n_checks;
    info.k_info = (uint32_t)s.k_info;
    info.n_tx = (uint32_t)(s.N - s.n_punct);
    info.z = (uint32_t)s.Z;
    info.n_layers = (uint32_t)s.n_layers;
    info.n_edges = (uint32_t)s.n_edges;
}

void channel_params(const ldpc_code_info &info, double ebn0_db, float *mu, float *sigma) {
    const double rate = (double)info.k_info / (double)info.n_tx;
    const double g = std::pow(10.0, ebn0_db / 10.0);
    *mu = (float)(4.0 * rate * g);
    *sigma = (float)std::sqrt(8.0 * rate * g);
}

void set_context_free_error(const std::string &msg) { g_create_error = msg; }

}  // namespace ldpc

extern "C" {

const char *ldpc_last_error(const ldpc_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int ldpc_row_levels(uint32_t n, uint32_t m, const int32_t *row_ptr, const int32_t *col, int32_t *level) {
    std::string why;
    if (!level || validate_rows(n, m, row_ptr, col, why)) return LDPC_ERR_ARG;
    return levelize(n, m, row_ptr, col, level);
}

int ldpc_row_passes(uint32_t n, uint32_t m, const int32_t *row_ptr, const int32_t *col, int32_t width,
                    int32_t *pass) {
    std::string why;
    if (!pass || width < 1 || validate_rows(n, m, row_ptr, col, why)) return LDPC_ERR_ARG;
    return schedule_passes(n, m, row_ptr, col, width, pass);
}

int ldpc_row_passes_by_degree(uint32_t n, uint32_t m, const int32_t *row_ptr, const int32_t *col, int32_t width,
                              int32_t *pass, uint8_t *kind, int32_t kind_cap) {
    std::string why;
    if (!pass || width < 1 || validate_rows(n, m, row_ptr, col, why)) return LDPC_ERR_ARG;
    std::vector<uint8_t> k;
    const int np = schedule_passes_by_degree(n, m, row_ptr, col, width, pass, k);
    if (np < 0) return LDPC_ERR_ARG;
    if (kind) {
        if (kind_cap < np) return LDPC_ERR_ARG;
        std::copy(k.begin(), k.end(), kind);
    }
    return np;
}

int ldpc_create(const ldpc_code_desc *code, const ldpc_decoder_cfg *cfg, int device, ldpc_ctx **out) {
    if (!code || !cfg || !out) return fail(nullptr, LDPC_ERR_ARG, "null argument");
    *out = nullptr;
    if (code->family != LDPC_CODE_AR4JA && code->family != LDPC_CODE_MATRIX)
        return fail(nullptr, LDPC_ERR_ARG, "unknown code family");
    int k_id = -1, rate_id = -1;
    if (code->family == LDPC_CODE_AR4JA && !ldpc::ar4ja_ids(*code, k_id, rate_id))
        return fail(nullptr, LDPC_ERR_ARG, "no AR4JA code with this (k_info, rate)");
    if (cfg->profile > LDPC_PROFILE_D) return fail(nullptr, LDPC_ERR_UNSUPPORTED, "unknown decoder profile");
    if (cfg->max_iter < 1 || cfg->max_iter > 100000) return fail(nullptr, LDPC_ERR_ARG, "max_iter out of range");
    const bool prof_t = cfg->profile == LDPC_PROFILE_T;
    const bool prof_d = cfg->profile == LDPC_PROFILE_D;
    if (prof_d) {
        if (code->family != LDPC_CODE_AR4JA)
            return fail(nullptr, LDPC_ERR_UNSUPPORTED, "Profile D kernels are built for the AR4JA codes");
        if (cfg->variant || cfg->l_max != 0.0f || cfg->r_max != 0.0f || cfg->l_post_max != 0.0f || cfg->gamma != 0.0f)
            return fail(nullptr, LDPC_ERR_ARG, "Profile D takes its settings from ldpc_set_dmin_params");
    } else if (prof_t) {
        if (!(cfg->beta > 0.0f) || !std::isfinite(cfg->beta)) return fail(nullptr, LDPC_ERR_ARG, "beta must be > 0");
        if (cfg->l_max != 0.0f || cfg->r_max != 0.0f || cfg->l_post_max != 0.0f || cfg->gamma != 0.0f)
            return fail(nullptr, LDPC_ERR_ARG, "clamps belong to Profile R; the golden Profile T has none");
        if (code->family != LDPC_CODE_AR4JA)
            return fail(nullptr, LDPC_ERR_UNSUPPORTED, "Profile T kernels are built for the AR4JA codes");
        if (cfg->variant && !ldpc::find_variant(k_id, rate_id, cfg->variant))
            return fail(nullptr, LDPC_ERR_UNSUPPORTED, "no such Profile T kernel variant for this code");
    } else {
        if (cfg->variant) return fail(nullptr, LDPC_ERR_ARG, "kernel variants belong to Profile T");
        for (float v : {cfg->l_max, cfg->r_max, cfg->l_post_max, cfg->gamma})
            if (!(v >= 0.0f) || !std::isfinite(v)) return fail(nullptr, LDPC_ERR_ARG, "clamps must be >= 0");
        if (code->family == LDPC_CODE_MATRIX &&
            (!code->row_ptr || !code->col || code->m == 0 || code->n == 0 || code->n_punct >= code->n))
            return fail(nullptr, LDPC_ERR_ARG, "matrix: bad dimensions or null arrays");
    }

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(nullptr, LDPC_ERR_NODEV, "no HIP device");
    if (device < 0 || device >= ndev) return fail(nullptr, LDPC_ERR_ARG, "device index out of range");

    ldpc_ctx *c = new ldpc_ctx();
    c->device = device;
    c->code = *code;
    c->code.row_ptr = nullptr;  // not kept: the matrix is copied into the pass table
    c->code.col = nullptr;
    c->cfg = *cfg;
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
    int cus = 0;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
        return bail(e, "hipDeviceGetAttribute");
    c->cus = cus;

    if (prof_d) {
        c->mv = ldpc::profile_d_kernel(k_id, rate_id);
        ldpc::fill_code_info(k_id, rate_id, c->info);
        int occ = 0;
        if ((e = c->mv->occupancy(&occ)) != hipSuccess) return bail(e, "occupancy");
        c->grid_mc = c->grid_dec = cus * (occ > 0 ? occ : 1);
        if (c->mv->work_per_wg &&
            (e = hipMalloc(&c->d_work, (size_t)c->grid_mc * c->mv->work_per_wg)) != hipSuccess)
            return bail(e, "hipMalloc");
    } else if (prof_t && !(is_r12_k1024(*code) && cfg->variant == 0)) {
        c->mv = ldpc::find_variant(k_id, rate_id, cfg->variant ? cfg->variant : ldpc::default_variant(k_id, rate_id));
        ldpc::fill_code_info(k_id, rate_id, c->info);
        int occ = 0;
        if ((e = c->mv->occupancy(&occ)) != hipSuccess) return bail(e, "occupancy");
        c->grid_mc = c->grid_dec = cus * (occ > 0 ? occ : 1);
    } else if (prof_t) {
        ldpc::fill_code_info(0, 0, c->info);
        int occ_mc = 0, occ_dec = 0;
        if ((e = ldpc::t12_occupancy(&occ_mc)) != hipSuccess) return bail(e, "occupancy");
        if ((e = ldpc::t12_occupancy_decode(&occ_dec)) != hipSuccess) return bail(e, "occupancy");
        c->grid_mc = cus * (occ_mc > 0 ? occ_mc : 1);
        c->grid_dec = cus * (occ_dec > 0 ? occ_dec : 1);
    } else {
        std::vector<int32_t> rp, col;
        uint32_t n = 0, m = 0, np = 0, k_info = 0;
        if (code->family == LDPC_CODE_AR4JA) {
            ar4ja_rows(k_id, rate_id, rp, col, n, m, np);
            k_info = code->k_info;
        } else {
            n = code->n;
            m = code->m;
            np = code->n_punct;
            rp.assign(code->row_ptr, code->row_ptr + m + 1);
            if (rp[m] <= 0 || rp[m] > (1 << 24)) {
                ldpc_destroy(c);
                return fail(nullptr, LDPC_ERR_ARG, "matrix: bad row_ptr");
            }
            col.assign(code->col, code->col + rp[m]);
            k_info = n > m ? n - m : 0;  // rate (N - M) / (N - P) (ldpc_sim_scalar.cpp:352)
        }
        const int rc = setup_profile_r(c, rp, col, n, m, np, k_info);
        if (rc) {
            const std::string msg = c->err.empty() ? g_create_error : c->err;
            ldpc_destroy(c);
            g_create_error = msg;
            return rc;
        }
        int occ = 0;
        if ((e = ldpc::r_occupancy(c->rl, ldpc::r_launch_lds(c->rl), &occ)) !=
            hipSuccess)
            return bail(e, "occupancy");
        c->grid_mc = c->grid_dec = cus * (occ > 0 ? occ : 1);
        c->rl.grid = c->grid_mc;
        c->rl.queue = c->d_queue;
    }
    *out = c;
    return LDPC_OK;
}

void ldpc_destroy(ldpc_ctx *c) {
    if (!c) return;
    // Teardown is best effort: errors here have no caller to report to.
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (DevBuf *b : {&c->in, &c->out, &c->iters, &c->ok, &c->tails})
        if (b->p) (void)hipFree(b->p);
    if (c->d_table) (void)hipFree(c->d_table);
    if (c->d_kinds) (void)hipFree(c->d_kinds);
    if (c->d_slot) (void)hipFree(c->d_slot);
    if (c->d_cw_gen) (void)hipFree(c->d_cw_gen);
    if (c->d_alpha) (void)hipFree(c->d_alpha);
    if (c->d_cap_count) (void)hipFree(c->d_cap_count);
    if (c->d_cap_ids) (void)hipFree(c->d_cap_ids);
    if (c->d_work) (void)hipFree(c->d_work);
    if (c->d_beta) (void)hipFree(c->d_beta);
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

int ldpc_get_schedule(const ldpc_code_desc *code, uint32_t layer_order, int32_t *out, int32_t cap) {
    if (!code || !out || cap <= 0 || layer_order > 1) return LDPC_ERR_ARG;
    int k_id, rate_id;
    if (!ldpc::ar4ja_ids(*code, k_id, rate_id)) return LDPC_ERR_UNSUPPORTED;
    if (k_id == 0 && rate_id == 0 && layer_order == 0) {
        // the specialised kernel's own compiled table
        const int n = ldpc::t12_schedule(out, cap);
        return n < 0 ? LDPC_ERR_ARG : n;
    }
    const ar4ja::Schedule s = ar4ja::make_schedule(k_id, rate_id, (int)layer_order);
    int n = 0;
    for (int li = 0; li < s.n_layers; ++li) {
        const ar4ja::Layer &L = s.layer[li];
        if (n + 2 + 2 * L.deg > cap) return LDPC_ERR_ARG;
        out[n++] = L.first_row;
        out[n++] = L.deg;
        for (int d = 0; d < L.deg; ++d) {
            out[n++] = L.slot[d].base;
            out[n++] = L.slot[d].shift;
        }
    }
    return n;
}

int ldpc_list_variants(const ldpc_code_desc *code, uint32_t *out, int32_t cap) {
    if (!code || (!out && cap > 0) || cap < 0) return LDPC_ERR_ARG;
    int k_id, rate_id;
    if (!ldpc::ar4ja_ids(*code, k_id, rate_id)) return LDPC_ERR_UNSUPPORTED;
    int n = 0;
    const ldpc::MultiVariant *t = ldpc::multi_table(k_id, rate_id, &n);
    for (int i = 0; i < n && i < cap; ++i) out[i] = ldpc::variant_code(t[i]);
    return n;
}

int ldpc_channel_params(const ldpc_ctx *c, double ebn0_db, float *mu, float *sigma) {
    if (!c || !mu || !sigma || !std::isfinite(ebn0_db)) return LDPC_ERR_ARG;
    ldpc::channel_params(c->info, ebn0_db, mu, sigma);
    return LDPC_OK;
}

int ldpc_set_capture(ldpc_ctx *c, uint32_t max_frames) {
    if (!c) return LDPC_ERR_ARG;
    if (max_frames > (1u << 24)) return fail(c, LDPC_ERR_ARG, "max_frames must be <= 2^24");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->d_cap_ids) HIP_TRY(c, hipFree(c->d_cap_ids));
    c->d_cap_ids = nullptr;
    c->cap_max = 0;
    if (!max_frames) return LDPC_OK;
    if (!c->d_cap_count) HIP_TRY(c, hipMalloc(&c->d_cap_count, sizeof(uint32_t)));
    HIP_TRY(c, hipMalloc(&c->d_cap_ids, (size_t)max_frames * sizeof(unsigned long long)));
    HIP_TRY(c, ldpc::launch_fill_u32(c->d_cap_count, 1, 0u, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->cap_max = max_frames;
    return LDPC_OK;
}

int ldpc_get_captures(ldpc_ctx *c, uint64_t *cw_ids, uint32_t cap, uint64_t *n_failed) {
    if (!c || (!cw_ids && cap)) return LDPC_ERR_ARG;
    if (!c->cap_max) {
        if (n_failed) *n_failed = 0;
        return 0;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    uint32_t count = 0;
    HIP_TRY(c, hipMemcpy(&count, c->d_cap_count, sizeof(count), hipMemcpyDeviceToHost));
    const uint32_t kept = std::min(count, c->cap_max);
    std::vector<uint64_t> ids(kept);
    if (kept) HIP_TRY(c, hipMemcpy(ids.data(), c->d_cap_ids, kept * sizeof(uint64_t), hipMemcpyDeviceToHost));
    std::sort(ids.begin(), ids.end());
    const uint32_t n = std::min(kept, cap);
    std::copy(ids.begin(), ids.begin() + n, cw_ids);
    if (n_failed) *n_failed = count;
    return (int)n;
}

int ldpc_set_dmin_params(ldpc_ctx *c, const float *alpha, const float *beta, uint32_t groups, uint32_t n_params,
                         int correction, float r_max) {
    if (!c) return LDPC_ERR_ARG;
    if (c->cfg.profile != LDPC_PROFILE_D) return fail(c, LDPC_ERR_ARG, "not a Profile D context");
    if (!alpha || !beta || groups == 0 || groups > 100000) return fail(c, LDPC_ERR_ARG, "bad parameter arrays");
    if (n_params != c->info.n_edges / c->info.z)
        return fail(c, LDPC_ERR_ARG, "n_params must be the sum of the layer degrees (n_edges / z)");
    if (!(r_max >= 0.0f) || !std::isfinite(r_max)) return fail(c, LDPC_ERR_ARG, "r_max must be >= 0");
    if (correction < 0 || correction > 2)
        return fail(c, LDPC_ERR_ARG, "correction: 0 (min-sum), 1 (two minima) or 2 (three minima)");
    const size_t nb = (size_t)groups * n_params * sizeof(float);
    for (size_t i = 0; i < (size_t)groups * n_params; ++i)
        if (!std::isfinite(alpha[i]) || !std::isfinite(beta[i])) return fail(c, LDPC_ERR_ARG, "non-finite parameter");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->d_alpha) HIP_TRY(c, hipFree(c->d_alpha));
    if (c->d_beta) HIP_TRY(c, hipFree(c->d_beta));
    c->d_alpha = c->d_beta = nullptr;
    c->d_groups = 0;
    HIP_TRY(c, hipMalloc(&c->d_alpha, nb));
    HIP_TRY(c, hipMalloc(&c->d_beta, nb));
    HIP_TRY(c, hipMemcpy(c->d_alpha, alpha, nb, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_beta, beta, nb, hipMemcpyHostToDevice));
    c->d_groups = (int)groups;
    c->d_corr = correction;
    c->d_rmax = r_max;
    return LDPC_OK;
}

int ldpc_noise_tails(ldpc_ctx *c, uint64_t seed, uint64_t first_cw, uint64_t n_cw, const float *thresholds,
                     uint32_t n_thr, uint64_t *counts) {
    if (!c) return LDPC_ERR_ARG;
    if (!thresholds || !counts || n_thr < 1 || n_thr > (uint32_t)ldpc::NOISE_TAIL_MAX)
        return fail(c, LDPC_ERR_ARG, "thresholds: 1..16 values and a counts array are required");
    for (uint32_t k = 0; k < n_thr; ++k)
        if (!std::isfinite(thresholds[k])) return fail(c, LDPC_ERR_ARG, "thresholds must be finite");
    if (n_cw > (1ull << 40)) return fail(c, LDPC_ERR_ARG, "n_cw must be <= 2^40");
    int rc;
    if ((rc = ensure(c, c->tails, ldpc::NOISE_TAIL_MAX * sizeof(unsigned long long)))) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    auto *d = static_cast<unsigned long long *>(c->tails.p);
    HIP_TRY(c, ldpc::launch_fill_u64(d, n_thr, 0ull, c->stream));
    HIP_TRY(c, ldpc::launch_noise_tails(seed, first_cw, n_cw, (int)c->info.n_tx, thresholds, (int)n_thr, d, c->cus,
                                        c->stream));
    HIP_TRY(c, hipMemcpyAsync(counts, d, n_thr * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LDPC_OK;
}

int ldpc_set_codeword_mode(ldpc_ctx *c, uint32_t mode) {
    if (!c) return LDPC_ERR_ARG;
    if (mode == LDPC_CODEWORD_ZERO) {
        c->cw_random = false;
        return LDPC_OK;
    }
    if (mode != LDPC_CODEWORD_RANDOM) return fail(c, LDPC_ERR_ARG, "unknown codeword mode");
    if (c->cfg.profile != LDPC_PROFILE_T || c->mv || !is_r12_k1024(c->code))
        return fail(c, LDPC_ERR_UNSUPPORTED,
                    "random codewords are built into the Profile T kernel of AR4JA r1/2 K=1024 (variant 0)");
    const uint32_t *w;
    int z, nu;
    std::string why;
    const int rc = ldpc::qc_generator(0, 0, &w, &z, &nu, why);
    if (rc) return fail(c, rc, why);
    // transmitted parity blocks r = 0..7 of [12][nu = 8][z / 32 = 4]
    const size_t nb = sizeof(uint32_t) * (size_t)ldpc::t12_cw_gen_words();
    if (8 * nu * (z / 32) != ldpc::t12_cw_gen_words()) return fail(c, LDPC_ERR_HIP, "generator shape");
    HIP_TRY(c, hipSetDevice(c->device));
    if (!c->d_cw_gen) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipMalloc(&c->d_cw_gen, nb));
        HIP_TRY(c, hipMemcpy(c->d_cw_gen, w, nb, hipMemcpyHostToDevice));
    }
    c->cw_random = true;
    return LDPC_OK;
}

int ldpc_sync(ldpc_ctx *c) {
    if (!c) return LDPC_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LDPC_OK;
}

void *ldpc_stream(ldpc_ctx *c) { return c ? (void *)c->stream : nullptr; }

static float l_max_at(const ldpc_ctx *c, float mu) { return c->cfg.gamma > 0.0f ? c->cfg.gamma * mu : c->cfg.l_max; }

static ldpc::MultiParams multi_params(const ldpc_ctx *c) {
    ldpc::MultiParams p{};
    p.d_alpha = c->d_alpha;
    p.d_beta = c->d_beta;
    p.d_groups = c->d_groups;
    p.d_corr = c->d_corr;
    p.d_rmax = c->d_rmax;
    p.cap = ldpc::Capture{c->d_cap_count, c->d_cap_ids, c->cap_max};
    p.work = static_cast<uint32_t *>(c->d_work);
    p.queue = c->d_queue;
    p.counters = c->d_counters;
    p.hist = c->d_hist;
    p.max_iter = (int)c->cfg.max_iter;
    p.beta = c->cfg.beta;
    p.beta3 = c->cfg.beta < 1.0f ? c->cfg.beta : 1.0f;  // stepper:300-304
    return p;
}

int ldpc_run_point(ldpc_ctx *c, const ldpc_point_cfg *pt, ldpc_point_result *res, uint64_t *hist) {
    if (!c || !pt || !res) return fail(c, LDPC_ERR_ARG, "null argument");
    if (!std::isfinite(pt->ebn0_db)) return fail(c, LDPC_ERR_ARG, "ebn0_db must be finite");
    const uint64_t chunk = pt->chunk_cw ? pt->chunk_cw : (1ull << 20);
    // the device queue counter is 32-bit and also counts one exit ticket per wave
    if (chunk > (1ull << 31)) return fail(c, LDPC_ERR_ARG, "chunk_cw must be <= 2^31");
    if (c->cfg.profile == LDPC_PROFILE_D && !c->d_groups)
        return fail(c, LDPC_ERR_ARG, "Profile D: call ldpc_set_dmin_params first");
    HIP_TRY(c, hipSetDevice(c->device));
    float mu, sigma;
    ldpc_channel_params(c, pt->ebn0_db, &mu, &sigma);
    const bool prof_t = c->cfg.profile == LDPC_PROFILE_T;
    ldpc::RLaunch rl = c->rl;
    rl.l_max = l_max_at(c, mu);
    rl.cap = ldpc::Capture{c->d_cap_count, c->d_cap_ids, c->cap_max};
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
        if (c->mv) {
            ldpc::MultiParams p = multi_params(c);
            p.seed = pt->seed;
            p.first_cw = next;
            p.n_cw = (uint32_t)n;
            p.mu = mu;
            p.sigma = sigma;
            HIP_TRY(c, c->mv->launch(p, c->grid_mc, c->stream));
        } else if (prof_t)
            HIP_TRY(c, ldpc::t12_launch_mc(pt->seed, next, (uint32_t)n, mu, sigma, (int)c->cfg.max_iter, c->cfg.beta,
                                           c->d_queue, c->d_counters, c->d_hist, c->grid_mc, rl.cap,
                                           c->cw_random ? c->d_cw_gen : nullptr, c->stream));
        else
            HIP_TRY(c, ldpc::r_launch_mc(rl, pt->seed, next, (uint32_t)n, mu, sigma, c->d_counters, c->d_hist,
                                         c->stream));
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
    if (c->cfg.profile == LDPC_PROFILE_D && !c->d_groups)
        return fail(c, LDPC_ERR_ARG, "Profile D: call ldpc_set_dmin_params first");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, ldpc::launch_fill_u32(c->d_queue, 1, 0u, c->stream));
    if (c->mv) {
        HIP_TRY(c, ldpc::launch_fill_u64(c->d_counters, 4, 0ull, c->stream));
        HIP_TRY(c, ldpc::launch_fill_u64(c->d_hist, c->cfg.max_iter + 1, 0ull, c->stream));
        ldpc::MultiParams p = multi_params(c);
        p.cap = ldpc::Capture{nullptr, nullptr, 0};
        p.llr_in = (const uint16_t *)d_llr;
        p.llr_out = (uint16_t *)d_out_llr;
        p.iters = (uint32_t *)d_out_iters;
        p.ok = (uint8_t *)d_out_ok;
        p.n_cw = n_cw;
        HIP_TRY(c, c->mv->launch(p, c->grid_dec, c->stream));
    } else if (c->cfg.profile == LDPC_PROFILE_T)
        HIP_TRY(c, ldpc::t12_launch_decode((const uint16_t *)d_llr, (uint16_t *)d_out_llr, (uint32_t *)d_out_iters,
                                           (uint8_t *)d_out_ok, n_cw, (int)c->cfg.max_iter, c->cfg.beta, c->d_queue,
                                           c->grid_dec, c->stream));
    else
        HIP_TRY(c, ldpc::r_launch_decode(c->rl, (const uint16_t *)d_llr, (uint16_t *)d_out_llr,
                                         (uint32_t *)d_out_iters, (uint8_t *)d_out_ok, n_cw, c->stream));
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
    if (c->mv)  // same Philox / Box-Muller stream, size-generic generator, no clamp
        HIP_TRY(c, ldpc::r_launch_gen(seed, first_cw, n_cw, mu, sigma, 0.0f, (int)c->info.n_tx, (int)c->info.n,
                                      (uint16_t *)c->out.p, c->stream));
    else if (c->cfg.profile == LDPC_PROFILE_T && c->cw_random)
        HIP_TRY(c, ldpc::t12_launch_gen_cw(seed, first_cw, n_cw, mu, sigma, (uint16_t *)c->out.p, c->d_cw_gen,
                                           c->stream));
    else if (c->cfg.profile == LDPC_PROFILE_T)
        HIP_TRY(c, ldpc::t12_launch_gen(seed, first_cw, n_cw, mu, sigma, (uint16_t *)c->out.p, c->stream));
    else
        HIP_TRY(c, ldpc::r_launch_gen(seed, first_cw, n_cw, mu, sigma, l_max_at(c, mu), c->rl.n_tx, c->rl.n,
                                      (uint16_t *)c->out.p, c->stream));
    HIP_TRY(c, hipMemcpyAsync(out_llr, c->out.p, nb, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LDPC_OK;
}

}  // extern "C"
