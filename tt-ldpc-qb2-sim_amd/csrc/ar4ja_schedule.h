// ar4ja_schedule.h -- compile-time CCSDS AR4JA layer schedules.
//
// A layer is one quarter of a block row: Z = M/4 consecutive checks that share
// no variable.  Every edge slot of a layer is a cyclic shift inside one
// Z-variable column quarter:  variable(j) = base + (shift + j) mod Z  for the
// layer's check j.  Slots are listed in ascending base, which is ascending
// variable order for every check of the layer -- the edge order the
// reference's golden tree update consumes
// (tools/generate_circulant_dma_tables.py:150-153;
//  engines/circulant_dma_4cw/ar4ja_circulant_descriptors_4cw.h:43-1016).
//
// Construction: matrices/makeH_AR4JA.m:74-102 (theta/phi tables :19-72).
// Block-row order 1, 2, 0 (order_id 0) is the golden's tile order
// (stepper:765-771: tiles 0..31 cover rows M..3M-1, tiles 32..39 rows
// 0..M-1); order_id 1 decodes block rows 0, 1, 2 (the makeH row order).
#pragma once
#include <stdint.h>

namespace ar4ja {

// theta_k, k = 1..26
constexpr int THETA[26] = {3, 0, 1, 2, 2, 3, 0, 1, 0, 1, 2, 0, 2, 3, 0, 1, 2, 0, 1, 2, 0, 1, 2, 1, 2, 3};

// phi_j(k, M) for M = 128 * 2^t, t = 0..6; [j][k-1][t]
constexpr int PHI[4][26][7] = {
    {{1, 59, 16, 160, 108, 226, 1148}, {22, 18, 103, 241, 126, 618, 2032}, {0, 52, 105, 185, 238, 404, 249},
     {26, 23, 0, 251, 481, 32, 1807}, {0, 11, 50, 209, 96, 912, 485}, {10, 7, 29, 103, 28, 950, 1044},
     {5, 22, 115, 90, 59, 534, 717}, {18, 25, 30, 184, 225, 63, 873}, {3, 27, 92, 248, 323, 971, 364},
     {22, 30, 78, 12, 28, 304, 1926}, {3, 43, 70, 111, 386, 409, 1241}, {8, 14, 66, 66, 305, 708, 1769},
     {25, 46, 39, 173, 34, 719, 532}, {25, 62, 84, 42, 510, 176, 768}, {2, 44, 79, 157, 147, 743, 1138},
     {27, 12, 70, 174, 199, 759, 965}, {7, 38, 29, 104, 347, 674, 141}, {7, 47, 32, 144, 391, 958, 1527},
     {15, 1, 45, 43, 165, 984, 505}, {10, 52, 113, 181, 414, 11, 1312}, {4, 61, 86, 250, 97, 413, 1840},
     {19, 10, 1, 202, 158, 925, 709}, {7, 55, 42, 68, 86, 687, 1427}, {9, 7, 118, 177, 168, 752, 989},
     {26, 12, 33, 170, 506, 867, 1925}, {17, 2, 126, 89, 489, 323, 270}},
    {{0, 0, 0, 0, 0, 0, 0}, {27, 32, 53, 182, 375, 767, 1822}, {30, 21, 74, 249, 436, 227, 203},
     {28, 36, 45, 65, 350, 247, 882}, {7, 30, 47, 70, 260, 284, 1989}, {1, 29, 0, 141, 84, 370, 957},
     {8, 44, 59, 237, 318, 482, 1705}, {20, 29, 102, 77, 382, 273, 1083}, {26, 39, 25, 55, 169, 886, 1072},
     {24, 14, 3, 12, 213, 634, 354}, {4, 22, 88, 227, 67, 762, 1942}, {12, 15, 65, 42, 313, 184, 446},
     {23, 48, 62, 52, 242, 696, 1456}, {15, 55, 68, 243, 188, 413, 1940}, {15, 39, 91, 179, 1, 854, 1660},
     {22, 11, 70, 250, 306, 544, 1661}, {31, 1, 115, 247, 397, 864, 587}, {3, 50, 31, 164, 80, 82, 708},
     {29, 40, 121, 17, 33, 1009, 1466}, {21, 62, 45, 31, 7, 437, 433}, {2, 27, 56, 149, 447, 36, 1345},
     {5, 38, 54, 105, 336, 562, 867}, {11, 40, 108, 183, 424, 816, 1551}, {26, 15, 14, 153, 134, 452, 2041},
     {9, 11, 30, 177, 152, 290, 1383}, {17, 18, 116, 19, 492, 778, 1790}},
    {{0, 0, 0, 0, 0, 0, 0}, {12, 46, 8, 35, 219, 254, 318}, {30, 45, 119, 167, 16, 790, 494},
     {18, 27, 89, 214, 263, 642, 1467}, {10, 48, 31, 84, 415, 248, 757}, {16, 37, 122, 206, 403, 899, 1085},
     {13, 41, 1, 122, 184, 328, 1630}, {9, 13, 69, 67, 279, 518, 64}, {7, 9, 92, 147, 198, 477, 689},
     {15, 49, 47, 54, 307, 404, 1300}, {16, 36, 11, 23, 432, 698, 148}, {18, 10, 31, 93, 240, 160, 777},
     {4, 11, 19, 20, 454, 497, 1431}, {23, 18, 66, 197, 294, 100, 659}, {5, 54, 49, 46, 479, 518, 352},
     {3, 40, 81, 162, 289, 92, 1177}, {29, 27, 96, 101, 373, 464, 836}, {11, 35, 38, 76, 104, 592, 1572},
     {4, 25, 83, 78, 141, 198, 348}, {8, 46, 42, 253, 270, 856, 1040}, {2, 24, 58, 124, 439, 235, 779},
     {11, 33, 24, 143, 333, 134, 476}, {11, 18, 25, 63, 399, 542, 191}, {3, 37, 92, 41, 14, 545, 1393},
     {15, 35, 38, 214, 277, 777, 1752}, {13, 21, 120, 70, 412, 483, 1627}},
    {{0, 0, 0, 0, 0, 0, 0}, {13, 44, 35, 162, 312, 285, 1189}, {19, 51, 97, 7, 503, 554, 458},
     {14, 12, 112, 31, 388, 809, 460}, {15, 15, 64, 164, 48, 185, 1039}, {20, 12, 93, 11, 7, 49, 1000},
     {17, 4, 99, 237, 185, 101, 1265}, {4, 7, 94, 125, 328, 82, 1223}, {4, 2, 103, 133, 254, 898, 874},
     {11, 30, 91, 99, 202, 627, 1292}, {17, 53, 3, 105, 285, 154, 1491}, {20, 23, 6, 17, 11, 65, 631},
     {8, 29, 39, 97, 168, 81, 464}, {22, 37, 113, 91, 127, 823, 461}, {19, 42, 92, 211, 8, 50, 844},
     {15, 48, 119, 128, 437, 413, 392}, {5, 4, 74, 82, 475, 462, 922}, {21, 10, 73, 115, 85, 175, 256},
     {17, 18, 116, 248, 419, 715, 1986}, {9, 56, 31, 62, 459, 537, 19}, {20, 9, 127, 26, 468, 722, 266},
     {18, 11, 98, 140, 209, 37, 471}, {31, 23, 23, 121, 311, 488, 1166}, {13, 8, 38, 12, 211, 179, 1300},
     {2, 7, 18, 41, 510, 430, 1033}, {18, 24, 62, 249, 320, 264, 1606}}};

// One term of the block structure: (block row, block column, k) with k = 0
// for the identity and k = 1..26 for pi_k (makeH_AR4JA.m:94-100).
struct Term {
    int br, bc, k;
};
constexpr Term BASE_12[15] = {{0, 2, 0}, {0, 4, 0}, {0, 4, 1}, {1, 0, 0}, {1, 1, 0}, {1, 3, 0}, {1, 4, 2}, {1, 4, 3},
                              {1, 4, 4},  {2, 0, 0}, {2, 1, 5}, {2, 1, 6}, {2, 3, 7}, {2, 3, 8}, {2, 4, 0}};
constexpr Term EXT_23[8] = {{1, 0, 9}, {1, 0, 10}, {1, 0, 11}, {1, 1, 0},
                            {2, 0, 0}, {2, 1, 12}, {2, 1, 13}, {2, 1, 14}};
constexpr Term EXT_45[16] = {{1, 0, 21}, {1, 0, 22}, {1, 0, 23}, {1, 1, 0}, {1, 2, 15}, {1, 2, 16},
                             {1, 2, 17}, {1, 3, 0},  {2, 0, 0},  {2, 1, 24}, {2, 1, 25}, {2, 1, 26},
                             {2, 2, 0},  {2, 3, 18}, {2, 3, 19}, {2, 3, 20}};

constexpr int MAX_DEG = 18;

struct Slot {
    int base, shift;
};
struct Layer {
    int first_row, deg;
    Slot slot[MAX_DEG];
};

// rate_id: 0 = 1/2, 1 = 2/3, 2 = 4/5;  k_id: 0 = K 1024, 1 = 4096, 2 = 16384
constexpr int block_size(int k_id, int rate_id) { return (512 << (2 * k_id)) >> rate_id; }
constexpr int n_block_cols(int rate_id) { return rate_id == 0 ? 5 : rate_id == 1 ? 7 : 11; }

struct Schedule {
    int M, Z, N, n_checks, n_punct, k_info, n_layers, n_edges;
    Layer layer[12];
};

constexpr int ilog2(int x) {
    int r = 0;
    while ((1 << r) < x) ++r;
    return r;
}

constexpr Schedule make_schedule(int k_id, int rate_id, int order_id = 0) {
    Schedule s{};
    const int M = block_size(k_id, rate_id), Z = M / 4, t = ilog2(M) - 7;
    s.M = M;
    s.Z = Z;
    s.N = n_block_cols(rate_id) * M;
    s.n_checks = 3 * M;
    s.n_punct = M;
    s.k_info = 1024 << (2 * k_id);
    s.n_layers = 12;
    s.n_edges = 0;
    const int order[2][3] = {{1, 2, 0}, {0, 1, 2}};
    int li = 0;
    for (int oi = 0; oi < 3; ++oi) {
        const int br = order[order_id][oi];
        for (int j = 0; j < 4; ++j) {
            Layer L{};
            L.first_row = br * M + j * Z;
            L.deg = 0;
            auto add = [&](const Term &tm, int col_off) {
                if (tm.br != br) return;
                int q = j, sh = 0;
                if (tm.k) {
                    q = (THETA[tm.k - 1] + j) % 4;
                    sh = PHI[j][tm.k - 1][t] % Z;
                }
                L.slot[L.deg].base = (tm.bc + col_off) * M + q * Z;
                L.slot[L.deg].shift = sh;
                L.deg++;
            };
            if (rate_id == 2)
                for (const Term &tm : EXT_45) add(tm, 0);
            if (rate_id >= 1)
                for (const Term &tm : EXT_23) add(tm, rate_id == 2 ? 4 : 0);
            for (const Term &tm : BASE_12) add(tm, rate_id == 2 ? 6 : rate_id == 1 ? 2 : 0);
            // ascending base (insertion sort; slots never share a base)
            for (int a = 1; a < L.deg; ++a)
                for (int b = a; b > 0 && L.slot[b - 1].base > L.slot[b].base; --b) {
                    Slot tmp = L.slot[b - 1];
                    L.slot[b - 1] = L.slot[b];
                    L.slot[b] = tmp;
                }
            s.n_edges += L.deg * Z;
            s.layer[li++] = L;
        }
    }
    return s;
}

}  // namespace ar4ja
