"""CCSDS AR4JA parity-check matrices, layer schedules and `.chinn.out` I/O.

The construction follows the CCSDS AR4JA definition as restated by the
reference's MATLAB builder (matrices/makeH_AR4JA.m:13-102): three block rows
of M x M submatrices, each submatrix a sum of identity and permutation
matrices pi_k, with

    pi_k(i) = (M/4) * ((theta_k + floor(4 i / M)) mod 4)
              + (phi_k(floor(4 i / M), M) + i) mod (M/4)

and the last M columns punctured.  The theta/phi constants are the standard's
tables (makeH_AR4JA.m:19-72); they are data that must be identical to the
standard, so they are kept here as plain tuples.

Every check's edges are listed in ascending variable order.  That is the
order the reference's golden decoder feeds its tree Approx-Min* update
(tools/generate_circulant_dma_tables.py:150-153 picks the d-th nonzero of a
row in ascending column order; the 40-tile slice tables in
engines/circulant_dma_4cw/ar4ja_circulant_descriptors_4cw.h:43-1016 gather
exactly that order).
"""
from __future__ import annotations

import dataclasses
import hashlib
import os
from typing import List, Sequence, Tuple

import numpy as np

# Standard block sizes M, indexed [K][rate]  (makeH_AR4JA.m:13-16).
_M_TABLE = {
    1024: {"1/2": 512, "2/3": 256, "4/5": 128},
    4096: {"1/2": 2048, "2/3": 1024, "4/5": 512},
    16384: {"1/2": 8192, "2/3": 4096, "4/5": 2048},
}
RATES = ("1/2", "2/3", "4/5")
BLOCK_LENGTHS = (1024, 4096, 16384)

# theta_k, k = 1..26  (makeH_AR4JA.m:19-45, column 2).
THETA = (3, 0, 1, 2, 2, 3, 0, 1, 0, 1, 2, 0, 2, 3, 0, 1, 2, 0, 1, 2, 0, 1, 2, 1, 2, 3)

# phi_k(j, M) for j = 0..3 and M = 128, 256, ..., 8192  (makeH_AR4JA.m:19-72).
# PHI[j][k-1] is the 7-tuple over M.
PHI = (
    (  # j = 0
        (1, 59, 16, 160, 108, 226, 1148), (22, 18, 103, 241, 126, 618, 2032),
        (0, 52, 105, 185, 238, 404, 249), (26, 23, 0, 251, 481, 32, 1807),
        (0, 11, 50, 209, 96, 912, 485), (10, 7, 29, 103, 28, 950, 1044),
        (5, 22, 115, 90, 59, 534, 717), (18, 25, 30, 184, 225, 63, 873),
        (3, 27, 92, 248, 323, 971, 364), (22, 30, 78, 12, 28, 304, 1926),
        (3, 43, 70, 111, 386, 409, 1241), (8, 14, 66, 66, 305, 708, 1769),
        (25, 46, 39, 173, 34, 719, 532), (25, 62, 84, 42, 510, 176, 768),
        (2, 44, 79, 157, 147, 743, 1138), (27, 12, 70, 174, 199, 759, 965),
        (7, 38, 29, 104, 347, 674, 141), (7, 47, 32, 144, 391, 958, 1527),
        (15, 1, 45, 43, 165, 984, 505), (10, 52, 113, 181, 414, 11, 1312),
        (4, 61, 86, 250, 97, 413, 1840), (19, 10, 1, 202, 158, 925, 709),
        (7, 55, 42, 68, 86, 687, 1427), (9, 7, 118, 177, 168, 752, 989),
        (26, 12, 33, 170, 506, 867, 1925), (17, 2, 126, 89, 489, 323, 270),
    ),
    (  # j = 1
        (0, 0, 0, 0, 0, 0, 0), (27, 32, 53, 182, 375, 767, 1822),
        (30, 21, 74, 249, 436, 227, 203), (28, 36, 45, 65, 350, 247, 882),
        (7, 30, 47, 70, 260, 284, 1989), (1, 29, 0, 141, 84, 370, 957),
        (8, 44, 59, 237, 318, 482, 1705), (20, 29, 102, 77, 382, 273, 1083),
        (26, 39, 25, 55, 169, 886, 1072), (24, 14, 3, 12, 213, 634, 354),
        (4, 22, 88, 227, 67, 762, 1942), (12, 15, 65, 42, 313, 184, 446),
        (23, 48, 62, 52, 242, 696, 1456), (15, 55, 68, 243, 188, 413, 1940),
        (15, 39, 91, 179, 1, 854, 1660), (22, 11, 70, 250, 306, 544, 1661),
        (31, 1, 115, 247, 397, 864, 587), (3, 50, 31, 164, 80, 82, 708),
        (29, 40, 121, 17, 33, 1009, 1466), (21, 62, 45, 31, 7, 437, 433),
        (2, 27, 56, 149, 447, 36, 1345), (5, 38, 54, 105, 336, 562, 867),
        (11, 40, 108, 183, 424, 816, 1551), (26, 15, 14, 153, 134, 452, 2041),
        (9, 11, 30, 177, 152, 290, 1383), (17, 18, 116, 19, 492, 778, 1790),
    ),
    (  # j = 2
        (0, 0, 0, 0, 0, 0, 0), (12, 46, 8, 35, 219, 254, 318),
        (30, 45, 119, 167, 16, 790, 494), (18, 27, 89, 214, 263, 642, 1467),
        (10, 48, 31, 84, 415, 248, 757), (16, 37, 122, 206, 403, 899, 1085),
        (13, 41, 1, 122, 184, 328, 1630), (9, 13, 69, 67, 279, 518, 64),
        (7, 9, 92, 147, 198, 477, 689), (15, 49, 47, 54, 307, 404, 1300),
        (16, 36, 11, 23, 432, 698, 148), (18, 10, 31, 93, 240, 160, 777),
        (4, 11, 19, 20, 454, 497, 1431), (23, 18, 66, 197, 294, 100, 659),
        (5, 54, 49, 46, 479, 518, 352), (3, 40, 81, 162, 289, 92, 1177),
        (29, 27, 96, 101, 373, 464, 836), (11, 35, 38, 76, 104, 592, 1572),
        (4, 25, 83, 78, 141, 198, 348), (8, 46, 42, 253, 270, 856, 1040),
        (2, 24, 58, 124, 439, 235, 779), (11, 33, 24, 143, 333, 134, 476),
        (11, 18, 25, 63, 399, 542, 191), (3, 37, 92, 41, 14, 545, 1393),
        (15, 35, 38, 214, 277, 777, 1752), (13, 21, 120, 70, 412, 483, 1627),
    ),
    (  # j = 3
        (0, 0, 0, 0, 0, 0, 0), (13, 44, 35, 162, 312, 285, 1189),
        (19, 51, 97, 7, 503, 554, 458), (14, 12, 112, 31, 388, 809, 460),
        (15, 15, 64, 164, 48, 185, 1039), (20, 12, 93, 11, 7, 49, 1000),
        (17, 4, 99, 237, 185, 101, 1265), (4, 7, 94, 125, 328, 82, 1223),
        (4, 2, 103, 133, 254, 898, 874), (11, 30, 91, 99, 202, 627, 1292),
        (17, 53, 3, 105, 285, 154, 1491), (20, 23, 6, 17, 11, 65, 631),
        (8, 29, 39, 97, 168, 81, 464), (22, 37, 113, 91, 127, 823, 461),
        (19, 42, 92, 211, 8, 50, 844), (15, 48, 119, 128, 437, 413, 392),
        (5, 4, 74, 82, 475, 462, 922), (21, 10, 73, 115, 85, 175, 256),
        (17, 18, 116, 248, 419, 715, 1986), (9, 56, 31, 62, 459, 537, 19),
        (20, 9, 127, 26, 468, 722, 266), (18, 11, 98, 140, 209, 37, 471),
        (31, 23, 23, 121, 311, 488, 1166), (13, 8, 38, 12, 211, 179, 1300),
        (2, 7, 18, 41, 510, 430, 1033), (18, 24, 62, 249, 320, 264, 1606),
    ),
)

# Block structure of H.  Each entry is (block_row, block_col, term) where term
# is "I" or a permutation index k (1-based).  Rate 1/2 occupies the last five
# block columns; rate 2/3 prepends two and rate 4/5 prepends four more
# (makeH_AR4JA.m:94-100).
_BASE_12 = (
    (0, 2, "I"), (0, 4, "I"), (0, 4, 1),
    (1, 0, "I"), (1, 1, "I"), (1, 3, "I"), (1, 4, 2), (1, 4, 3), (1, 4, 4),
    (2, 0, "I"), (2, 1, 5), (2, 1, 6), (2, 3, 7), (2, 3, 8), (2, 4, "I"),
)
_EXT_23 = (  # two extra block columns (relative index 0, 1)
    (1, 0, 9), (1, 0, 10), (1, 0, 11), (1, 1, "I"),
    (2, 0, "I"), (2, 1, 12), (2, 1, 13), (2, 1, 14),
)
_EXT_45 = (  # four extra block columns (relative index 0..3)
    (1, 0, 21), (1, 0, 22), (1, 0, 23), (1, 1, "I"), (1, 2, 15), (1, 2, 16), (1, 2, 17), (1, 3, "I"),
    (2, 0, "I"), (2, 1, 24), (2, 1, 25), (2, 1, 26), (2, 2, "I"), (2, 3, 18), (2, 3, 19), (2, 3, 20),
)


def block_size(K: int, rate: str) -> int:
    """Submatrix size M for (K, rate) (makeH_AR4JA.m:77-83)."""
    try:
        return _M_TABLE[K][rate]
    except KeyError:
        raise ValueError(f"no AR4JA code for K={K} rate={rate}") from None


def pi_shift(k: int, j: int, M: int) -> Tuple[int, int]:
    """(target quarter, in-quarter shift) of pi_k for row quarter j.

    pi_k maps row i of quarter j = floor(4 i / M) to column
    (M/4) * ((theta_k + j) mod 4) + (phi_k(j, M) + i) mod (M/4)
    (makeH_AR4JA.m:88-91).
    """
    phitbl = M.bit_length() - 1 - 7  # log2(M) - 7 -> 0 for M = 128
    return (THETA[k - 1] + j) % 4, PHI[j][k - 1][phitbl] % (M // 4)


def pi_perm(k: int, M: int) -> np.ndarray:
    i = np.arange(M)
    Z = M // 4
    out = np.empty(M, dtype=np.int64)
    for j in range(4):
        q, s = pi_shift(k, j, M)
        rows = i[j * Z:(j + 1) * Z]
        out[j * Z:(j + 1) * Z] = Z * q + (s + rows) % Z
    return out


@dataclasses.dataclass(frozen=True)
class LayerEdge:
    """One edge slot of a layer: variable = base + (shift + j) mod Z for check j."""
    base: int
    shift: int


@dataclasses.dataclass
class Code:
    """A parity-check code with ascending edge lists and a layered schedule.

    Attributes:
        name: human name, e.g. ``"AR4JA r1/2 K=1024"``.
        n: number of variables N (transmitted + punctured).
        m: number of checks.
        k: information length.
        n_punct: number of punctured variables (always the last ones).
        z: circulant (layer) size, M/4 for AR4JA.
        rows: list of ascending variable lists, one per check (makeH row order).
        layers: list of layers; each is (first_row, [LayerEdge, ...]) with rows
            first_row .. first_row+z-1.  Layer order is the decoding order.
    """
    name: str
    n: int
    m: int
    k: int
    n_punct: int
    z: int
    rows: List[List[int]]
    layers: List[Tuple[int, List[LayerEdge]]]

    @property
    def n_tx(self) -> int:
        return self.n - self.n_punct

    @property
    def rate(self) -> float:
        return self.k / self.n_tx

    @property
    def n_edges(self) -> int:
        return sum(len(r) for r in self.rows)

    def row_degrees(self) -> np.ndarray:
        return np.array([len(r) for r in self.rows], dtype=np.int64)

    def col_degrees(self) -> np.ndarray:
        deg = np.zeros(self.n, dtype=np.int64)
        for r in self.rows:
            deg[r] += 1
        return deg

    def check_to_var(self, pad: int = 0xFFFF) -> np.ndarray:
        dmax = int(self.row_degrees().max())
        out = np.full((self.m, dmax), pad, dtype=np.int64)
        for i, r in enumerate(self.rows):
            out[i, :len(r)] = r
        return out

    def dense(self) -> np.ndarray:
        H = np.zeros((self.m, self.n), dtype=np.uint8)
        for i, r in enumerate(self.rows):
            H[i, r] = 1
        return H

    def schedule_rows(self) -> np.ndarray:
        """Check indices in decoding order (layer by layer, ascending within)."""
        return np.concatenate([np.arange(f, f + self.z) for f, _ in self.layers])

    def digest(self) -> str:
        h = hashlib.sha256()
        for r in self.rows:
            h.update(np.asarray(r, dtype="<u4").tobytes())
            h.update(b"|")
        return h.hexdigest()


def _block_terms(rate: str):
    """Block terms for a rate with block-column indices shifted to absolute."""
    if rate == "1/2":
        return list(_BASE_12)
    if rate == "2/3":
        return [(r, c, t) for r, c, t in _EXT_23] + [(r, c + 2, t) for r, c, t in _BASE_12]
    if rate == "4/5":
        return ([(r, c, t) for r, c, t in _EXT_45] + [(r, c + 4, t) for r, c, t in _EXT_23]
                + [(r, c + 6, t) for r, c, t in _BASE_12])
    raise ValueError(f"unknown rate {rate}")


# Profile T decodes block row 1, then 2, then 0 (the golden's 40-tile order:
# tiles 0..31 are rows M..3M-1, tiles 32..39 rows 0..M-1,
# tests/test_circulant_dma_stepper.cpp:765-771 with
# tools/generate_circulant_dma_tables.py:227-246).
BLOCK_ROW_ORDER_T = (1, 2, 0)


def ar4ja(K: int = 1024, rate: str = "1/2", block_row_order: Sequence[int] = BLOCK_ROW_ORDER_T) -> Code:
    """Build the CCSDS AR4JA code (K, rate) with its quarter-layer schedule."""
    return ar4ja_m(block_size(K, rate), rate, block_row_order)


def ar4ja_m(M: int, rate: str = "1/2", block_row_order: Sequence[int] = BLOCK_ROW_ORDER_T) -> Code:
    """AR4JA parity-check structure of a rate for submatrix size M (a power of
    two, 128..8192); K follows as M * (number of block columns - 3)."""
    if M < 128 or M > 8192 or M & (M - 1):
        raise ValueError(f"submatrix size M={M} is not one of 128, 256, ..., 8192")
    Z = M // 4
    terms = _block_terms(rate)
    n_blk_cols = max(c for _, c, _ in terms) + 1
    N = n_blk_cols * M
    rows: List[List[int]] = [[] for _ in range(3 * M)]
    # per (block_row, quarter) -> list of LayerEdge
    layer_edges = {(b, j): [] for b in range(3) for j in range(4)}
    for b, c, t in terms:
        for j in range(4):
            if t == "I":
                q, s = j, 0
            else:
                q, s = pi_shift(t, j, M)
            base = c * M + q * Z
            layer_edges[(b, j)].append(LayerEdge(base, s))
            for i in range(Z):
                rows[b * M + j * Z + i].append(base + (s + i) % Z)
    for r in rows:
        r.sort()
        assert len(set(r)) == len(r)
    layers = []
    for b in block_row_order:
        for j in range(4):
            # Edge slots in ascending variable order; the order is the same for
            # every check of the quarter because slots live in distinct
            # (block column, quarter) cells.
            edges = sorted(layer_edges[(b, j)], key=lambda e: e.base)
            bases = [e.base for e in edges]
            assert len(set(bases)) == len(bases)
            layers.append((b * M + j * Z, edges))
    K = N - 3 * M
    return Code(name=f"AR4JA r{rate} K={K}", n=N, m=3 * M, k=K, n_punct=M, z=Z, rows=rows, layers=layers)


def all_ar4ja() -> List[Code]:
    return [ar4ja(K, r) for K in BLOCK_LENGTHS for r in RATES]


# ---------------------------------------------------------------------------
# .chinn.out / .chinn files (run_simulation.py:9-55, ldpc_sim_scalar.cpp:116-268)
# ---------------------------------------------------------------------------

def read_chinn_out(path: str) -> Tuple[int, int, List[List[int]]]:
    """Parse a ``.chinn.out`` triplet file into (M, N, rows-in-file-order).

    Lines are ``row col val`` (1-indexed); the last line gives (M, N, 0).
    Entries with val == 0 are ignored.  Within a row, edges keep the order in
    which they appear in the file (ldpc_sim_scalar.cpp:116-202).
    """
    if not os.path.exists(path):
        raise FileNotFoundError(f"Matrix file not found: {path}")
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip() and not ln.startswith("#")]
    if not lines:
        raise ValueError(f"the matrix file {path} holds no lines")
    M, N = int(lines[-1][0]), int(lines[-1][1])
    rows: List[List[int]] = [[] for _ in range(M)]
    for parts in lines[:-1]:
        if len(parts) < 3 or float(parts[2]) == 0.0:
            continue
        r, c = int(parts[0]) - 1, int(parts[1]) - 1
        if not (0 <= r < M and 0 <= c < N):
            raise ValueError(f"entry ({r + 1},{c + 1}) outside {M}x{N}")
        rows[r].append(c)
    return M, N, rows


def write_chinn_out(path: str, code: Code) -> None:
    with open(path, "w") as f:
        for i, r in enumerate(code.rows):
            for v in r:
                f.write(f"{i + 1}\t{v + 1}\t1.000000 \n")
        f.write(f"{code.m}\t{code.n}\t0.000000\n")


def parse_matrix_header(path: str) -> Tuple[int, int, int, int]:
    """(N, M, max_var_deg, max_check_deg) of a matrix file (run_simulation.py:9-55)."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"Matrix file not found: {path}")
    with open(path) as f:
        lines = [ln.strip() for ln in f if ln.strip() and not ln.startswith("#")]
    if not lines:
        raise ValueError(f"the matrix file {path} holds no lines")
    first = lines[0].split()
    if len(first) == 3:
        M, N = int(lines[-1].split()[0]), int(lines[-1].split()[1])
        cdeg = [0] * M
        vdeg = [0] * N
        for ln in lines[:-1]:
            p = ln.split()
            if len(p) >= 3 and float(p[2]) != 0.0:
                r, c = int(p[0]) - 1, int(p[1]) - 1
                if 0 <= r < M:
                    cdeg[r] += 1
                if 0 <= c < N:
                    vdeg[c] += 1
        return N, M, (max(vdeg) if vdeg else 0), (max(cdeg) if cdeg else 0)
    deg = lines[1].split()
    N, M = int(first[0]), int(first[1])
    if len(first) >= 4:
        return N, M, int(first[2]), int(first[3])
    return N, M, int(deg[0]), int(deg[1])
