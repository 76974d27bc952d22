"""AR4JA construction: dimensions, degrees, reference tables, schedule."""
import hashlib
import os

import numpy as np
import pytest

from conftest import REFERENCE
from ldpc_mi355x import codes

# Expected properties of the 9 codes (tools/validate_all_9_ar4ja_codes.m:31-41):
# K, rate, M, M/4, checks, variables, nnz, degree-3 rows, other degree, other-degree rows
VALIDATE_ALL_9 = [
    (1024, "1/2", 512, 128, 1536, 2560, 7680, 512, 6, 1024),
    (1024, "2/3", 256, 64, 768, 1792, 5888, 256, 10, 512),
    (1024, "4/5", 128, 32, 384, 1408, 4992, 128, 18, 256),
    (4096, "1/2", 2048, 512, 6144, 10240, 30720, 2048, 6, 4096),
    (4096, "2/3", 1024, 256, 3072, 7168, 23552, 1024, 10, 2048),
    (4096, "4/5", 512, 128, 1536, 5632, 19968, 512, 18, 1024),
    (16384, "1/2", 8192, 2048, 24576, 40960, 122880, 8192, 6, 16384),
    (16384, "2/3", 4096, 1024, 12288, 28672, 94208, 4096, 10, 8192),
    (16384, "4/5", 2048, 512, 6144, 22528, 79872, 2048, 18, 4096),
]

# sha256 over the ascending edge lists of r1/2 K=1024; equal to the digest of
# the reference's CHECK_TO_VAR table (ar4ja_circulant_descriptors_4cw.h:1019-2556)
R12_K1024_DIGEST = "508d62f7ea9200464cdd4717edaba26605807318fbe2194fe2b75e1a1ee59c47"


@pytest.mark.parametrize("row", VALIDATE_ALL_9, ids=lambda r: f"K{r[0]}_r{r[1].replace('/', '')}")
def test_nine_codes_match_validate_all_9(row):
    K, rate, M, L, m, n, nnz, n3, dother, nother = row
    c = codes.ar4ja(K, rate)
    assert codes.block_size(K, rate) == M and c.z == L
    assert (c.m, c.n, c.n_edges, c.k, c.n_punct) == (m, n, nnz, K, M)
    rd = c.row_degrees()
    assert (rd == 3).sum() == n3 and (rd == dother).sum() == nother
    assert c.n_tx == n - M and abs(c.rate - {"1/2": .5, "2/3": 2 / 3, "4/5": .8}[rate]) < 1e-12


def test_nested_family_property():
    """At one submatrix size M, H_1/2 is the last 5M columns and H_2/3 the last
    7M columns of H_4/5 (validate_all_9_ar4ja_codes.m:194-208)."""
    for M in (128, 512, 2048):
        h12, h23, h45 = (codes.ar4ja_m(M, r).dense() for r in ("1/2", "2/3", "4/5"))
        assert np.array_equal(h45[:, 6 * M:], h12)
        assert np.array_equal(h45[:, 4 * M:], h23)


def test_r12_k1024_digest_and_degrees(code12):
    assert code12.digest() == R12_K1024_DIGEST
    cd = code12.col_degrees()
    # transmitted: 512 of degree 1, 512 of degree 2, 1024 of degree 3; punctured: degree 6
    assert np.bincount(cd[:2048]).tolist() == [0, 512, 512, 1024]
    assert np.bincount(cd[2048:]).tolist() == [0, 0, 0, 0, 0, 0, 512]


def test_pi_is_permutation():
    for M in (128, 512, 2048, 8192):
        for k in range(1, 27):
            p = codes.pi_perm(k, M)
            assert np.array_equal(np.sort(p), np.arange(M))


def test_layers_are_collision_free_and_cover_h():
    for c in codes.all_ar4ja():
        seen_rows = []
        for first, slots in c.layers:
            seen_rows.extend(range(first, first + c.z))
            vars_in_layer = np.concatenate([s.base + (s.shift + np.arange(c.z)) % c.z for s in slots])
            assert len(np.unique(vars_in_layer)) == len(vars_in_layer)
            for j in (0, 1, c.z - 1):
                assert [s.base + (s.shift + j) % c.z for s in slots] == c.rows[first + j]
        assert sorted(seen_rows) == list(range(c.m))


def test_schedule_is_golden_tile_order(code12):
    """12 quarter layers = block rows 1, 2, 0; the golden's 40 tiles visit rows
    512..1535 then 0..511 (stepper:765-771)."""
    order = code12.schedule_rows()
    assert np.array_equal(order, np.concatenate([np.arange(512, 1536), np.arange(0, 512)]))


def test_chinn_roundtrip(tmp_path, code12):
    p = tmp_path / "h.chinn.out"
    codes.write_chinn_out(str(p), code12)
    M, N, rows = codes.read_chinn_out(str(p))
    assert (M, N) == (1536, 2560) and rows == code12.rows
    assert codes.parse_matrix_header(str(p)) == (2560, 1536, 6, 6)


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference tree not mounted")
def test_against_reference_tables(code12):
    """Re-read the reference's CHECK_TO_VAR and slice tables (text) and compare."""
    import re
    src = open(os.path.join(REFERENCE, "engines/circulant_dma_4cw/ar4ja_circulant_descriptors_4cw.h")).read()
    i = src.index("CHECK_TO_VAR[1536][6] = {")
    rows = re.findall(r"\{\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+) \}", src[i:src.index("};", i)])
    c2v = np.array(rows, dtype=np.int64)
    assert np.array_equal(code12.check_to_var(), c2v)
    # the shipped .chinn.out has the same dimensions and edge count
    M, N, r = codes.read_chinn_out(os.path.join(REFERENCE, "matrices/AR4JA_r45_4c_128c_r12.chinn.out"))
    assert (M, N, sum(map(len, r))) == (1536, 2560, 7680)
    assert codes.parse_matrix_header(os.path.join(REFERENCE, "matrices/AR4JA_r45_4c_128c_r12.chinn.out")) == (
        2560, 1536, 6, 6)
