"""Level schedule of the Profile R (scalar_5w) row order.

The reference decodes rows one after another in file order
(engines/scalar_5w/ldpc_decoder_core.h:313-521).  A row's level is one more
than the highest level of the earlier rows it shares a variable with; rows of
one level touch disjoint variables and all their earlier neighbours are in
earlier levels, so decoding level by level reads exactly the LLRs the
sequential order reads.  libldpc builds the same levels in C++
(ldpc_row_levels); this module is the host-side view used by tests and tools.
"""
from __future__ import annotations

from typing import List, Sequence


def row_levels(rows: Sequence[Sequence[int]], n: int) -> List[int]:
    last = [0] * n
    out = []
    for r in rows:
        lv = 1 + max(last[v] for v in r)
        out.append(lv)
        for v in r:
            last[v] = lv
    return out


def levels(rows: Sequence[Sequence[int]], n: int) -> List[List[int]]:
    """Rows grouped by level, file order inside a level."""
    lv = row_levels(rows, n)
    out: List[List[int]] = [[] for _ in range(max(lv) if lv else 0)]
    for i, l in enumerate(lv):
        out[l - 1].append(i)
    return out
