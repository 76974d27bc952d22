"""Parallel schedules of the Profile R (scalar_5w) row order.

The reference decodes rows one after another in file order
(engines/scalar_5w/ldpc_decoder_core.h:313-521).  A row's level is one more
than the highest level of the earlier rows it shares a variable with; rows of
one level touch disjoint variables and all their earlier neighbours are in
earlier levels, so decoding level by level reads exactly the LLRs the
sequential order reads.  The kernel runs list-scheduled passes instead
(passes): a row joins a pass of at most `width` rows once every earlier row
sharing a variable with it is in an earlier pass, longest remaining chain
first.  libldpc builds both in C++ (ldpc_row_levels, ldpc_row_passes); this
module is the host-side view used by tests and tools.
"""
from __future__ import annotations

import heapq
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


def passes(rows: Sequence[Sequence[int]], n: int, width: int = 64) -> List[int]:
    """Pass (0-based) of every row in the list schedule described above."""
    m = len(rows)
    last = [-1] * n
    succ: List[List[int]] = [[] for _ in range(m)]
    indeg = [0] * m
    for i, r in enumerate(rows):
        for v in r:
            if last[v] >= 0 and last[v] != i:  # a repeated column is no edge to itself
                succ[last[v]].append(i)
                indeg[i] += 1
            last[v] = i
    height = [0] * m
    for i in range(m - 1, -1, -1):
        height[i] = 1 + max((height[s] for s in succ[i]), default=0)
    ready = [(-height[i], i) for i in range(m) if indeg[i] == 0]
    heapq.heapify(ready)
    out = [0] * m
    p = 0
    while ready:
        cur = [heapq.heappop(ready)[1] for _ in range(min(width, len(ready)))]
        for r in cur:
            out[r] = p
            for s in succ[r]:
                indeg[s] -= 1
                if indeg[s] == 0:
                    heapq.heappush(ready, (-height[s], s))
        p += 1
    if any(indeg):
        raise ValueError("rows could not be scheduled")
    return out
