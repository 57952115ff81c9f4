"""Independent formulations used to pin the oracle and the HIP library.

None of these retype the oracle's procedure: the build order is the closed-form
key tuple (not per-rank radix rounds), CPBT is a brute-force DP, BFS is a
textbook queue BFS, CC is scipy's connected components, SSSP is a float32
Dijkstra and PageRank a dense-matrix FP64 power iteration.
"""
from __future__ import annotations

import heapq

import numpy as np

REC8 = np.dtype([("src", "<u4"), ("dst", "<u4")])
REC12 = np.dtype([("src", "<u4"), ("dst", "<u4"), ("w", "<f4")])
PAGE = 4096


def random_edges(rng, V, E, weighted=False, skew=True):
    """Random edge list with hubs (skew), self-loops and duplicates allowed."""
    if skew:
        # power-law-ish: low ids are hubs
        src = np.minimum((rng.pareto(1.2, E) * V / 50).astype(np.int64), V - 1)
        dst = np.minimum((rng.pareto(1.1, E) * V / 50).astype(np.int64), V - 1)
        perm = rng.permutation(V)
        src, dst = perm[src], perm[dst]
    else:
        src = rng.integers(0, V, E)
        dst = rng.integers(0, V, E)
    rec = np.zeros(E, dtype=REC12 if weighted else REC8)
    rec["src"] = src
    rec["dst"] = dst
    if weighted:
        rec["w"] = rng.random(E, dtype=np.float32)
    return rec


def logical_list(edges, P, symmetrize):
    """Rank slices (core/graph.hpp:1024-1030) with per-rank symmetrisation
    (:1057-1078).  Returns (records in logical order, rank of each)."""
    E0 = edges.shape[0]
    parts, ranks = [], []
    for r in range(P):
        n = E0 // P + (E0 % P if r == P - 1 else 0)
        off = (E0 // P) * r
        sl = edges[off:off + n].copy()
        if symmetrize:
            rev = sl.copy()
            rev["src"], rev["dst"] = sl["dst"], sl["src"]
            sl = np.concatenate([sl, rev])
        parts.append(sl)
        ranks.append(np.full(sl.shape[0], r, dtype=np.int64))
    return np.concatenate(parts), np.concatenate(ranks)


def sub_partition(v, partition_offset, local_offsets):
    """q = p*S + s of each vertex id."""
    P = len(partition_offset) - 1
    S = len(local_offsets[0]) - 1
    p = np.searchsorted(partition_offset, v, side="right") - 1
    p = np.clip(p, 0, P - 1)
    s = np.zeros_like(p)
    lo = np.stack(local_offsets)  # [P, S+1]
    for k in range(1, S):
        s += (v >= lo[p, k]).astype(p.dtype)
    return p * S + s


def expected_order(recs, ranks, P, partition_offset, local_offsets, pull):
    """Closed-form order of the two-level build.

    CSC (pull): sub-partition q(src); P>1 key (q, dst, r, src, j), P=1 (q, dst, j).
    CSR (push): sub-partition q(dst); P>1 key (q, src, r, dst, j), P=1 (q, src, j).
    Returns (permutation, q of each record after permutation).
    """
    src = recs["src"].astype(np.int64)
    dst = recs["dst"].astype(np.int64)
    j = np.arange(recs.shape[0])
    own, key = (src, dst) if pull else (dst, src)
    q = sub_partition(own, partition_offset, local_offsets)
    if P > 1:
        perm = np.lexsort((j, own, ranks, key, q))
    else:
        perm = np.lexsort((j, key, q))
    return perm, q[perm]


def compressed_index(dsts, V):
    """[(first position, dst) for each distinct dst in a dst-sorted run] + (n, V)."""
    n = dsts.shape[0]
    if n == 0:
        return np.array([0], np.uint64), np.array([V], np.uint32)
    starts = np.flatnonzero(np.concatenate([[True], dsts[1:] != dsts[:-1]]))
    return (np.concatenate([starts, [n]]).astype(np.uint64),
            np.concatenate([dsts[starts], [V]]).astype(np.uint32))


def cpbt_bruteforce(a, k):
    """Minimal bottleneck over all splits of a into k contiguous groups (DP)."""
    a = [int(x) for x in a]
    n = len(a)
    pre = [0]
    for x in a:
        pre.append(pre[-1] + x)
    INF = float("inf")
    best = [[INF] * (n + 1) for _ in range(k + 1)]
    best[0][0] = 0
    for g in range(1, k + 1):
        for i in range(n + 1):
            for j in range(i + 1):
                c = max(best[g - 1][j], pre[i] - pre[j])
                if c < best[g][i]:
                    best[g][i] = c
    return best[k][n], pre


def bfs_queue(V, src, dst, root):
    """Textbook queue BFS depths (uint32 max = unreached)."""
    order = np.argsort(src, kind="stable")
    s, d = src[order], dst[order]
    start = np.searchsorted(s, np.arange(V + 1))
    depth = np.full(V, np.iinfo(np.uint32).max, dtype=np.uint32)
    depth[root] = 0
    q = [root]
    h = 0
    while h < len(q):
        u = q[h]
        h += 1
        for v in d[start[u]:start[u + 1]]:
            if depth[v] == np.iinfo(np.uint32).max:
                depth[v] = depth[u] + 1
                q.append(int(v))
    return depth


def min_id_parents(V, src, dst, depth, root):
    """parent[v] = min{u : u->v, depth[u] = depth[v]-1}; V if unreached; root->root."""
    UNR = np.iinfo(np.uint32).max
    parent = np.full(V, V, dtype=np.int64)
    ok = (depth[src] != UNR) & (depth[dst] != UNR) & (depth[src].astype(np.int64) + 1 == depth[dst])
    np.minimum.at(parent, dst[ok], src[ok])
    parent[root] = root
    return parent.astype(np.uint32)


def dijkstra_f32(V, src, dst, w, root):
    """Dijkstra with float32 left-fold path sums; 1e9 for unreachable."""
    order = np.argsort(src, kind="stable")
    s, d, ww = src[order], dst[order], w[order].astype(np.float32)
    start = np.searchsorted(s, np.arange(V + 1))
    dist = np.full(V, np.float32(1e9), dtype=np.float32)
    dist[root] = np.float32(0)
    done = np.zeros(V, bool)
    heap = [(0.0, root)]
    while heap:
        du, u = heapq.heappop(heap)
        if done[u]:
            continue
        done[u] = True
        base = dist[u]
        for k in range(start[u], start[u + 1]):
            v = d[k]
            nd = np.float32(base + ww[k])
            if nd < dist[v]:
                dist[v] = nd
                heapq.heappush(heap, (float(nd), int(v)))
    return dist


def pagerank_dense(V, src, dst, iterations, d=0.85, with_deltas=False):
    """Dense FP64 power iteration of toolkits/pagerank.cpp's update.  With
    with_deltas, also the printed delta(i) of every iteration: 1 before the
    first (the init returns 1 per vertex), then sum over vertices of
    |x' - x| * deg (|x' - x| where deg = 0) divided by V, and 0 after the
    last, undivided iteration (its process_vertices returns 0)."""
    deg = np.bincount(src, minlength=V).astype(np.float64)
    A = np.zeros((V, V))
    np.add.at(A, (dst, src), 1.0)  # A[v, u] = multiplicity of u->v
    x = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 1.0)
    deltas = [1.0]
    for it in range(iterations):
        y = (1 - d) + d * (A @ x)
        if it == iterations - 1:
            x = y
            deltas.append(0.0)
        else:
            xn = np.where(deg > 0, y / np.maximum(deg, 1), y)
            deltas.append(float((np.abs(xn - x) * np.where(deg > 0, deg, 1.0)).sum() / V))
            x = xn
    return (x, np.array(deltas)) if with_deltas else x


def widest_paths(V, src, dst, w, root):
    """Maximum-bottleneck widths from root by synchronous relaxation to the
    fixpoint: width[v] = max(width[v], max over edges u->v of min(width[u],
    w)); root 1e9, unreached 0 (float32, min/max only: exact)."""
    width = np.zeros(V, np.float32)
    width[root] = np.float32(1e9)
    w = w.astype(np.float32)
    while True:
        new = width.copy()
        np.maximum.at(new, dst, np.minimum(width[src], w))
        if np.array_equal(new, width):
            return width
        width = new
