"""The vertex order of the generated sphere mesh, pinned by properties of the
reference's ordering rather than by a second restatement of it.

utils/mesh_builder.cpp:316-402 (Graph::rcm, applied by reorder_mesh at
:553-585) numbers the inner points by a breadth-first search over the inner
adjacency graph, from point 0: a point is numbered when it leaves the FIFO
queue, and each numbered point appends its not-yet-numbered neighbours in
ascending order of their inner degree.  The outer points follow the inner
ones (full_order, :534-549).  Hence, on the inner graph of the final mesh:

  1. numbers never decrease with the breadth-first distance from point 0;
  2. each point's lowest-numbered neighbour (its parent) is one step closer
     to point 0, and parents are non-decreasing along the numbering (the
     queue is first in, first out);
  3. points with the same parent are numbered in ascending inner degree.

A depth-first order, the reversed (Cuthill-McKee reversed) order or a
descending degree sort each break one of these; the planted variants below
check that the test notices.  The zone-rule quirk of utils/icosahedron.cpp:
387-389 (the v2 test) is reproduced as read and is pinned by reading only.
"""
from collections import deque

import numpy as np
import pytest

import mke_amd as P


def inner_graph(tri, n_inner):
    adj = [set() for _ in range(n_inner)]
    for a, b, c in tri:
        for u, v in ((a, b), (b, c), (c, a)):
            if u < n_inner and v < n_inner:
                adj[u].add(v)
                adj[v].add(u)
    return adj


def order_violations(adj):
    """Which of the three properties the numbering 0..n-1 of graph adj
    breaks (empty list: none)."""
    n = len(adj)
    dist = [-1] * n
    dist[0] = 0
    q = deque([0])
    while q:
        v = q.popleft()
        for w in adj[v]:
            if dist[w] < 0:
                dist[w] = dist[v] + 1
                q.append(w)
    bad = []
    if any(d < 0 for d in dist):
        bad.append("disconnected")
    if any(dist[i + 1] < dist[i] for i in range(n - 1)):
        bad.append("distance decreases")
    parent = [min(adj[u]) if u else -1 for u in range(n)]
    if any(dist[parent[u]] != dist[u] - 1 for u in range(1, n)):
        bad.append("parent not one step closer")
    if any(parent[u + 1] < parent[u] for u in range(1, n - 1)):
        bad.append("parents out of order")
    deg = [len(a) for a in adj]
    if any(parent[u + 1] == parent[u] and deg[u + 1] < deg[u] for u in range(1, n - 1)):
        bad.append("siblings not by ascending degree")
    return bad


def relabel(adj, order):
    """adj renumbered so that old point order[k] becomes k."""
    new = {old: k for k, old in enumerate(order)}
    return [set(new[w] for w in adj[old]) for old in order]


@pytest.mark.parametrize("level", [2, 3, 4, 5])
def test_generated_order_is_the_reference_bfs(level):
    m = P.Mesh.sphere(level)
    a = m.arrays()
    n = m.inner_size
    assert list(a["outer"]) == list(range(n, m.n_points))   # outer points last
    adj = inner_graph(a["tri"], n)
    deg = sorted(set(len(s) for s in adj))
    assert deg == [5, 6]
    assert order_violations(adj) == []


def test_planted_orders_are_caught():
    m = P.Mesh.sphere(3)
    a = m.arrays()
    adj = inner_graph(a["tri"], m.inner_size)
    n = len(adj)
    # reversed numbering
    assert order_violations(relabel(adj, list(range(n))[::-1]))
    # depth-first numbering with the same neighbour order
    seen, order, stack = set(), [], [0]
    while stack:
        v = stack.pop()
        if v in seen:
            continue
        seen.add(v)
        order.append(v)
        stack.extend(sorted(adj[v], key=lambda w: (len(adj[w]), w), reverse=True))
    assert order_violations(relabel(adj, order))
    # breadth-first with neighbours in descending degree
    num = [-1] * n
    q, k = deque([0]), 0
    while q:
        v = q.popleft()
        if num[v] >= 0:
            continue
        num[v] = k
        k += 1
        q.extend(w for w in sorted(adj[v], key=lambda w: (-len(adj[w]), w)) if num[w] < 0)
    order = sorted(range(n), key=lambda v: num[v])
    assert order_violations(relabel(adj, order)) == ["siblings not by ascending degree"]
