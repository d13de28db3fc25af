"""Known answers of the greedy aggregation (aggregation/base.hpp:448-707) on
small graphs, traced by hand from its rules:

- grow: while the open aggregate holds fewer than 6 members, add every front
  vertex that ranks first by (two-way couplings into the aggregate, or
  one-way ones while no front vertex has a two-way one; the share of its
  neighbours in aggregates connected to the open one; its neighbours in the
  front), ties taken together in front order up to the room left of 9;
- round: while below 9 members, add the first front vertex with a strong
  dependency into the aggregate and fewer unaggregated neighbours than
  neighbours inside it;
- a single-vertex aggregate merges into the aggregate of its first strong
  neighbour with room, or else stays on its own.

In CASES every off-diagonal entry is -1 (one case has a weak coupling) and
every diagonal the same, so every other edge is strong in both directions.
NONSYMMETRIC_CASES use matrices with a_ij != a_ji, where the strength of
i -> j is w_ij w_ji / a_jj with w = -a (dependency/criterion.hpp:42-60), so
that an edge can be strong for one end only: "one-way" (i depends on j, j
not on i; graph/properties.hpp:33-41) or "influences only" (j depends on i,
i not on j).  Each case lists the aggregates as the vertices in insertion
order.
"""

import numpy as np


def laplacian(n, edges, diag):
    """CSR rows in column order, diagonal inline, of the graph Laplacian-like
    matrix with -1 on the given undirected edges and `diag` on the diagonal."""
    nbrs = [set() for _ in range(n)]
    for a, b in edges:
        nbrs[a].add(b)
        nbrs[b].add(a)
    rp, cols, vals = [0], [], []
    for i in range(n):
        for j in sorted(nbrs[i] | {i}):
            cols.append(j)
            vals.append(diag if j == i else -1.0)
        rp.append(len(cols))
    return np.array(rp, np.int64), np.array(cols, np.int64), np.array(vals, np.float64)


def chain(n):
    return laplacian(n, [(i, i + 1) for i in range(n - 1)], 2.0)


def weak_chain(n, k):
    """chain(n) with the coupling between k and k + 1 set to -0.01."""
    rp, cols, vals = chain(n)
    for i, j in ((k, k + 1), (k + 1, k)):
        vals[rp[i] + np.flatnonzero(cols[rp[i]:rp[i + 1]] == j)[0]] = -0.01
    return rp, cols, vals


def grid_edges(nx, ny):
    """Vertex i*ny + j; 5-point couplings."""
    e = [(i * ny + j, i * ny + j + 1) for i in range(nx) for j in range(ny - 1)]
    return e + [(i * ny + j, (i + 1) * ny + j) for i in range(nx - 1) for j in range(ny)]


def grid(nx, ny):
    return laplacian(nx * ny, grid_edges(nx, ny), 4.0)


# (name, csr builder, aggregates in insertion order,
#  (multi-vertex, single-vertex, merged) counts of the aggregation statistics)
CASES = [
    # six in a row by grow; the round rule then takes the chain's end 6 (no
    # unaggregated neighbour, one inside)
    ("chain7", lambda: chain(7), [[0, 1, 2, 3, 4, 5, 6]], (1, 0, 0)),
    # the round rule rejects 6 (one neighbour inside, one unaggregated); the
    # second aggregate stops growing when its front runs empty
    ("chain10", lambda: chain(10), [[0, 1, 2, 3, 4, 5], [6, 7, 8, 9]], (2, 0, 0)),
    ("chain20", lambda: chain(20),
     [[0, 1, 2, 3, 4, 5], [6, 7, 8, 9, 10, 11], [12, 13, 14, 15, 16, 17], [18, 19]], (4, 0, 0)),
    # 4x4 grid: 1 and 4 tie and enter together; 5 wins by two couplings;
    # 2 and 8 tie on all three keys; in the second aggregate 6 wins on its
    # connected share (3/4), 11 beats 10 on it (1/3 > 1/4), then 10 wins by
    # two couplings and 9 on its share; the round rule rejects every vertex
    ("grid4x4", lambda: grid(4, 4),
     [[0, 1, 4, 5, 2, 8], [3, 7, 6, 11, 10, 9], [12, 13, 14, 15]], (3, 0, 0)),
    # 3x3 grid with a pendant vertex 9 on 8: grow ends at six, the round rule
    # fills the aggregate to nine (5, 7, 8); 9 then finds its only neighbour's
    # aggregate full and stays a single-vertex aggregate
    ("grid3x3_pendant", lambda: laplacian(10, grid_edges(3, 3) + [(8, 9)], 4.0),
     [[0, 1, 3, 4, 2, 6, 5, 7, 8], [9]], (1, 1, 0)),
    # chain of 13 whose coupling 6-7 is -0.01 (weak both ways: 0.01^2 is
    # below a third of the row maximum 1): the round rule rejects 6 (7 is
    # unaggregated); 6 then starts an aggregate, its front vertex 7 has no
    # strong coupling into it, so 6 merges into its strong neighbour 5's
    # aggregate; 7..12 grow as a chain
    ("chain13_weak", lambda: weak_chain(13, 6), [[0, 1, 2, 3, 4, 5, 6], [7, 8, 9, 10, 11, 12]],
     (2, 0, 1)),
]


def directed(n, weights, diag):
    """CSR rows in column order, diagonal inline, of the nonsymmetric matrix
    with a_ij = -w for every (i, j): w in `weights` and `diag` on the
    diagonal (a scalar or one value per row)."""
    d = np.broadcast_to(np.asarray(diag, np.float64), (n,))
    rows = [{i: float(d[i])} for i in range(n)]
    for (i, j), w in weights.items():
        rows[i][j] = -float(w)
    rp, cols, vals = [0], [], []
    for i in range(n):
        for j in sorted(rows[i]):
            cols.append(j)
            vals.append(rows[i][j])
        rp.append(len(cols))
    return np.array(rp, np.int64), np.array(cols, np.int64), np.array(vals, np.float64)


def pairs(edges):
    """{(i, j): w_ij, (j, i): w_ji} from (i, j, w_ij, w_ji) tuples."""
    out = {}
    for i, j, wij, wji in edges:
        out[(i, j)] = wij
        out[(j, i)] = wji
    return out


# A 7-vertex matrix whose strength graph is traced by hand below (theta 1/3):
#   row 0: a00 = 1,    a01 = -2,  a03 = -1      (3 lists no entry for 0)
#   row 1: a10 = -0.4, a11 = 2,   a12 = -3
#   row 2: a21 = -1,   a22 = 2,   a24 = +1      (w_24 = -1 < 0)
#   row 3: a33 = 1,    a34 = -1,  a36 = -8
#   row 4: a42 = -2,   a43 = -4,  a44 = 5
#   row 5: a50 = -1,   a55 = 3                  (0 lists no entry for 5)
#   row 6: a63 = -1,   a66 = 1
# Strengths s_ij = max(w_ij, 0) max(w_ji, 0) / a_jj:
#   0->1: 2*0.4/2 = 0.4   0->3: 1*0/1 = 0 (no reverse entry)
#   1->0: 0.4*2/1 = 0.8   1->2: 3*1/2 = 1.5
#   2->1: 1*3/2 = 1.5     2->4: 0 (w_24 < 0)
#   3->4: 1*4/5 = 0.8     3->6: 8*1/1 = 8
#   4->2: 2*0/2 = 0       4->3: 4*1/1 = 4
#   5->0: 1*0/1 = 0       6->3: 1*8/1 = 8
# Row thresholds max/3: 0.1333, 0.5, 0.5, 2.667, 1.333, 0, 2.667, so the
# strong (depends) edges are 0->1, 1->0, 1->2, 2->1, 3->6, 4->3, 6->3, and
# vertex 5 (all strengths 0, none above 0) is isolated.  influences(j->i) is
# depends(i->j).  Flags per off-diagonal entry in row order (bit 0 depends,
# bit 1 influences):
#   0: [0->1 3, 0->3 0]  1: [1->0 3, 1->2 3]  2: [2->1 3, 2->4 0]
#   3: [3->4 2, 3->6 3]  4: [4->2 0, 4->3 1]  5: [5->0 0]  6: [6->3 3]
# 4->3 is one-way; 3->4 is "influences only".  Taking w_ij^2 for w_ij w_ji
# makes 0->3 strong; the row's own diagonal for a_jj makes 1->0 weak;
# w_ji^2 makes 1->2 weak.
STRENGTH_CASE = (
    lambda: directed(7, {(0, 1): 2, (0, 3): 1, (1, 0): 0.4, (1, 2): 3, (2, 1): 1, (2, 4): -1,
                         (3, 4): 1, (3, 6): 8, (4, 2): 2, (4, 3): 4, (5, 0): 1, (6, 3): 1},
                     [1.0, 2.0, 2.0, 1.0, 5.0, 3.0, 1.0]),
    np.array([3, 0, 3, 3, 3, 0, 2, 3, 0, 1, 0, 3], np.uint8),
    np.array([0, 0, 0, 0, 0, 1, 0], np.uint8),
)

# Nonsymmetric aggregation cases.  All diagonals are equal, so s_ij is
# proportional to the product p = w_ij w_ji of the two weights of an edge
# and a vertex depends on a neighbour iff that edge's product exceeds a
# third of the largest product at the vertex.  The two weights of each
# edge differ, so the products are not the squares of either weight.
NONSYMMETRIC_CASES = [
    # chain 0-..-7 with products 9, 1, 1, 1, 1, 1, 1: vertex 1 depends on 0
    # only (1 < 9/3), so 2->1 is one-way and 1->2 influences only.  grow from
    # 0 adds 1 (two-way), then 2 on its one-way count alone (no front vertex
    # has a two-way coupling into the aggregate), then 3, 4, 5 (two-way); the
    # round rule rejects 6 (one neighbour inside, one unaggregated); 6 and 7
    # grow into the second aggregate.  Ignoring one-way couplings stops the
    # first aggregate at [0, 1].
    ("oneway_chain8",
     lambda: directed(8, pairs([(0, 1, 1, 9), (1, 2, 0.5, 2), (2, 3, 2, 0.5), (3, 4, 1, 1),
                                (4, 5, 0.25, 4), (5, 6, 4, 0.25), (6, 7, 1, 1)]), 100.0),
     [[0, 1, 2, 3, 4, 5], [6, 7]], (2, 0, 0)),
    # hub 0 with neighbours 1 (product 1) and 2 (product 9), 1 also coupled
    # to 7 (product 1), 2-3-4-5-6 a chain of products 9: vertex 0 depends
    # on 2 only, vertex 1 on both 0 and 7, so 1->0 is one-way.  In every grow
    # step the front vertex 1 (one-way, listed first) is dropped as soon as
    # a two-way candidate appears (2, then 3, 4, 5, 6); the round rule
    # rejects 1 (7 unaggregated, 0 inside); 1 and 7 then form the second
    # aggregate.  Counting one-way couplings as two-way lets 1 tie with 2
    # and enter the first aggregate together with it.
    ("oneway_hub8",
     lambda: directed(8, pairs([(0, 1, 4, 0.25), (0, 2, 3, 3), (1, 7, 0.5, 2), (2, 3, 1, 9),
                                (3, 4, 9, 1), (4, 5, 3, 3), (5, 6, 4.5, 2)]), 100.0),
     [[0, 2, 3, 4, 5, 6], [1, 7]], (2, 0, 0)),
    # chain 0-..-8 with products 1, 1, 1, 1, 1, 1, 9, 81: 5 depends on 6 but
    # 6 not on 5 (1 < 9/3) and 7 not on 6 (9 < 81/3).  0..5 grow into the
    # first aggregate; the round rule rejects 6 (it does not depend on 5).
    # 6 starts an aggregate that cannot grow (7 has no coupling into it that
    # 7 depends on) and merges into 5's aggregate: 6->5 is strong through
    # "influences" alone (find_merge_aggregate takes is_strong, depends or
    # influences).  7 and 8 (two-way) form the second aggregate.  Taking
    # "depends" alone for the merge leaves 6 a single-vertex aggregate.
    ("influence_merge9",
     lambda: directed(9, pairs([(0, 1, 2, 0.5), (1, 2, 0.5, 2), (2, 3, 4, 0.25), (3, 4, 1, 1),
                                (4, 5, 0.25, 4), (5, 6, 2, 0.5), (6, 7, 9, 1),
                                (7, 8, 27, 3)]), 100.0),
     [[0, 1, 2, 3, 4, 5, 6], [7, 8]], (2, 0, 1)),
    # chain 0-..-5 of products 100, chain 6-..-11 of products 1, and vertex
    # 12 coupled to 5 (product 81), 10 and 11 (products 1): 12 depends on 5
    # only (1 < 81/3), 10 and 11 depend on 12, and 5 on 12 (81 > 100/3).
    # 0..5 grow into the first aggregate; the round rule rejects 12 (10 and
    # 11 still unaggregated).  6..11 grow into the second; the round rule
    # rejects 12 again: its couplings into the aggregate are "influences
    # only" and the rule counts only those it depends on (counting any strong
    # coupling would accept it: no unaggregated neighbour, two inside).  12
    # then merges into 5's aggregate through its strong coupling to 5.
    ("influence_round13",
     lambda: directed(13, pairs([(0, 1, 10, 10), (1, 2, 20, 5), (2, 3, 4, 25), (3, 4, 50, 2),
                                 (4, 5, 10, 10), (5, 12, 9, 9), (6, 7, 2, 0.5), (7, 8, 1, 1),
                                 (8, 9, 4, 0.25), (9, 10, 0.5, 2), (10, 11, 1, 1),
                                 (10, 12, 2, 0.5), (11, 12, 0.25, 4)]), 1000.0),
     [[0, 1, 2, 3, 4, 5, 12], [6, 7, 8, 9, 10, 11]], (2, 0, 1)),
]


def expected_map(n, aggregates):
    agg = np.full(n, 0xFFFFFFFE, np.uint32)
    for a, members in enumerate(aggregates):
        agg[members] = a
    return agg
