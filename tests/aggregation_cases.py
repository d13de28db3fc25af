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

Every off-diagonal entry is -1 (one case has a weak coupling) and every
diagonal the same, so every other edge is strong in both directions.  Each
case lists the aggregates as the vertices in insertion order.
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


def expected_map(n, aggregates):
    agg = np.full(n, 0xFFFFFFFE, np.uint32)
    for a, members in enumerate(aggregates):
        agg[members] = a
    return agg
