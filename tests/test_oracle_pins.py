"""Pins the CPU oracle (oracle/h2h_oracle.cpp) to what the mathematics fixes,
with algorithms independent of the oracle's own code:

* L4 labels are shortest-path distances to the ancestors (scipy Dijkstra);
* L3 labels are shortest upward paths: Dijkstra on the bag DAG restricted to
  the nodes at or below the target's depth;
* both query kinds equal Dijkstra (UsingLCA with L4; noLCA with L3 and L4);
* the LCA equals an ancestor walk;
* the partition-order record equals a brute-force interior-set scan of the
  Cut ranges; the tree invariants hold (bag within ancestors, father = the
  min-rank bag member, Euler length 2n - 1);
* the two restatements of the GPU partition contraction (literal bubble
  sorts, heap order) agree;
* the serial, per-node and per-entry label rules give identical labels.

A plausible oracle mistake (a dropped L4 term, an off-by-one position, a
wrong father) breaks at least one of these.  CPU only."""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import dijkstra

import cases
import oracle as O


def _ref13(method):
    H = cases.REF13_METHODS[method]
    return O.OracleIndex.from_files(*cases.ref13_paths(method), H), H


def _ref13_case(method):
    """The 13-vertex fixture as a Case (0-based ids), via the oracle's loader."""
    edge, idf, cut = cases.ref13_paths(method)
    rows = [list(map(int, l.split()[1:])) for l in open(edge).read().split("\n")[1:] if l.strip()]
    u = np.array([r[0] - 1 for r in rows], np.int32)   # "a u v w", 1-based
    v = np.array([r[1] - 1 for r in rows], np.int32)
    w = np.array([r[2] for r in rows], np.int32)
    return u, v, w


def _apsp(n, u, v, w):
    A = csr_matrix((w.astype(np.float64), (u, v)), shape=(n, n))
    return dijkstra(A, directed=True).astype(np.int64)


def _ancestors(father, x):
    out = [x]
    while father[out[-1]] >= 0:
        out.append(int(father[out[-1]]))
    return out[::-1]  # root first: index j = depth - 1


def _small_cases():
    out = []
    for m in sorted(cases.REF13_METHODS):
        u, v, w = _ref13_case(m)
        for C in range(1, cases.REF13_METHODS[m] + 1):
            out.append(("ref13-" + m, C, lambda m=m: _ref13(m)[0], (13, u, v, w)))
    for seed in range(8):
        c = cases.small_random(seed)
        for C in sorted({1, c.H}):
            out.append((c.name, C, lambda c=c: c.oracle(2), (c.n, c.u, c.v, c.w)))
    return out


SMALL = _small_cases()


@pytest.mark.parametrize("name,C,make,graph", SMALL, ids=[f"{s[0]}-C{s[1]}" for s in SMALL])
def test_l4_labels_are_ancestor_distances(name, C, make, graph):
    n, u, v, w = graph
    D = _apsp(n, u, v, w)
    for mode in (O.CPU, O.GPU_FAST):
        o = make()
        o.contract(mode, C)
        o.labels(O.SERIAL, l4=True)
        fa, dh, dis = o.father(), o.dis_hash(), o.dis()
        for x in range(n):
            anc = _ancestors(fa, x)
            row = dis[dh[x]:dh[x + 1]]
            assert len(row) == len(anc)
            np.testing.assert_array_equal(row, D[x, anc], err_msg=f"{name} C={C} vertex {x}")


@pytest.mark.parametrize("name,C,make,graph", SMALL[::2], ids=[f"{s[0]}-C{s[1]}" for s in SMALL[::2]])
def test_l3_labels_are_upward_paths(name, C, make, graph):
    n = graph[0]
    o = make()
    o.contract(O.GPU_FAST, C)
    o.labels(O.SERIAL, l4=False)
    fa, dep, dh, dis = o.father(), o.depth(), o.dis_hash(), o.dis()
    off, nbr, bw = o.bags()
    src = np.repeat(np.arange(n), np.diff(off))
    for a in range(n):
        # reversed bag edges among the nodes at depth >= depth(a): a path from a
        # down to x is an upward bag path from x to a through those nodes
        keep = (dep[src] >= dep[a]) & (dep[nbr] >= dep[a])
        A = csr_matrix((bw[keep].astype(np.float64), (nbr[keep], src[keep])), shape=(n, n))
        d = dijkstra(A, directed=True, indices=a)
        for x in range(n):
            anc = _ancestors(fa, x)
            if len(anc) >= dep[a] and anc[dep[a] - 1] == a:
                assert dis[dh[x] + dep[a] - 1] == d[x], f"{name} C={C} x={x} a={a}"


@pytest.mark.parametrize("name,C,make,graph", SMALL, ids=[f"{s[0]}-C{s[1]}" for s in SMALL])
def test_queries_and_lca(name, C, make, graph):
    n, u, v, w = graph
    D = _apsp(n, u, v, w)
    xs, ys = np.meshgrid(np.arange(n), np.arange(n))
    x = xs.ravel().astype(np.int32)
    y = ys.ravel().astype(np.int32)
    o = make()
    o.contract(O.GPU_FAST, C)
    fa = o.father()
    L = O.lib()
    for a, b in zip(x[:: max(1, len(x) // 500)], y[:: max(1, len(x) // 500)]):
        pa, pb = _ancestors(fa, int(a)), _ancestors(fa, int(b))
        k = 0
        while k < min(len(pa), len(pb)) and pa[k] == pb[k]:
            k += 1
        assert L.orc_lca(o.h, int(a), int(b)) == pa[k - 1]
    o.labels(O.PF, l4=True, threads=2)
    for kind in (O.Q_LCA, O.Q_NOLCA):
        got, _ = o.query(kind, x, y, threads=2)
        np.testing.assert_array_equal(got, D[x, y], err_msg=f"kind {kind}")
    o.labels(O.PF, l4=False, threads=2)
    got, _ = o.query(O.Q_NOLCA, x, y, threads=2)
    np.testing.assert_array_equal(got, D[x, y])


def test_grid64_sampled_against_dijkstra():
    c = cases.grid64()
    o = c.oracle(4)
    o.contract(O.CPU, 1)
    o.labels(O.PF, l4=True, threads=4)
    rng = np.random.default_rng(3)
    src = rng.integers(0, c.n, 64)
    D = c.dijkstra_rows(src)
    fa, dh, dis = o.father(), o.dis_hash(), o.dis()
    for i, s in enumerate(src):
        anc = _ancestors(fa, int(s))
        A = c.dijkstra_rows(anc)            # distances from every ancestor
        np.testing.assert_array_equal(dis[dh[s]:dh[s + 1]], A[:, s])
        got, _ = o.query(O.Q_LCA, np.full(c.n, s, np.int32), np.arange(c.n, dtype=np.int32), threads=4)
        np.testing.assert_array_equal(got, D[i])


def _interior_record(n, adj, ids, cut, H):
    """Partition-order record by brute force: for levels H..1 and heap nodes
    from the highest index down, the not-yet-taken vertices of the range
    whose every neighbour is taken or inside the range, ascending id."""
    taken = np.zeros(n, bool)
    rec = []
    for h in range(H, 0, -1):
        for p in range((1 << h) - 2, (1 << (h - 1)) - 2, -1):
            lo, hi = cut[p]
            inside = set(int(x) for x in ids[lo:hi])
            mine = sorted(x for x in inside if not taken[x] and all(taken[y] or y in inside for y in adj[x]))
            for x in mine:
                taken[x] = True
            rec.extend(mine)
    return np.array(rec, np.int32), taken


@pytest.mark.parametrize("seed", range(6))
def test_partition_record_is_interior_scan(seed):
    c = cases.small_random(seed)
    adj = [[] for _ in range(c.n)]
    for a, b in zip(c.u, c.v):
        adj[a].append(int(b))
    rec, taken = _interior_record(c.n, adj, c.ids, c.cut, c.H)
    assert taken.all() and len(rec) == c.n
    _, elim, frac = c.oracle().partition()
    np.testing.assert_array_equal(elim, rec)
    assert abs(frac[-1] - 1.0) < 1e-12 if frac[0] <= frac[-1] else abs(frac[0] - 1.0) < 1e-12


@pytest.mark.parametrize("seed", range(10))
def test_tree_invariants(seed):
    c = cases.small_random(seed)
    o = c.oracle()
    o.contract(O.GPU_FAST, 1)
    n = c.n
    fa, dep, elim = o.father(), o.depth(), o.elimination_order()
    rank = np.empty(n, np.int64)
    rank[elim] = np.arange(n)
    off, nbr, _ = o.bags()
    roots = np.nonzero(fa < 0)[0]
    assert len(roots) == 1 and dep[roots[0]] == 1
    for x in range(n):
        bag = nbr[off[x]:off[x + 1]]
        anc = set(_ancestors(fa, x)[:-1])
        assert set(bag.tolist()) <= anc, f"bag of {x} not within its ancestors"
        if fa[x] >= 0:
            assert fa[x] == bag[np.argmin(rank[bag])]
            assert dep[x] == dep[fa[x]] + 1
    eu = o.euler()
    assert len(eu) == 2 * n - 1
    first = o.first_appear()
    np.testing.assert_array_equal(eu[first], np.arange(n))


@pytest.mark.parametrize("seed", range(12))
def test_gpu_contraction_restatements_agree(seed):
    c = cases.small_random(seed)
    for C in sorted({1, c.H}):
        a = c.oracle()
        a.contract(O.GPU_LITERAL, C)
        b = c.oracle()
        b.contract(O.GPU_FAST, C)
        np.testing.assert_array_equal(a.elimination_order(), b.elimination_order())
        np.testing.assert_array_equal(a.recorded_height(), b.recorded_height())
        np.testing.assert_array_equal(a.father(), b.father())


@pytest.mark.parametrize("seed", range(8))
def test_label_rules_agree(seed):
    """Serial (deepest entry first), per-node and per-entry rules give the same
    labels for both contraction orders.  The two orders themselves give
    different trees on most graphs: the GPU kernel's first pick of a
    partition uses the stale key and its slot lists keep holes, so the
    restricted orders differ; each order is reproduced on its own."""
    c = cases.small_random(seed)
    for mode in (O.CPU, O.GPU_FAST):
        o = c.oracle(3)
        o.contract(mode, 1)
        for l4 in (False, True):
            labs = []
            for method in (O.SERIAL, O.PF, O.LL):
                o.labels(method, l4=l4, threads=3)
                labs.append(o.dis())
            np.testing.assert_array_equal(labs[0], labs[1])
            np.testing.assert_array_equal(labs[0], labs[2])
