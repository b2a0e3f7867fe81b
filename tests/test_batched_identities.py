"""The identities the batched L4 build relies on, checked on the CPU against
the oracle's labels (computed by the reference's own recurrence) on small
random road-like graphs.  Each one is a rewriting that changes the order of
work, never the values, so a kernel built on a wrong identity would differ
from the oracle; these tests pin the identity itself, independently of any
kernel:

* the reverse-side rule: d(v, a) = min(U(v, a), min over u in X(a) of
  w(a, u) + d(v, u)), with U the upward (L3) distance;
* the chain triangle as one min-plus product of the chain closure P with
  the cross terms through the entries above the chain;
* dropping a bag entry that a two-hop path through another bag member
  beats strictly leaves every L4 label unchanged;
* the closure form of a subtree's outer columns, reading only the ancestors
  that have entries above the subtree;
* the level form of a subtree's outer columns: a node's own bound over its
  entries above the subtree, then its bag members inside the subtree.

CPU only."""
import numpy as np
import pytest

import cases
import oracle as O

INF = 1 << 40


def _index(seed):
    c = cases.small_random(seed)
    o = c.oracle(2)
    o.contract(O.GPU_FAST, 1)
    o.labels(O.SERIAL, l4=False)
    d3 = o.dis().copy()
    o.labels(O.SERIAL, l4=True)
    d4 = o.dis().copy()
    off, nbr, bw = o.bags()
    return c.n, o.father(), o.depth(), o.dis_hash(), d3, d4, off, nbr, bw


def _path_to_root(fa, v):
    p = [v]
    while fa[p[-1]] >= 0:
        p.append(int(fa[p[-1]]))
    return p[::-1]  # root first


@pytest.mark.parametrize("seed", range(6))
def test_reverse_side_rule(seed):
    n, fa, dep, dh, d3, d4, off, nbr, bw = _index(seed)
    for v in range(n):
        a = fa[v]
        while a >= 0:
            best = d3[dh[v] + dep[a] - 1]
            for e in range(off[a], off[a + 1]):
                best = min(best, bw[e] + d4[dh[v] + dep[nbr[e]] - 1])
            assert best == d4[dh[v] + dep[a] - 1], f"v={v} a={a}"
            a = fa[a]


def _segment_closure(ch, h, off, nbr, bw, dep):
    """Upward closure over the bag entries inside the segment ch (root first):
    P[i][m], m <= i."""
    L = len(ch)
    P = np.full((L, L), INF, np.int64)
    for i in range(L):
        P[i][i] = 0
        for m in range(i - 1, -1, -1):
            best = INF
            for e in range(off[ch[i]], off[ch[i] + 1]):
                k = dep[nbr[e]] - 1 - h
                if k >= m:
                    best = min(best, bw[e] + P[k][m])
            P[i][m] = best
    return P


@pytest.mark.parametrize("seed", range(6))
def test_chain_triangle_is_a_product(seed):
    n, fa, dep, dh, d3, d4, off, nbr, bw = _index(seed)
    rng = np.random.default_rng(seed)
    checked = 0
    for _ in range(40):
        path = _path_to_root(fa, int(rng.integers(n)))
        if len(path) < 3:
            continue
        a0 = int(rng.integers(0, len(path) - 1))
        ch = path[a0:a0 + int(rng.integers(2, len(path) - a0 + 1))]
        h, L = dep[ch[0]] - 1, len(ch)
        P = _segment_closure(ch, h, off, nbr, bw, dep)
        # t'_q(i): through q's entries above the segment, into row i
        tp = np.full((L, L), INF, np.int64)
        for q in range(L):
            for e in range(off[ch[q]], off[ch[q] + 1]):
                pu = dep[nbr[e]] - 1
                if pu < h:
                    for i in range(q + 1, L):
                        tp[i][q] = min(tp[i][q], bw[e] + d4[dh[ch[i]] + pu])
        for i in range(L):
            for m in range(i):
                T = min(P[m][q] + min(P[i][q], tp[i][q]) for q in range(m + 1))
                assert T == d4[dh[ch[i]] + h + m], f"segment {ch} i={i} m={m}"
                checked += 1
    assert checked > 0


@pytest.mark.parametrize("seed", range(6))
def test_two_hop_dominated_entries_are_redundant(seed):
    n, fa, dep, dh, d3, d4, off, nbr, bw = _index(seed)
    bag = [dict(zip(nbr[off[x]:off[x + 1]].tolist(), bw[off[x]:off[x + 1]].tolist())) for x in range(n)]
    dropped = 0
    for v in range(n):
        keep = []
        for u, wu in bag[v].items():
            # beaten strictly through another member k: w(v,k) + w(k,u) < w(v,u),
            # w(k,u) read from whichever of the two bags holds the pair
            dom = any(k != u and wk + bag[k].get(u, bag[u].get(k, INF)) < wu for k, wk in bag[v].items())
            dropped += dom
            if not dom:
                keep.append((u, wu))
        assert keep or not bag[v], f"every entry of {v} dropped"  # the root's bag is empty
        for j in range(dep[v] - 1):
            a_j = _path_to_root(fa, v)[j]
            best = INF
            for u, wu in keep:
                pu = dep[u] - 1
                best = min(best, wu + (d4[dh[u] + j] if j <= pu else d4[dh[a_j] + pu]))
            assert best == d4[dh[v] + j], f"v={v} j={j}"
    assert dropped > 0  # the rule removes something on these graphs


@pytest.mark.parametrize("seed", range(6))
def test_subtree_outer_columns_skip_ancestors_without_outer_entries(seed):
    n, fa, dep, dh, d3, d4, off, nbr, bw = _index(seed)
    kids = [[] for _ in range(n)]
    for x in range(n):
        if fa[x] >= 0:
            kids[fa[x]].append(x)
    for root in range(n):
        if fa[root] < 0 or dep[root] < 3:
            continue
        r = dep[root]
        sub, stack = [], [root]
        while stack:
            x = stack.pop()
            sub.append(x)
            stack.extend(kids[x])
        has_outer = {x: any(dep[nbr[e]] - 1 < r - 1 for e in range(off[x], off[x + 1])) for x in sub}
        assert has_outer[root]  # the father of the root is above the subtree
        for v in sub:
            path = _path_to_root(fa, v)
            anc = path[r - 1:]  # in-subtree ancestors, root first, v last
            # closure over in-subtree entries: upward distances from v to anc
            dist = {u: INF for u in anc}
            dist[v] = 0
            for x in reversed(anc):  # v first, then up: upward paths from v
                if dist[x] == INF:
                    continue
                for e in range(off[x], off[x + 1]):
                    u = nbr[e]
                    if u in dist and dep[u] < dep[x]:
                        dist[u] = min(dist[u], dist[x] + bw[e])
            for j in range(r - 1):
                best = INF
                for u in anc:
                    if not has_outer[u] or dist[u] == INF:
                        continue
                    # B_u(j): u's entries above the subtree
                    for e in range(off[u], off[u + 1]):
                        z = nbr[e]
                        pz = dep[z] - 1
                        if pz < r - 1:
                            g = d4[dh[z] + j] if j <= pz else d4[dh[path[j]] + pz]
                            best = min(best, dist[u] + bw[e] + g)
                assert best == d4[dh[v] + j], f"root={root} v={v} j={j}"
        break  # one subtree per graph keeps the test fast


@pytest.mark.parametrize("seed", range(6))
def test_subtree_outer_columns_level_form(seed):
    n, fa, dep, dh, d3, d4, off, nbr, bw = _index(seed)
    rng = np.random.default_rng(seed)
    checked = 0
    for v in rng.choice(n, min(n, 120), replace=False):
        path = _path_to_root(fa, int(v))  # root first, v last
        for root in path[1:]:             # every ancestor below the tree root as the subtree root
            R = dep[root]
            for j in range(R - 1):
                b = INF  # the bound: entries above the subtree
                inner = INF  # bag members inside the subtree (their rows hold the final values)
                for e in range(off[v], off[v + 1]):
                    u, w, pu = nbr[e], bw[e], dep[nbr[e]] - 1
                    if dep[u] < R:
                        b = min(b, w + (d4[dh[u] + j] if j <= pu else d4[dh[path[j]] + pu]))
                    else:
                        inner = min(inner, w + d4[dh[u] + j])
                assert min(b, inner) == d4[dh[v] + j], f"v={v} root={root} j={j}"
                checked += 1
    assert checked > 0
