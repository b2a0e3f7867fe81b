"""Shortest-path retrieval (h2h_shortest_paths): every returned path starts
at s, ends at t, follows graph edges, never repeats a vertex, and its length
is the Dijkstra distance -- on road-like graphs and on a graph with
zero-weight edges (where the walk may have to step back)."""
import numpy as np
import pytest

import cases
import h2h
import h2h_gen as G

pytestmark = pytest.mark.gpu


def _check_paths(case, paths, s, t):
    wt = {}
    for a, b, w in zip(case.u.tolist(), case.v.tolist(), case.w.tolist()):
        wt[(a, b)] = w
    D = case.dijkstra_rows(np.unique(s))
    row = {v: i for i, v in enumerate(np.unique(s).tolist())}
    for P, a, b in zip(paths, s.tolist(), t.tolist()):
        assert P[0] == a and P[-1] == b
        assert len(set(P.tolist())) == len(P)
        length = sum(wt[(int(x), int(y))] for x, y in zip(P[:-1], P[1:]))
        assert length == D[row[a], b]


def _index(case, kind=h2h.L4):
    p = h2h.H2HIndex()
    case.product(p)
    p.contract(1, h2h.CONTRACT_GPU_ORDER, 8)
    p.build_labels(kind, h2h.BATCHED)
    return p


@pytest.mark.parametrize("name,kind", [("grid64", h2h.L4), ("delaunay", h2h.L4), ("delaunay", h2h.L3)])
def test_paths_are_shortest(name, kind):
    case = cases.grid64() if name == "grid64" else cases.delaunay(30000, 2.6, 5, 7)
    rng = np.random.default_rng(3)
    s = rng.integers(0, case.n, 400).astype(np.int32)
    t = rng.integers(0, case.n, 400).astype(np.int32)
    s[:3] = t[:3]   # s == t gives [s]
    with _index(case, kind) as p:
        paths, _ = p.shortest_paths(s, t)
        assert all(len(P) == 1 for P in paths[:3])
        _check_paths(case, paths, s, t)


def test_paths_with_zero_weight_edges():
    g = G.grid_graph(24, 24, drop=0.1, seed=9)
    u, v, w = g.arcs()
    w = w.copy()
    rng = np.random.default_rng(5)
    # about a third of the edges get weight 0, the same in both directions
    key = np.minimum(u, v).astype(np.int64) * g.n + np.maximum(u, v)
    zero = np.isin(key, rng.choice(np.unique(key), len(np.unique(key)) // 3, replace=False))
    w[zero] = 0
    case = cases.Case("zero", g.n, u, v, w, *G.bisect(g.xy, 4), 4)
    s = rng.integers(0, case.n, 300).astype(np.int32)
    t = rng.integers(0, case.n, 300).astype(np.int32)
    with _index(case) as p:
        paths, _ = p.shortest_paths(s, t)
        _check_paths(case, paths, s, t)


def test_paths_need_the_graph(tmp_path):
    case = cases.grid64()
    with _index(case) as p:
        p.save(str(tmp_path / "g.h2h"), with_labels=False)
    with h2h.H2HIndex() as q:
        q.load(str(tmp_path / "g.h2h"))
        with pytest.raises(h2h.H2HError) as e:
            q.shortest_paths(np.array([0], np.int32), np.array([5], np.int32))
        assert e.value.status == -7
