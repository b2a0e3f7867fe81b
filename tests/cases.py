"""Shared graph cases for the tests: the reference's 13-vertex fixtures and
seeded synthetic graphs (both directions listed, simple, connected)."""
import os

import numpy as np

import h2h_gen as G
import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF13 = os.path.join(GOLDEN, "ref13")
REF13_METHODS = {"Latitude": 4, "Minimum": 4, "metis": 3, "scotch": 3}


def ref13_paths(method):
    return (os.path.join(REF13, "13Edge.txt"), os.path.join(REF13, f"{method}ID.txt"),
            os.path.join(REF13, f"{method}Cut.txt"))


class Case:
    def __init__(self, name, n, u, v, w, ids, cut, H):
        self.name, self.n, self.u, self.v, self.w = name, n, u, v, w
        self.ids, self.cut, self.H = ids, cut, H

    def oracle(self, threads=2):
        return O.OracleIndex(self.n, self.u, self.v, self.w, self.ids, self.cut, self.H, threads=threads)

    def product(self, idx):
        idx.set_graph(self.n, self.u, self.v, self.w, self.ids, self.cut, self.H)

    def dijkstra_rows(self, sources):
        from scipy.sparse import csr_matrix
        from scipy.sparse.csgraph import dijkstra
        A = csr_matrix((self.w.astype(np.float64), (self.u, self.v)), shape=(self.n, self.n))
        return dijkstra(A, directed=True, indices=np.asarray(sources)).astype(np.int64)


def from_graph(name, g, H, ids=None, cut=None):
    if ids is None:
        ids, cut = G.bisect(g.xy, H)
    u, v, w = g.arcs()
    return Case(name, g.n, u, v, w, ids, cut, H)


def grid64():
    g, ids, cut, H = G.make_workload("grid64")
    return from_graph("grid64", g, H, ids, cut)


def small_random(seed):
    rng = np.random.default_rng(1000 + seed)
    H = int(rng.integers(2, 7))
    if seed % 2 == 0:
        g = G.grid_graph(int(rng.integers(5, 20)), int(rng.integers(5, 20)), drop=float(rng.uniform(0, 0.3)),
                         seed=seed)
    else:
        g = G.delaunay_graph(int(rng.integers(40, 400)), float(rng.uniform(2.1, 3.0)), seed=seed)
    return from_graph(f"rand{seed}", g, H)


def delaunay(n, deg, seed, H):
    return from_graph(f"del{n}", G.delaunay_graph(n, deg, seed=seed), H)


def uneven(n, deg, seed, H, leaf_range):
    """A Delaunay graph whose bottom partitions have uneven sizes in leaf_range."""
    g = G.delaunay_graph(n, deg, seed=seed)
    ids, cut = G.bisect(g.xy, H, leaf_range, seed=seed)
    return from_graph(f"uneven{n}", g, H, ids, cut)
