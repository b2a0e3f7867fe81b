"""Index persistence (h2h_index_save / h2h_index_load; the reference keeps
its index in memory only): a saved and reloaded index -- with its label
rows, or without them and rebuilt on load -- answers every query the built
one answers, holds the same label rows, and rebuilds labels equal to the
oracle's from the saved bags; damaged files are refused."""
import numpy as np
import pytest

import cases
import h2h
import oracle as O

pytestmark = pytest.mark.gpu


def _rows(lab, depth):
    dh, dis = lab["dis_hash"], lab["dis"]
    return [dis[dh[v]:dh[v] + depth[v]] for v in range(len(depth))]


def _built(case, threads=8):
    o = case.oracle(threads)
    o.contract(O.GPU_FAST, 1)
    p = h2h.H2HIndex()
    case.product(p)
    p.contract(1, h2h.CONTRACT_GPU_ORDER, threads)
    return o, p


@pytest.mark.parametrize("kind,with_labels", [(h2h.L4, True), (h2h.L3, True), (h2h.L4, False), (h2h.L3, False)])
def test_save_load_roundtrip(tmp_path, kind, with_labels):
    case = cases.delaunay(30000, 2.6, 5, 7)
    o, p = _built(case)
    rng = np.random.default_rng(17)
    x = rng.integers(0, case.n, 200000).astype(np.int32)
    y = rng.integers(0, case.n, 200000).astype(np.int32)
    qkinds = (h2h.QUERY_LCA, h2h.QUERY_NOLCA) if kind == h2h.L4 else (h2h.QUERY_NOLCA,)
    path = str(tmp_path / "index.h2h")
    with p:
        p.build_labels(kind, h2h.BATCHED)
        depth = p.tree()["depth"]
        rows = _rows(p.labels(), depth)
        ans = {qk: p.query(x, y, qk)[0] for qk in qkinds}
        p.save(path, with_labels)
    with h2h.H2HIndex() as q:
        q.load(path)
        assert q.info()["label_entries"] == sum(len(r) for r in rows)
        np.testing.assert_array_equal(q.tree()["depth"], depth)
        for a, b in zip(_rows(q.labels(), depth), rows):
            np.testing.assert_array_equal(a, b)
        for qk in qkinds:
            for ex in (h2h.EXEC_PLAIN, h2h.EXEC_BATCHED):
                np.testing.assert_array_equal(q.query(x, y, qk, ex)[0], ans[qk])
        # labels rebuilt from the saved bags (and two-hop marks) equal the oracle's
        o.labels(O.SERIAL, l4=(kind == h2h.L4))
        q.build_labels(kind, h2h.BATCHED)
        odh, odis = o.dis_hash(), o.dis()
        for v, r in enumerate(_rows(q.labels(), depth)):
            np.testing.assert_array_equal(r, odis[odh[v]:odh[v] + depth[v]])


def test_damaged_files_are_refused(tmp_path):
    case = cases.grid64()
    _, p = _built(case, threads=4)
    good = tmp_path / "good.h2h"
    with p:
        with pytest.raises(h2h.H2HError) as e:
            p.save(str(good))                       # no labels yet
        assert e.value.status == -7
        p.build_labels(h2h.L4, h2h.BATCHED)
        p.save(str(good))
        p.save(str(tmp_path / "small.h2h"), with_labels=False)
        assert (tmp_path / "small.h2h").stat().st_size < good.stat().st_size / 2
        ref, _ = p.query(np.arange(case.n, dtype=np.int32), np.zeros(case.n, np.int32))
    data = bytearray(good.read_bytes())
    flipped = bytearray(data)
    flipped[len(data) // 2] ^= 0x40
    huge = bytearray(data)
    huge[24:32] = (1 << 61).to_bytes(8, "little")   # bag entries: 8 x this wraps a 64-bit size
    cases_bad = {"flip.h2h": bytes(flipped), "cut.h2h": bytes(data[:-100]), "magic.h2h": b"XXXX" + bytes(data[4:]),
                 "empty.h2h": b"", "huge.h2h": bytes(huge)}
    with h2h.H2HIndex() as q:
        for name, blob in cases_bad.items():
            (tmp_path / name).write_bytes(blob)
            with pytest.raises(h2h.H2HError) as e:
                q.load(str(tmp_path / name))
            assert e.value.status == -2, name
        with pytest.raises(h2h.H2HError) as e:
            q.load(str(tmp_path / "missing.h2h"))
        assert e.value.status == -2
        q.load(str(good))                           # a good file still loads after the refusals
        got, _ = q.query(np.arange(case.n, dtype=np.int32), np.zeros(case.n, np.int32))
        np.testing.assert_array_equal(got, ref)
