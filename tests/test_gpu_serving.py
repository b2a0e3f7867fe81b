"""Serving loop: continuous batching over the resident index.  Every
request's answer is the same alone and in the mix (bit-exact: queries are
independent reads of a read-only index)."""
import threading

import numpy as np
import pytest

import cases
import h2h

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def index():
    case = cases.delaunay(20000, 2.6, 9, 6)
    p = h2h.H2HIndex()
    case.product(p)
    p.contract(1, h2h.CONTRACT_GPU_ORDER, 4)
    p.build_labels(h2h.L4, h2h.PF)
    yield case, p
    p.close()


@pytest.mark.parametrize("exec_mode", [h2h.EXEC_PLAIN, h2h.EXEC_BATCHED])
def test_alone_equals_in_mix(index, exec_mode):
    case, p = index
    rng = np.random.default_rng(3)
    sizes = np.exp(rng.uniform(0, np.log(6000), 120)).astype(int) + 1
    reqs = [(rng.integers(0, case.n, s).astype(np.int32), rng.integers(0, case.n, s).astype(np.int32)) for s in sizes]
    alone = [p.query(x, y, h2h.QUERY_LCA, h2h.EXEC_PLAIN)[0] for x, y in reqs]
    p.server_start(streams=4, max_inflight=20000, max_request=8192, kind=h2h.QUERY_LCA, exec_mode=exec_mode)
    try:
        results = [None] * len(reqs)

        def client(k0):
            for k in range(k0, len(reqs), 4):
                t, out = p.submit(*reqs[k])
                st = p.wait(t, timeout_ms=60000)
                assert st["n"] == len(reqs[k][0])
                assert st["submit_us"] <= st["start_us"] <= st["done_us"]
                results[k] = out

        ths = [threading.Thread(target=client, args=(i,)) for i in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        for k in range(len(reqs)):
            np.testing.assert_array_equal(results[k], alone[k])
    finally:
        p.server_stop()


def test_admission_and_cancel(index):
    case, p = index
    ref, _ = p.query(np.arange(100, dtype=np.int32), np.zeros(100, np.int32))
    p.server_start(streams=2, max_inflight=4096, max_request=1024)
    try:
        x = np.zeros(2048, np.int32)
        with pytest.raises(h2h.H2HError) as e:
            p.submit(x, x)
        assert e.value.status == -9  # H2H_ERR_ADMIT
        with pytest.raises(h2h.H2HError) as e:
            p.submit(np.array([case.n], np.int32), np.array([0], np.int32))
        assert e.value.status == -1
        t, out = p.submit(np.arange(100, dtype=np.int32), np.arange(100, dtype=np.int32))
        p.wait(t)
        np.testing.assert_array_equal(out, 0)
        # nothing may replace the device index under the running server
        for call in (lambda: p.contract(1), lambda: p.build_labels(h2h.L4, h2h.BATCHED),
                     lambda: case.product(p), lambda: p.partition_rank_tree()):
            with pytest.raises(h2h.H2HError) as e:
                call()
            assert e.value.status == -7  # H2H_ERR_STATE
        t, out = p.submit(np.arange(100, dtype=np.int32), np.zeros(100, np.int32))
        p.wait(t)
        np.testing.assert_array_equal(out, ref)
    finally:
        p.server_stop()
