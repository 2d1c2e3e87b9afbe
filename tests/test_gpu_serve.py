"""Serving loop on the GPU: queries admitted while others run give the same answer
alone, in the mix, and on the plain one-query-per-launch path; admission respects
the HBM budget."""
import threading
import time

import pytest

from shadowvm_amd import pipeline as P
from shadowvm_amd import ssb

from .helpers import assert_same_columns

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def serve_tables(engine):
    wl = ssb.four_way(seed=31, fact_rows=8_000_000, n_cust=300_000, n_supp=20_000, n_part=140_000,
                      prefix="sv")
    ssb.load_device(wl)
    return wl


def _serve(queries, mode, workers=2, budget=0, stagger_s=0.0):
    P.server_start(workers=workers, budget=budget, max_batch=64, mode=mode)
    try:
        tickets = []
        for q in queries:
            tickets.append(P.server_submit(q.plans()))
            if stagger_s:
                time.sleep(stagger_s)
        return [P.server_wait(t, timeout_ms=120_000) for t in tickets]
    finally:
        P.server_stop()


def test_serving_same_alone_in_mix_and_plain(serve_tables):
    wl = serve_tables
    qs = ssb.serve_queries(wl, 30, seed=5, light_rows=4_000_000, prefix="sa")
    ssb.make_views(qs)
    print(f"[serve] {sum(q.shape['kind'] == 'heavy' for q in qs)} heavy of {len(qs)}", flush=True)
    mix = _serve(qs, mode=0, stagger_s=0.002)
    plain = _serve(qs, mode=1)
    alone = [_serve([q], mode=0)[0] for q in qs[:5]]
    for a, b in zip(mix, plain):
        assert a.rows == b.rows
        assert_same_columns(a.columns, b.columns)
    for a, b in zip(alone, mix[:5]):
        assert a.rows == b.rows
        assert_same_columns(a.columns, b.columns)


def test_admission_respects_budget(serve_tables):
    wl = serve_tables
    qs = ssb.serve_queries(wl, 12, seed=6, light_frac=1.0, light_rows=4_000_000, prefix="sb")
    ssb.make_views(qs)
    peak = [0]
    stop = threading.Event()
    budget = 8 << 20

    def watch():
        while not stop.is_set():
            try:
                peak[0] = max(peak[0], P.server_stats()["bytes_in_use"])
            except Exception:
                pass
            time.sleep(0.0005)

    P.server_start(workers=3, budget=budget, max_batch=64, mode=0)
    th = threading.Thread(target=watch)
    th.start()
    try:
        ts = [P.server_submit(q.plans()) for q in qs]
        res = [P.server_wait(t, timeout_ms=120_000) for t in ts]
    finally:
        stop.set()
        th.join()
        P.server_stop()
    assert all(r is not None for r in res)
    assert peak[0] <= budget


def test_query_larger_than_budget_is_refused(serve_tables):
    wl = serve_tables
    q = ssb.serve_queries(wl, 1, seed=7, light_frac=0.0, prefix="sc")[0]
    ssb.make_views([q])
    P.server_start(workers=1, budget=1 << 16, max_batch=8, mode=0)
    try:
        with pytest.raises(P._lib.SvmError) as e:
            P.server_submit(q.plans())
        assert e.value.code == P._lib.ERR_BUSY
    finally:
        P.server_stop()


def test_sequential_submissions_reach_a_worker_that_serves_their_lane(serve_tables):
    """One query at a time (each waited for before the next), heavy and light, to a
    batching server whose worker 0 serves the light lane only: every query is taken by
    a worker of its lane (a wake-up that reached only the light-lane worker would leave
    a heavy query waiting; the bounded wait turns that into a failure)."""
    wl = serve_tables
    qs = ssb.serve_queries(wl, 12, seed=8, light_frac=0.5, light_rows=4_000_000, prefix="sw")
    ssb.make_views(qs)
    assert any(q.shape["kind"] == "heavy" for q in qs) and any(q.shape["kind"] == "light" for q in qs)
    want = _serve(qs, mode=1)
    P.server_start(workers=3, budget=0, max_batch=64, mode=0)
    try:
        for q, w in zip(qs, want):
            got = P.server_wait(P.server_submit(q.plans()), timeout_ms=20_000)
            assert got.rows == w.rows
            assert_same_columns(got.columns, w.columns)
    finally:
        P.server_stop()
