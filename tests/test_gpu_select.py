"""Dimension selects (Filter + Map, order-kept compaction) against the oracle at the
sizes where the compaction changes form.

A source of one bulk with at most 256 tiles of 1,024 rows is compacted in one launch
(tiles taken from a ticket, each finding the rows of earlier tiles by a look-back over
per-tile status words); larger sources count the tiles first and write in a second
launch.  Both must keep source row order exactly, whatever the row count (ragged last
tiles, one row, exactly 256 tiles, one row past it) and whatever passes (nothing,
everything, half).  Outputs are compared byte for byte, string projections included,
and the same select is run repeatedly on one stream (new launch epochs each time).
"""
import numpy as np
import pytest

from oracle import oracle as O
from shadowvm_amd import pipeline as P

pytestmark = pytest.mark.gpu

I, S = P.INT32, P.STRING


def _table(n, seed):
    rng = np.random.default_rng(seed)
    key = rng.permutation(n).astype(np.int32) + 1
    attr = rng.integers(0, 100, n).astype(np.int32)
    names = np.frombuffer(b"".join(b"NAME_%05d" % (v % 100000) for v in rng.integers(0, 1 << 30, n)),
                          np.uint8).reshape(n, 10)
    return [key, attr, names]


def _plan(pid, filt):
    pt = P.PipelineTest(pid)
    osch = P.create_schema([I, S, I, S], [4, 10, 4, 40])
    pt.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema([I, I, S], [4, 4, 10]))
    if filt:
        pt.AddOperator(P.FILTER, {"predication": filt}, osch)
    pt.AddOperator(P.MAP, {"select": "$(0.0), $(0.2), $(0.1)"}, osch)
    pt.AddOperator(P.SINK, {}, osch)
    return pt


@pytest.mark.parametrize("n", [1, 3, 1023, 1024, 1025, 262_143, 262_144, 262_145, 1_000_003])
@pytest.mark.parametrize("filt", ["", "$(0.1)<0", "$(0.1)<50", 'StrCmp($(0.2),"NAME_0",6)==0'])
def test_select_forms_equal_oracle(engine, n, filt):
    tid = f"sel_src_{n}"
    cols = _table(n, seed=n)
    if not P.table_exists(tid):
        P.table_upload(tid, [I, I, S], [4, 4, 10], cols)
    want = O.run_pipeline(_plan("sel_out", filt).plan(tid), O.OTable([I, I, S], [4, 4, 10], cols,
                                                                      max_tuple_count=3_000_000), [])
    print(f"[select] n={n} filter={filt!r} oracle rows={want.rows}", flush=True)
    for rep in range(3):
        got = _plan("sel_out", filt).Run(tid)
        assert got.rows == want.rows, (rep, got.rows, want.rows)
        for q, (g, w) in enumerate(zip(got.columns, want.columns)):
            assert np.array_equal(np.asarray(g), np.asarray(w)), (rep, q)
