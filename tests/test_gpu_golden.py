"""The device runtime against golden vectors of the reference's own CPU runtime
(tests/golden/runtime_vectors.json, from oracle/_ref, compiled from
com/cpp/cpu-runtime.hpp):

* a string group key is Hash(bytes, width) over every byte of the field, NULs
  included (codegen/codegen.cpp:262-264): a join onto a dimension of the golden
  strings, grouped by the string, yields exactly the golden hashes;
* StrCmp (cpu-runtime.hpp:37-45) returns the reference's value, not only its sign:
  a Map projecting StrCmp of two string columns yields the golden results row by row.
"""
import collections
import json
import os

import numpy as np
import pytest

from shadowvm_amd import pipeline as P

pytestmark = pytest.mark.gpu

I, S = P.INT32, P.STRING
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "runtime_vectors.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_string_group_keys_are_reference_hashes(engine, golden):
    by_width = collections.defaultdict(dict)
    for hx, h, _ in golden["hash"]:
        b = bytes.fromhex(hx)
        by_width[len(b)][b] = h
    checked = 0
    for w, table in sorted(by_width.items()):
        strs = list(table)
        n = len(strs)
        print(f"[golden] hash width {w}: {n} strings", flush=True)
        keys = np.arange(1, n + 1, dtype=np.int32)
        P.table_upload("gh_dim", [I, S], [4, w], [keys, np.frombuffer(b"".join(strs), np.uint8).reshape(n, w)])
        P.table_upload("gh_fact", [I, I], [4, 4], [keys, np.ones(n, np.int32)])
        try:
            pt = P.PipelineTest("gh_q")
            res = P.create_schema([I] * 4, [4] * 4)
            pt.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema([I, I], [4, 4]))
            pt.AddOperator(P.JOINWITH, {"lkey": "$(0.0)", "rkey": "$(1.0)"}, res)
            pt.AddOperator(P.REDUCEBY, {"dim": "$(1.1)", "metric": "$(0.1)"}, res)
            pt.AddOperator(P.SINK, {}, res)
            got = pt.Run("gh_fact", ["gh_dim"])
        finally:
            P.table_drop("gh_dim")
            P.table_drop("gh_fact")
        want = collections.Counter(np.int32(h).item() for h in table.values())
        assert np.asarray(got.columns[0]).tolist() == sorted(want), w
        assert np.asarray(got.columns[1]).tolist() == [want[k] for k in sorted(want)], w
        checked += n
    assert checked == len({bytes.fromhex(x[0]) for x in golden["hash"]})


def test_strcmp_values_are_reference_values(engine, golden):
    """Select lists split on ',' (core/pipeline.cu:11-80), so StrCmp's value is read
    through filters: bit b of StrCmp + 256 (in [1, 511]) passes the rows that have it."""
    by_n = collections.defaultdict(list)
    for ax, cx, n, want in golden["strcmp"]:
        by_n[n].append((bytes.fromhex(ax), bytes.fromhex(cx), want))
    for n, rows in sorted(by_n.items()):
        m = len(rows)
        print(f"[golden] StrCmp width {n}: {m} pairs", flush=True)
        a = np.frombuffer(b"".join(r[0] for r in rows), np.uint8).reshape(m, n)
        c = np.frombuffer(b"".join(r[1] for r in rows), np.uint8).reshape(m, n)
        P.table_upload("gs_src", [S, S, I], [n, n, 4], [a, c, np.arange(m, dtype=np.int32)])
        value = np.full(m, -256, np.int64)
        try:
            for b in range(9):
                pt = P.PipelineTest("gs_out")
                osch = P.create_schema([I, S], [4, 40])
                pt.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema([S, S, I], [n, n, 4]))
                pt.AddOperator(P.FILTER, {"predication": f"((StrCmp($(0.0),$(0.1),{n})+256)>>{b})&1"}, osch)
                pt.AddOperator(P.MAP, {"select": "$(0.2)"}, osch)
                pt.AddOperator(P.SINK, {}, osch)
                got = pt.Run("gs_src")
                value[np.asarray(got.columns[0][:got.rows], np.int64)] += 1 << b
        finally:
            P.table_drop("gs_src")
        assert value.tolist() == [r[2] for r in rows], n
