"""The Loader operator on the GPU: '|'-separated files parsed in HBM equal the
oracle (oracle/loader.py, which calls the C library's atoi/atol as the reference
does) byte for byte on seeded tricky text; SSB-style files written from the
generator load back into tables on which the star-join query gives the same
answer as on the generated tables."""
import random

import numpy as np
import pytest

from oracle import loader as OL
from shadowvm_amd import _lib
from shadowvm_amd import pipeline as P
from shadowvm_amd import ssb

pytestmark = pytest.mark.gpu

I, S, I16, I64, B = P.INT32, P.STRING, P.INT16, P.INT64, P.BYTE


def _loader(pid, path, types, sizes, col="|", row="\n"):
    sch = P.create_schema(types, sizes)
    pt = P.PipelineTest(pid)
    pt.AddOperator(P.SOURCE, {"type": "file", "path": str(path)}, sch)
    pt.AddOperator(P.LOADER, {"col_spilter": col, "row_spilter": row}, sch)
    pt.AddOperator(P.SINK, {}, sch)
    return pt


def _tricky_text(rng, nrows, ncols):
    alphabet = "abcXYZ019 _-+#"
    lines = []
    for _ in range(nrows):
        if rng.random() < 0.03:
            lines.append("")  # an empty row
            continue
        fields = []
        for _ in range(max(0, ncols + rng.choice([-2, -1, 0, 0, 0, 0, 1]))):
            k = rng.random()
            if k < 0.45:
                v = rng.randint(-2 ** 40, 2 ** 40) if rng.random() < 0.3 else rng.randint(-99999, 99999)
                pre = rng.choice(["", "", " ", "  ", "\t", "+"])
                fields.append(f"{pre}{v}{rng.choice(['', '', 'x', ' 7'])}")
            elif k < 0.55:
                fields.append(rng.choice(["", "-", "+", "99999999999999999999999", "-9223372036854775809"]))
            else:
                fields.append("".join(rng.choice(alphabet) for _ in range(rng.randint(0, 14))))
        end = rng.choice(["|\n", "|\n", "|\n", "\n", "|\r\n"])
        lines.append("|".join(fields) + end)
    return "".join(l if l.endswith("\n") else l + "\n" for l in lines)


def _check_same(pid, types, sizes, want):
    rows, ncols = P.table_info(pid)
    assert ncols == len(types)
    assert rows == len(want[0])
    for c in range(ncols):
        got = P.table_column(pid, c)
        assert np.array_equal(np.asarray(got), np.asarray(want[c])), f"column {c}"


def test_tricky_text_equals_oracle(engine, tmp_path):
    rng = random.Random(7)
    types = [I, S, I16, I64, B, I, S]
    sizes = [4, 10, 2, 8, 1, 4, 3]
    d = tmp_path / "tricky"
    d.mkdir()
    for name in ["c.tbl", "a.tbl", "b.tbl"]:  # rows are concatenated in name order
        (d / name).write_bytes(_tricky_text(rng, 3000, len(types)).encode() + b"5|partial")
    res = _loader("ld_tricky", d, types, sizes).Run(None)
    want = OL.load(str(d), types, sizes)
    print(f"[loader] {res.rows} rows from 3 files", flush=True)
    assert res.rows == len(want[0]) > 8000
    _check_same("ld_tricky", types, sizes, want)
    # one file, other separators
    f = tmp_path / "one.csv"
    f.write_bytes(b"1,22,x;\n;3,-4,yz,;")
    _loader("ld_one", f, [I, I, S], [4, 4, 4], col=",", row=";").Run(None)
    _check_same("ld_one", [I, I, S], [4, 4, 4], OL.load(str(f), [I, I, S], [4, 4, 4], ",", ";"))


def test_ssb_files_round_trip(engine, tmp_path):
    wl = ssb.single_join(seed=5, fact_rows=1_000_000, prefix="lr")
    ssb.load_device(wl)
    loaded = {}
    for t in wl.tables:
        types, sizes, cols = ssb.host_table(wl, t)
        assert all(ty == I for ty in types)
        text = "\n".join("|".join(map(str, row)) + "|" for row in zip(*[c.tolist() for c in cols])) + "\n"
        d = tmp_path / t.id
        d.mkdir()
        half = len(text) // 2
        cut = text.index("\n", half) + 1  # two files, split at a row boundary
        (d / "part0.tbl").write_text(text[:cut])
        (d / "part1.tbl").write_text(text[cut:])
        _loader(t.id + "_ld", d, types, sizes).Run(None, want_result=False)
        _check_same(t.id + "_ld", types, sizes, cols)
        loaded[t.id] = t.id + "_ld"
    direct = ssb.run_query(wl)
    for pt, src, refs in wl.dims:
        pt.Run(loaded[src], refs, want_result=False)
    fq, src, refs = wl.fact
    via_files = fq.Run(loaded[src], refs)
    assert via_files.rows == direct.rows
    for a, b in zip(via_files.columns, direct.columns):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_loader_errors(engine, tmp_path):
    with pytest.raises(_lib.SvmError) as e:
        _loader("ld_missing", tmp_path / "nope", [I], [4]).Run(None)
    assert e.value.code == _lib.ERR_NOT_FOUND
    f = tmp_path / "f.tbl"
    f.write_bytes(b"1.5|\n")
    with pytest.raises(_lib.SvmError) as e:
        _loader("ld_float", f, [P.FLOAT32], [4]).Run(None)
    assert e.value.code == _lib.ERR_UNSUPPORTED
