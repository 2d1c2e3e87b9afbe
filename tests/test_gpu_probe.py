"""Joins that probe the reference's hash table itself (Probe<T>, Probe<char>,
com/cpp/cpu-runtime.hpp:125-152) equal the oracle byte for byte: String keys (the
Yahoo streaming query of test/test-yahoo.cu:33-51), Int64 and Int16 keys, and Int32
keys whose range is too wide for a dense map.  At a bulk capacity of 16 the string
probe's quirks are reached on purpose: StrCmp's prefix equality, and row 0 (not a
miss) after 10 occupied slots of other keys."""
import numpy as np
import pytest

from oracle import oracle as O
from shadowvm_amd import pipeline as P

from .helpers import assert_same_columns
from .test_gpu_refhash import engine_with

pytestmark = pytest.mark.gpu

I16, I32, I64, S = P.INT16, P.INT32, P.INT64, P.STRING


def strings(values, width):
    out = np.zeros((len(values), width), np.uint8)
    for r, v in enumerate(values):
        b = v.encode()[:width]
        out[r, :len(b)] = np.frombuffer(b, np.uint8)
    return out


def str_hash(b):
    """StrHash: Hash31 over the bytes up to the first NUL (com/cpp/cpu-runtime.hpp:60-66)."""
    h = 0
    for x in b:
        if x == 0:
            break
        h = (h * 31 + (x - 256 if x > 127 else x)) & 0xFFFFFFFF
    return h & 0x7FFFFFFF


def both(plan_dims, plan_fact, fact_tab, dim_tabs, tpb):
    """Run dimension pipelines then the fact pipeline on the GPU and on the oracle.
    plan_dims: [(PipelineTest, source id)], dim_tabs: {source id: (types, sizes, cols)}."""
    odims = {}
    for pt, src in plan_dims:
        pt.Run(src, want_result=False)
        t, sz, cols = dim_tabs[src]
        odims[pt.pipe_info["id"]] = O.run_pipeline(pt.plan(src), O.OTable(t, sz, cols, tpb), tuple_per_bulk=tpb).table
    fq, src, refs = plan_fact
    got = fq.Run(src, refs)
    t, sz, cols = fact_tab
    want = O.run_pipeline(fq.plan(src, refs), O.OTable(t, sz, cols, tpb), [odims[r] for r in refs], tuple_per_bulk=tpb)
    return got, want


def yahoo_tables(seed, n_ad=1200, n_ev=200_000):
    rng = np.random.default_rng(seed)
    hexs = lambda n: ["".join(rng.choice(list("0123456789abcdef"), 32)) for _ in range(n)]
    ad_ids = hexs(n_ad)
    camp = [f"campaign-{int(x):04d}" for x in rng.integers(0, 100, n_ad)]
    ad = [strings(ad_ids, 40), strings(camp, 40), strings(hexs(n_ad), 40)]
    unknown = hexs(50)
    pick = rng.random(n_ev)
    ev_ad = [ad_ids[i] if p < 0.9 else unknown[i % 50] for i, p in zip(rng.integers(0, n_ad, n_ev), pick)]
    ev = [strings(hexs(1)[0:1] * n_ev, 40), strings([f"page-{x}" for x in rng.integers(0, 500, n_ev)], 40),
          strings(ev_ad, 40), strings(rng.choice(["banner", "modal", "sponsored"], n_ev).tolist(), 10),
          strings(rng.choice(["view", "click", "purchase"], n_ev).tolist(), 10),
          strings([str(x) for x in rng.integers(0, 10**9, n_ev)], 10), strings(["1.2.3.4"] * n_ev, 10)]
    return ad, ev


def yahoo_plans(prefix):
    ad_s = P.create_schema([S, S, S], [40, 40, 40])
    ev_s = P.create_schema([S] * 7, [40, 40, 40, 10, 10, 10, 10])
    res = P.create_schema([S, S, S], [40, 40, 40])
    q = P.PipelineTest(f"{prefix}_q")
    q.AddOperator(P.SOURCE, {"type": "object"}, ev_s)
    q.AddOperator(P.FILTER, {"predication": 'StrCmp($(0.4),"view",10)==0'}, res)
    q.AddOperator(P.JOINWITH, {"lkey": "$(0.2)", "rkey": "$(1.0)"}, res)
    q.AddOperator(P.MAP, {"select": "$(1.1),$(1.0), $(0.2)"}, res)
    q.AddOperator(P.SINK, {}, res)
    return ad_s, ev_s, q


def test_yahoo_string_key_join(engine):
    """test/test-yahoo.cu:33-51 on generated events and ads: 40-byte string ad ids."""
    ad, ev = yahoo_tables(5)
    P.table_upload("yh_ad", [S, S, S], [40, 40, 40], ad)
    P.table_upload("yh_ev", [S] * 7, [40, 40, 40, 10, 10, 10, 10], ev)
    _, _, q = yahoo_plans("yh")
    got = q.Run("yh_ev", ["yh_ad"])
    want = O.run_pipeline(q.plan("yh_ev", ["yh_ad"]), O.OTable([S] * 7, [40, 40, 40, 10, 10, 10, 10], ev),
                          [O.OTable([S, S, S], [40, 40, 40], ad)])
    print(f"[probe] yahoo: {got.rows} joined view events", flush=True)
    assert got.rows == want.rows > 10_000
    assert_same_columns(got.columns, want.columns)


def _keyed_case(prefix, ktype, ksize, keys, seed, n_fact=400_000):
    """A (key, attr) dimension with the given keys, filtered attr < 30, and a fact of
    (key, value) rows: grouped sum of value by the dimension's attr."""
    rng = np.random.default_rng(seed)
    n = len(keys)
    attr = rng.integers(0, 50, n).astype(np.int32)
    misses = (keys.max() // 2 + 7 + np.arange(20)).astype(keys.dtype)
    fk = np.where(rng.random(n_fact) < 0.95, keys[rng.integers(0, n, n_fact)], misses[rng.integers(0, 20, n_fact)])
    fk = fk.astype(keys.dtype)
    val = rng.integers(-1000, 1000, n_fact).astype(np.int32)
    P.table_upload(f"{prefix}_dim", [ktype, I32], [ksize, 4], [keys, attr])
    P.table_upload(f"{prefix}_fact", [ktype, I32], [ksize, 4], [fk, val])
    d = P.PipelineTest(f"{prefix}_d")
    dsch = P.create_schema([ktype, I32, S], [ksize, 4, 40])
    d.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema([ktype, I32], [ksize, 4]))
    d.AddOperator(P.FILTER, {"predication": "$(0.1)<30"}, dsch)
    d.AddOperator(P.MAP, {"select": "$(0.0), $(0.1)"}, dsch)
    d.AddOperator(P.SINK, {}, dsch)
    f = P.PipelineTest(f"{prefix}_q")
    res = P.create_schema([I32] * 4, [4] * 4)
    f.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema([ktype, I32], [ksize, 4]))
    f.AddOperator(P.JOINWITH, {"lkey": "$(0.0)", "rkey": "$(1.0)"}, res)
    f.AddOperator(P.REDUCEBY, {"dim": "$(1.1)", "metric": "$(0.1)"}, res)
    f.AddOperator(P.SINK, {}, res)
    tabs = {f"{prefix}_dim": ([ktype, I32], [ksize, 4], [keys, attr])}
    return both([(d, f"{prefix}_dim")], (f, f"{prefix}_fact", [f"{prefix}_d"]),
                ([ktype, I32], [ksize, 4], [fk, val]), tabs, 3_000_000)


def scan_config(pipeline_id):
    import ctypes
    from shadowvm_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    f = L.svm_debug_scan_config
    f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]
    buf = ctypes.create_string_buffer(4096)
    assert f(pipeline_id.encode(), buf, len(buf)) == 0
    return buf.value.decode()


def test_sparse_int32_keys_probe(engine):
    """20,000 Int32 keys spread over 2^26 (one key in 3,355 of the range): a dense map
    would need 2^26 slots, so the join probes the reference's hash table, and the
    answer is the oracle's."""
    rng = np.random.default_rng(15)
    keys = np.unique(rng.integers(0, 2**26, 20_000, dtype=np.int64)).astype(np.int32)
    got, want = _keyed_case("pks", I32, 4, rng.permutation(keys), 16)
    cfg = scan_config("pks_q")
    print(f"[probe] sparse int32: {got.rows} groups, scan {cfg}", flush=True)
    assert "probe(0)" in cfg, cfg
    assert got.rows == want.rows == 30
    assert_same_columns(got.columns, want.columns)


@pytest.mark.parametrize("case", ["int64", "int16", "int32_wide"])
def test_typed_key_probes(engine, case):
    rng = np.random.default_rng(11)
    if case == "int64":
        keys = np.unique(rng.integers(-2**62, 2**62, 60_000, dtype=np.int64))
        got, want = _keyed_case("pk64", I64, 8, rng.permutation(keys), 12)
    elif case == "int16":
        keys = np.arange(-3000, 3000, dtype=np.int16)
        got, want = _keyed_case("pk16", I16, 2, rng.permutation(keys), 13)
    else:
        keys = np.unique(rng.integers(-2**31, 2**31 - 1, 80_000, dtype=np.int64)).astype(np.int32)
        got, want = _keyed_case("pk32", I32, 4, rng.permutation(keys), 14)
    print(f"[probe] {case}: {got.rows} groups", flush=True)
    assert got.rows == want.rows == 30
    assert_same_columns(got.columns, want.columns)


def test_string_probe_quirks(engine):
    """Bulk capacity 16 (80 slots): 12 ad ids share one home slot, so 10 fill its
    window and 2 are dropped; an event id with that home that equals none of them
    walks 10 occupied slots and joins row 0; an id that is a NUL-terminated prefix of
    a stored id, with the same home, joins that row (StrCmp stops at the first NUL)."""
    slots = 80
    base = "ad-"
    home = str_hash(b"ad-seed") % slots
    same = []
    k = 0
    while len(same) < 14:
        cand = f"{base}{k:06d}"
        if str_hash(cand.encode()) % slots == home:
            same.append(cand)
        k += 1
    stored, ghost, outsider = same[:12], same[12], same[13]
    far = lambda x: min((str_hash(x.encode()) - home) % slots, (home - str_hash(x.encode())) % slots) > 12
    # a stored id and a prefix of it with the same home, away from the crowded window
    pairs = ((f"zz-{k}", f"zz-{k}-{j}") for k in range(1000) for j in range(1000))
    pre, full = next((p, f) for p, f in pairs
                     if str_hash(f.encode()) % slots == str_hash(p.encode()) % slots and far(f))
    k = 0
    while not far(f"none-{k}") or str_hash(f"none-{k}".encode()) % slots == str_hash(full.encode()) % slots:
        k += 1
    nothing = f"none-{k}"
    ad_ids = stored + [full]
    n_ad = len(ad_ids)
    ad = [strings(ad_ids, 40), strings([f"camp-{i}" for i in range(n_ad)], 40), strings(["x"] * n_ad, 40)]
    ev_ids = stored + [ghost, outsider, pre, nothing]
    ev = [strings(["u"] * len(ev_ids), 40), strings(["p"] * len(ev_ids), 40), strings(ev_ids, 40),
          strings(["banner"] * len(ev_ids), 10), strings(["view"] * len(ev_ids), 10),
          strings(["0"] * len(ev_ids), 10), strings(["ip"] * len(ev_ids), 10)]
    with engine_with(engine, 16):
        P.table_upload("yq_ad", [S, S, S], [40, 40, 40], ad)
        P.table_upload("yq_ev", [S] * 7, [40, 40, 40, 10, 10, 10, 10], ev)
        _, _, q = yahoo_plans("yq")
        got = q.Run("yq_ev", ["yq_ad"])
        want = O.run_pipeline(q.plan("yq_ev", ["yq_ad"]), O.OTable([S] * 7, [40, 40, 40, 10, 10, 10, 10], ev, 16),
                              [O.OTable([S, S, S], [40, 40, 40], ad, 16)], tuple_per_bulk=16)
        assert want.dropped[1] == 2
        joined = [bytes(r).rstrip(b"\0").decode() for r in want.columns[2]]
        assert ghost in joined and outsider in joined and pre in joined  # the quirks fired
        assert nothing not in joined
        assert got.rows == want.rows
        assert_same_columns(got.columns, want.columns)


def test_yahoo_from_files(engine, tmp_path):
    """The whole of test/test-yahoo.cu: both tables loaded from '|'-separated files by
    Loader pipelines (object/object.cpp:122-169), then the filtered string-key join and
    projection, against the oracle's loader and query."""
    from oracle import loader as OL
    from .test_gpu_loader import _loader
    ad, ev = yahoo_tables(6, n_ad=800, n_ev=60_000)

    def text(cols):
        rows = zip(*[[bytes(r).rstrip(b"\0").decode() for r in c] for c in cols])
        return "".join("|".join(r) + "|\n" for r in rows)

    (tmp_path / "ad-camp-map").mkdir()
    (tmp_path / "ad-camp-map" / "part0").write_text(text(ad))
    (tmp_path / "events").mkdir()
    body = text(ev)
    cut = body.index("\n", len(body) // 2) + 1  # two files, split at a row boundary
    (tmp_path / "events" / "part0").write_text(body[:cut])
    (tmp_path / "events" / "part1").write_text(body[cut:])
    ad_t, ad_s = [S, S, S], [40, 40, 40]
    ev_t, ev_s = [S] * 7, [40, 40, 40, 10, 10, 10, 10]
    _loader("yf_ad", tmp_path / "ad-camp-map", ad_t, ad_s).Run(None, want_result=False)
    _loader("yf_ev", tmp_path / "events", ev_t, ev_s).Run(None, want_result=False)
    _, _, q = yahoo_plans("yf")
    got = q.Run("yf_ev", ["yf_ad"])
    o_ad = OL.load(str(tmp_path / "ad-camp-map"), ad_t, ad_s)
    o_ev = OL.load(str(tmp_path / "events"), ev_t, ev_s)
    want = O.run_pipeline(q.plan("yf_ev", ["yf_ad"]), O.OTable(ev_t, ev_s, o_ev), [O.OTable(ad_t, ad_s, o_ad)])
    print(f"[probe] yahoo from files: {got.rows} rows", flush=True)
    assert got.rows == want.rows > 5_000
    assert_same_columns(got.columns, want.columns)
