"""The drop-in boundary without a GPU: libsvm.so loads and exports every symbol
include/*.h declares; PipelineInfo bytes from the protobuf runtime decode to the
operator chain the reference's Pipeline::Pipeline builds (core/pipeline.cu:11-80);
expression lowering and error codes; generated kernels compile for gfx950."""
import ctypes
import glob
import json
import os
import re

import pytest

from shadowvm_amd import _lib, proto
from shadowvm_amd import pipeline as P
from shadowvm_amd import ssb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(svm_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers the public ABI
    public = {n for n in names if not n.startswith("svm_debug")}
    assert public <= set(_lib.SIGNATURES), sorted(public - set(_lib.SIGNATURES))


def describe(plan):
    L = ctypes.CDLL(_lib.LIB_PATH)
    f = L.svm_debug_describe_plan
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]
    data = proto.serialize(plan) if isinstance(plan, dict) else plan
    buf = ctypes.create_string_buffer(1 << 20)
    rc = f(data, len(data), buf, len(buf))
    return rc, buf.value.decode()


def test_decode_single_join_plan():
    wl = ssb.single_join(seed=3, fact_rows=1000)
    fq, src, refs = wl.fact
    rc, js = describe(fq.plan(src, refs))
    assert rc == 0, js
    d = json.loads(js)
    c = wl.constants
    assert d["id"] == "sj_query"
    assert d["pre"] == ["sj_lineorder", "sj_query_date"]
    assert d["join_lkey"] == [[0, 0]] and d["join_rkey"] == [[1, 0]]
    assert d["is_reduce"] and d["reduce_key"] == [] and d["metric"] == ["$(0.3)*$(0.2)"]
    # Int32 source operands become registers, decimal literals become parameters
    assert d["filter"][0][1] == " (f2)>=PI(0) && (f2)<=PI(1) && (f1)<PI(2) "
    assert d["int_literals"] == [c["discount"], c["discount"] + 2, c["quantity"]]
    assert d["output"] == [[2, 4]] * 9
    # the C++ re-encoder produces the same number of bytes as protobuf
    assert d["reencoded_len"] == len(proto.serialize(fq.plan(src, refs)))


def test_decode_reference_quirks():
    # StrSplit / GenKeyIndex tolerance: trailing ')' after the last key, spaces,
    # an empty dim list, SortBy ignored, several Filters AND-ed in order.
    pt = P.PipelineTest("q")
    s = P.create_schema([P.INT32, P.STRING, P.INT32], [4, 10, 4])
    pt.AddOperator(P.SOURCE, {"type": "object"}, s)
    pt.AddOperator(P.FILTER, {"predication": 'StrCmp($(0.1),"MFGR#12",8)==0'}, s)
    pt.AddOperator(P.FILTER, {"predication": "$(0.2) < 010 || $(0.0)>3000000000"}, s)
    pt.AddOperator(P.SORTBY, {"key": "$(0.0)"}, s)
    pt.AddOperator(P.JOINWITH, {"lkey": "$(0.0)", "rkey": "$(1.0)"}, s)
    pt.AddOperator(P.REDUCEBY, {"dim": "$(1.1), $(2.1))", "metric": "($(0.2)-$(0.0))/100"}, s)
    pt.AddOperator(P.SINK, {}, s)
    rc, js = describe(pt.plan("src", ["a", "b"]))
    assert rc == 0, js
    d = json.loads(js)
    assert d["reduce_key"] == [[1, 1], [2, 1]]
    assert d["metric"] == ["($(0.2)-$(0.0))/100"]
    f0, f1 = d["filter"]
    # StrCmp's length stays a literal (the kernel unrolls over it); other literals
    # become launch parameters
    assert f0[1] == "StrCmp(((const char*)(FCOL1+(long long)i*10)),PS(0),8)==PI(0)"
    # octal and out-of-int32 literals keep their C meaning verbatim
    assert f1[1] == "(f2) < 010 || (f0)>3000000000"
    assert d["string_literals"].startswith("MFGR#12")


def _lowered_filter(pred, types=(P.INT32, P.INT32, P.STRING), sizes=(4, 4, 10)):
    pt = P.PipelineTest("w")
    s = P.create_schema(list(types), list(sizes))
    pt.AddOperator(P.SOURCE, {"type": "object"}, s)
    pt.AddOperator(P.FILTER, {"predication": pred}, s)
    pt.AddOperator(P.MAP, {"select": "$(0.0)"}, s)
    pt.AddOperator(P.SINK, {}, s)
    rc, js = describe(pt.plan("src", []))
    assert rc == 0, js
    return json.loads(js)["filter"][0][1]


@pytest.mark.parametrize("pred", [
    "$(0.0) > svm_lds[0]",                 # a kernel variable
    "atomicAdd((unsigned*)0, 1) > 0",      # a device function
    "$(0.0) > a.pi[0]",                    # the launch arguments
    "$(0.0) > 0; }",                       # closing the generated statement
    "$(0.0) > 0 /* && 0 */",               # a comment
    "$(0.0) > 0 // rest of the line",
    "5 = 6",                               # an assignment (to a literal parameter)
    "$(0.0) += 1",
    "$(0.0) <<= 1",
    "$(0.0)++ > 0",
    "$(0.0) > 1.ptr",                      # a malformed number
    "#define x 1",
])
def test_lowering_rejects_foreign_code(pred):
    low = _lowered_filter(pred)
    assert low.startswith("!"), low
    assert "not allowed" in low or "malformed" in low, low


def test_lowering_accepts_the_reference_language():
    # C operators, casts, the alternative spellings the Spark control plane emits
    # (java_vRDD/src/test/java/SSBHelper.java:371) and the runtime functions
    low = _lowered_filter('StrCmp($(0.2),"MFGR#2221",10)>=0 and not ($(0.0) != 3) or '
                          '(long long)$(0.0)*$(0.1) >= 0x10 && $(0.1) % 7 <= 2u && !($(0.0) == 1) ? 1 : 0')
    assert not low.startswith("!"), low
    assert "and not" in low and "(long long)(f0)*(f1)" in low


def test_decode_rejects_malformed_bytes():
    rc, _ = describe(b"\x0a\xff\xff\xff\xff\x0f")  # length beyond the buffer
    assert rc == _lib.ERR_INVALID
    rc, _ = describe(b"\x0b")  # wire type 3 (groups) is not proto3
    assert rc == _lib.ERR_INVALID


def test_run_before_init_is_state_error():
    data = proto.serialize(ssb.single_join(seed=1, fact_rows=10).fact[0].plan("x", []))
    L = _lib.lib()
    buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
    out = ctypes.c_void_p()
    rc = L.svm_pipeline_run(buf, len(data), ctypes.byref(out))
    assert rc == _lib.ERR_STATE
    assert b"svm_init" in L.svm_last_error()


def test_generated_kernels_compile_for_gfx950(tmp_path, monkeypatch):
    monkeypatch.setenv("SVM_JIT_CACHE", str(tmp_path))
    L = ctypes.CDLL(_lib.LIB_PATH)
    buf = ctypes.create_string_buffer(1 << 20)
    rc = L.svm_debug_jit_selftest(buf, len(buf))
    report = buf.value.decode()
    assert rc == 0, report
    assert report.count(": ok") == 20
    assert len([f for f in os.listdir(tmp_path) if f.endswith(".co")]) == 20  # code objects cached per source
    # the host engine's code of five shapes, compiled by the host compiler and loaded
    assert report.count(": host ok") == 5
    assert len([f for f in os.listdir(tmp_path) if f.startswith("host_") and f.endswith(".so")]) == 5


def test_variant_counts_match_the_variant_tests():
    """The GPU variant tests force every autotuner candidate by index; their counts
    must be the library's, or a candidate would go unchecked."""
    from . import test_gpu_variants as V
    L = ctypes.CDLL(_lib.LIB_PATH)
    scan, batch = ctypes.c_int(0), ctypes.c_int(0)
    assert L.svm_debug_variant_counts(ctypes.byref(scan), ctypes.byref(batch)) == 0
    assert (scan.value, batch.value) == (V.N_SCAN_CANDIDATES, V.N_BATCH_VARIANTS)


def test_ssb_query_plans_decode_and_lower():
    """Every plan of the thirteen SSB query tests (dimension and fact pipelines) is
    decoded and its expressions lowered without a GPU."""
    from .ssb_queries import NAMES, queries
    ids = {k: f"ssbq_{k}" for k in ("lineorder", "date", "customer", "supplier", "part")}
    q = queries("ssbq_", ids)
    for name in NAMES:
        dims, fact = q[name]
        for pt, src in dims:
            rc, js = describe(pt.plan(src))
            assert rc == 0, (name, pt.pipe_info["id"], js)
        rc, js = describe(fact.plan(ids["lineorder"], [pt.pipe_info["id"] for pt, _ in dims]))
        assert rc == 0, (name, js)
