"""CPU oracle for the star-join path.  TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module, and only as the checker.  The product
(``shadowvm_amd``) never imports, links or runs anything under ``oracle/``.

What it restates
----------------
The reference's *generated CPU path* for one pipeline:

* plan decoding into filter / select / metric / reduce-key / join lists,
  ``Pipeline::Pipeline``  (core/pipeline.cu:11-80), with ``StrSplit``
  (com/cpp/com.cu:79-93) and ``CodeGen::GenKeyIndex`` (codegen/codegen.cpp:231-238);
* expression lowering ``CodeGen::GenExperssion`` (codegen/codegen.cpp:240-277):
  ``$(o.c)`` becomes a typed load at row ``i`` (object 0) or at
  ``join_index[o]`` (object o > 0); string operands become ``char*`` or, as a
  reduce key, ``Hash(ptr, size)``; the rest of the text is emitted verbatim so
  C int32 semantics apply;
* the blocks ``GenFilterBlock`` / ``GenJoinBlock`` / ``GenSelectBlock`` /
  ``GenReduceBlock`` (codegen/codegen.cpp:100-229) of the CPU variant, inside
  the ``HostQueryTransform`` row loop (codegen/template.cpp:61-73);
* compiled with ``g++ -O2 -fwrapv`` as the reference JIT compiles with ``clang -O2``
  (codegen/compiler.cpp:30), then run one task per bulk with per-bulk sort +
  reduce and the documented merge (oracle/svm_oracle_rt.h).

Deliberate deviations (the same ones the product documents):
  D1 metric sums are int64 (each row's value is the expression converted to
     the declared output column type, as the reference stores it, then widened);
  D2 per-bulk partials are merged into one result ordered by key;
  D4 join ``k`` writes ``join_index[rkey object]`` (the reference writes
     ``join_index[k+1]`` at codegen/codegen.cpp:140 but reads
     ``join_index[obj]`` at :258-261).
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess
import tempfile
import threading
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# com/svm.proto:62-72 and :101-111
BYTE, INT16, INT32, INT64, FLOAT16, FLOAT32, FLOAT64, STRING, TENSOR = range(9)
SOURCE, LOADER, FILTER, MAP, REDUCEBY, SORTBY, JOINWITH, SINK = range(8)

# CodeGen::ToNativeType (codegen/codegen.cpp:12-21)
NATIVE = {BYTE: "char", INT16: "short", INT32: "int", INT64: "long long",
          FLOAT32: "float", FLOAT64: "double", STRING: "char", TENSOR: "char"}
NP_DTYPE = {BYTE: np.int8, INT16: np.int16, INT32: np.int32, INT64: np.int64}

MAX_OBJ, MAX_COL = 8, 24
PAD = 64  # zero bytes after every column (StrCopy may read past the last row)


class OracleError(Exception):
    pass


# --------------------------------------------------------------------------- tables
@dataclass
class OTable:
    """A columnar table on the host.  As a source it is cut into bulks of the
    driver's ``tuple_per_bulk`` rows; as a ref object only its bulk 0 (the first
    ``bulk0_rows`` rows) is read."""

    types: list
    sizes: list
    columns: list  # np.ndarray per column, or None for an unmaterialised column
    max_tuple_count: int = 3_000_000
    rows: int = field(default=-1)
    # rows of bulk 0, the only bulk a join reads (core/pipeline.cu:91); -1: the table
    # is cut into bulks of max_tuple_count rows
    bulk0_rows: int = field(default=-1)

    def __post_init__(self):
        if self.rows < 0:
            self.rows = 0
            for c in self.columns:
                if c is not None:
                    self.rows = int(c.shape[0])
                    break
        if self.bulk0_rows < 0:
            self.bulk0_rows = min(self.rows, self.max_tuple_count) if self.max_tuple_count else self.rows
        self.bulk0_rows = min(self.bulk0_rows, self.rows)
        self._row0 = 0
        self._bufs = []
        for t, s, c in zip(self.types, self.sizes, self.columns):
            if c is None:
                self._bufs.append(None)
                continue
            raw = np.ascontiguousarray(c).view(np.uint8).reshape(-1)
            if raw.size != self.rows * int(s):
                raise OracleError(f"column has {raw.size} bytes, expected {self.rows}x{s}")
            buf = np.zeros(raw.size + PAD, np.uint8)
            buf[: raw.size] = raw
            self._bufs.append(buf)

    def ptr(self, c):
        b = self._bufs[c]
        return None if b is None else b.ctypes.data + self._row0 * int(self.sizes[c])

    def view(self, row0, rows):
        """Rows [row0, row0+rows) of this table without a copy (a fact slice, as the
        device's zero-copy views are): reads past its last row see the following
        rows of this table, as on the device."""
        if row0 < 0 or row0 + rows > self.rows:
            raise OracleError("view rows exceed the table")
        v = OTable.__new__(OTable)
        v.types, v.sizes = self.types, self.sizes
        v.columns = [None if c is None else c[row0:row0 + rows] for c in self.columns]
        v.max_tuple_count = self.max_tuple_count
        v.rows = rows
        v.bulk0_rows = min(rows, self.max_tuple_count) if self.max_tuple_count else rows
        v._bufs = self._bufs
        v._row0 = self._row0 + row0
        return v


def strings(values, width):
    """Fixed-width NUL-padded byte strings as a (rows, width) uint8 array."""
    out = np.zeros((len(values), width), np.uint8)
    for r, v in enumerate(values):
        b = v.encode() if isinstance(v, str) else bytes(v)
        b = b[:width]
        out[r, : len(b)] = np.frombuffer(b, np.uint8)
    return out


# ------------------------------------------------------------------ plan decoding
def str_split(code: str, sep: str = ","):
    """com/cpp/com.cu:79-93 (an empty string gives no items)."""
    ret, pos1 = [], 0
    pos2 = code.find(sep)
    while pos2 != -1:
        ret.append(code[pos1:pos2])
        pos1 = pos2 + len(sep)
        pos2 = code.find(sep, pos1)
    if pos1 != len(code):
        ret.append(code[pos1:])
    return ret


def _stoi(s: str) -> int:
    s = s.lstrip()
    k = 0
    if k < len(s) and s[k] in "+-":
        k += 1
    while k < len(s) and s[k].isdigit():
        k += 1
    if k == 0 or (k == 1 and s[0] in "+-"):
        raise OracleError(f"bad integer in {s!r}")
    return int(s[:k])


def gen_key_index(exp: str):
    """codegen/codegen.cpp:231-238 (trailing junk after the column is ignored)."""
    p1 = exp.find("$(")
    p2 = exp.find(".", p1)
    if p1 < 0 or p2 < 0:
        raise OracleError(f"bad key {exp!r}")
    return _stoi(exp[p1 + 2:p2]), _stoi(exp[p2 + 1:])


@dataclass
class Decoded:
    filter_exp: list
    select_exp: list
    metric_exp: list
    reduce_key: list
    join_lkey: list
    join_rkey: list
    is_reduce: bool
    source_info: dict
    output_info: dict


def decode(plan: dict) -> Decoded:
    """Pipeline::Pipeline (core/pipeline.cu:11-80)."""
    source_info = output_info = None
    d = Decoded([], [], [], [], [], [], False, None, None)
    for op in plan["workflow"]:
        t = op["type"]
        if t == SOURCE:
            source_info = op.get("output")
        elif t == SINK:
            output_info = op.get("output")
        elif t == LOADER:
            raise OracleError("Loader pipelines are not part of the star-join path")
    for op in plan["workflow"]:
        t, a = op["type"], op.get("args", {})
        if t == MAP:
            d.select_exp = str_split(a["select"], ",")
        elif t == FILTER:
            d.filter_exp.append(a["predication"])
        elif t == REDUCEBY:
            d.metric_exp = str_split(a["metric"], ",")
            if "dim" in a:
                d.reduce_key = [gen_key_index(e) for e in str_split(a["dim"], ",")]
            d.is_reduce = True
        elif t == JOINWITH:
            d.join_lkey.append(gen_key_index(a["lkey"]))
            d.join_rkey.append(gen_key_index(a["rkey"]))
    d.source_info, d.output_info = source_info, output_info
    return d


# ------------------------------------------------------------------ code generation
def _cols(info):
    return [(c["type"], int(c["val_size"])) for c in info["column_list"]]


class _Gen:
    def __init__(self, schemas):
        self.schemas = schemas  # per object: list of (type, size)
        self.used = set()

    def col(self, o, c):
        if o >= len(self.schemas) or c >= len(self.schemas[o]):
            raise OracleError(f"$({o}.{c}) out of range")
        self.used.add((o, c))
        return f"in_{o}_{c}"

    def expr(self, code: str, is_reduce=False) -> str:
        """codegen/codegen.cpp:240-277."""
        pos1 = 0
        while True:
            pos1 = code.find("$(", pos1)
            if pos1 < 0:
                return code
            pos2 = code.find(".", pos1)
            pos3 = code.find(")", pos2)
            if pos2 < 0 or pos3 < 0:
                raise OracleError(f"bad operand in {code!r}")
            o, c = _stoi(code[pos1 + 2:pos2]), _stoi(code[pos2 + 1:pos3])
            var = self.col(o, c)
            t, size = self.schemas[o][c]
            if t in (FLOAT16, FLOAT32, FLOAT64, TENSOR):
                raise OracleError("floating point / tensor operands are out of scope")
            pos = f"{var}+i*{size}" if o == 0 else f"{var}+join_index[{o}]*{size}"
            if t == STRING:
                value = f"Hash({pos},{size})" if is_reduce else f"(const char*)({pos})"
            else:
                value = f"(*(const {NATIVE[t]}*)({pos}))"
            code = code[:pos1] + value + code[pos3 + 1:]
            pos1 += len(value)


def generate(plan: dict, schemas) -> tuple:
    """C++ source of the task for one pipeline.  ``schemas[o]`` lists (type, size)
    for input object o (0 = source, k = k-th ref object)."""
    d = decode(plan)
    g = _Gen(schemas)
    out = _cols(d.output_info)
    body = []
    # GenFilterBlock, CPU variant (codegen/codegen.cpp:100-110)
    cond = "&&".join(g.expr(e) for e in d.filter_exp) if d.filter_exp else "true"
    body.append(f"bool cond = {cond};")
    # GenJoinBlock (codegen/codegen.cpp:121-149) with D4
    builds = []
    seen = set()
    for (lo, lc), (ro, rc) in zip(d.join_lkey, d.join_rkey):
        if ro in seen or ro == 0:
            raise OracleError("each JoinWith must name a distinct ref object")
        seen.add(ro)
        t, size = schemas[ro][rc]
        if t not in (INT16, INT32, INT64, STRING):
            raise OracleError("join key type unsupported")
        ty = NATIVE[t]
        li, ri = g.col(lo, lc), g.col(ro, rc)
        body.append(
            f"if(cond) {{ join_index[{ro}] = Probe<{ty}>((const {ty}*)({li}+{size}*i),{size},"
            f"(const {ty}*)({ri}), ctx->ht[{ro}], (unsigned)ctx->ht_size[{ro}]);"
            f" cond = cond & (join_index[{ro}] >=0 ? true:false); }}")
        builds.append(f"ort_build_join<{ty}>(c, {ro}, {rc});")
    if d.select_exp and d.is_reduce:
        raise OracleError("a pipeline with both Map and ReduceBy is not well defined")
    if d.select_exp:
        kind, nk, nm = 0, 0, 0
        widths = []
        blk = []
        for q, e in enumerate(d.select_exp):
            t, size = out[q]
            ex = g.expr(e)
            if t == STRING:
                blk.append(f"StrCopy(out->col[{q}]+j*{size},{ex},{size});")
            elif t in NATIVE and t not in (FLOAT32, FLOAT64, TENSOR):
                blk.append(f"*({NATIVE[t]}*)(out->col[{q}]+j*{size}) = {ex};")
            else:
                raise OracleError("projection type unsupported")
            widths.append(size)
        body.append("if(cond){ " + " ".join(blk) + " }")
        body.append("j+=cond;")
    elif d.is_reduce:
        nk, nm = len(d.reduce_key), len(d.metric_exp)
        kind = 1 if nk > 0 else 2
        if nk > 3 or (nk > 0 and nm > 3) or nm > 8 or nm == 0:
            raise OracleError("ReduceBy supports 1-3 keys x 1-3 metrics, or a scalar")
        widths = []
        blk = []
        for q, (o, c) in enumerate(d.reduce_key):
            if out[q][0] != INT32:
                raise OracleError("reduce key columns must be declared Int32")
            blk.append(f"out->key[{q}][j] = (int32_t)({g.expr(f'$({o}.{c})', True)});")
        for q, e in enumerate(d.metric_exp):
            t, _ = out[nk + q]
            if t not in (INT16, INT32, INT64, BYTE):
                raise OracleError("metric columns must be integer")
            val = f"(int64_t)({NATIVE[t]})({g.expr(e)})"
            if kind == 2:
                blk.append(f"out->met[{q}][0] += {val};")
            else:
                blk.append(f"out->met[{q}][j] = {val};")
        body.append("if(cond){ " + " ".join(blk) + " }")
        body.append("j+=cond;")
    else:
        raise OracleError("pipeline has neither Map nor ReduceBy")
    decls = []
    for (o, c) in sorted(g.used):
        if o == 0:
            decls.append(f"const char* in_0_{c} = fact[{c}];")
        else:
            decls.append(f"const char* in_{o}_{c} = ctx->obj[{o}].col[{c}];")
    wl = ", ".join(str(w) for w in widths) or "0"
    src = f"""#include "svm_oracle_rt.h"
static uint64 task(const ORT_Ctx* ctx, const char* const* fact, uint64 tuple_ct, ORT_BulkOut* out) {{
  (void)ctx; (void)fact;
  {' '.join(decls)}
  uint64 j = 0;
  int join_index[8];
  (void)join_index;
  for (uint64 i = 0; i < tuple_ct; i++) {{
    {chr(10).join('    ' + b for b in body)}
  }}
  return j;
}}
extern "C" int oracle_open(ORT_Ctx* c) {{
  auto t0 = std::chrono::steady_clock::now();
  {' '.join(builds)}
  c->open_ms = ort_ms_since(t0);
  return 0;
}}
extern "C" int oracle_run(ORT_Ctx* c, ORT_Result* r) {{
  static const int64_t w[] = {{{wl}}};
  return ort_drive(c, r, task, {kind}, {nk}, {nm}, {len(widths)}, w);
}}
extern "C" void oracle_free(ORT_Ctx* c, ORT_Result* r) {{ ort_free_result(r); ort_free_ctx(c); }}
"""
    return src, kind, nk, nm, d


# ------------------------------------------------------------------ compile + ctypes
class _Obj(ctypes.Structure):
    _fields_ = [("col", ctypes.c_void_p * MAX_COL), ("rows", ctypes.c_uint64),
                ("max_tuple_ct", ctypes.c_uint64), ("ncols", ctypes.c_int32),
                ("type", ctypes.c_int32 * MAX_COL), ("size", ctypes.c_int64 * MAX_COL)]


class _Ctx(ctypes.Structure):
    _fields_ = [("nobj", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("tuple_per_bulk", ctypes.c_uint64), ("obj", _Obj * MAX_OBJ),
                ("ht", ctypes.c_void_p * MAX_OBJ), ("ht_size", ctypes.c_uint64 * MAX_OBJ),
                ("dropped", ctypes.c_uint64 * MAX_OBJ), ("inserted", ctypes.c_uint64 * MAX_OBJ),
                ("open_ms", ctypes.c_double), ("run_ms", ctypes.c_double),
                ("merge_ms", ctypes.c_double)]


class _Res(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("ncols", ctypes.c_int32), ("rows", ctypes.c_uint64),
                ("col", ctypes.c_void_p * MAX_COL), ("width", ctypes.c_int64 * MAX_COL),
                ("bulk0_rows", ctypes.c_uint64)]


_lock = threading.Lock()
_cache = {}


def jit_dir():
    d = os.environ.get("SVM_ORACLE_JIT") or os.path.join(tempfile.gettempdir(), "svm_oracle_jit")
    os.makedirs(d, exist_ok=True)
    return d


def compile_source(src: str):
    """g++ -O2 -shared (the reference JIT uses clang -O2 -fPIC -shared,
    codegen/compiler.cpp:30-40)."""
    key = hashlib.sha1((src + open(os.path.join(HERE, "svm_oracle_rt.h")).read()).encode()).hexdigest()[:16]
    with _lock:
        if key in _cache:
            return _cache[key]
    # compiled outside the lock, so that threads running different plans compile
    # concurrently (two threads with the same plan both compile; the rename is atomic)
    d = jit_dir()
    so = os.path.join(d, f"q_{key}.so")
    if not os.path.exists(so):
        tag = f"{os.getpid()}_{threading.get_ident()}"
        cpp = os.path.join(d, f"q_{key}_{tag}.cpp")
        with open(cpp, "w") as f:
            f.write(src)
        tmp = so + f".tmp{tag}"
        # -fwrapv: signed int32 overflow in a plan's expressions wraps, as in the
        # generated GPU kernels (the reference's clang / nvcc builds wrap in practice)
        r = subprocess.run(["g++", "-O2", "-fwrapv", "-std=c++17", "-fPIC", "-shared", "-pthread",
                            "-I", HERE, cpp, "-o", tmp], capture_output=True, text=True)
        os.unlink(cpp)
        if r.returncode != 0:
            raise OracleError("oracle compile failed:\n" + r.stderr[-4000:])
        os.replace(tmp, so)
    with _lock:
        if key in _cache:
            return _cache[key]
        lib = ctypes.CDLL(so)
        lib.oracle_open.argtypes = [ctypes.POINTER(_Ctx)]
        lib.oracle_run.argtypes = [ctypes.POINTER(_Ctx), ctypes.POINTER(_Res)]
        lib.oracle_free.argtypes = [ctypes.POINTER(_Ctx), ctypes.POINTER(_Res)]
        _cache[key] = lib
        return lib


@dataclass
class OResult:
    kind: str            # "select" | "group" | "scalar"
    columns: list        # numpy arrays (group: K int32 keys then M int64 metrics)
    rows: int
    dropped: dict        # ref object -> dropped inserts
    inserted: dict
    open_ms: float
    run_ms: float
    merge_ms: float
    table: OTable = None  # select pipelines: the output object


def run_pipeline(plan: dict, source: OTable, refs=(), tuple_per_bulk=3_000_000, threads=None):
    """PipelineTest::Run semantics (test/test.hpp:122-146) on the CPU oracle."""
    refs = list(refs)
    schemas = [list(zip(source.types, source.sizes))] + [list(zip(r.types, r.sizes)) for r in refs]
    if len(schemas) > MAX_OBJ:
        raise OracleError("at most 7 ref objects")
    src, kind, nk, nm, d = generate(plan, schemas)
    lib = compile_source(src)
    ctx = _Ctx()
    ctx.nobj = len(schemas)
    ctx.threads = threads or os.cpu_count() or 1
    ctx.tuple_per_bulk = tuple_per_bulk
    for o, t in enumerate([source] + refs):
        ob = ctx.obj[o]
        ob.rows = t.rows if o == 0 else t.bulk0_rows  # a ref object is read through bulk 0 only
        ob.max_tuple_ct = t.max_tuple_count
        ob.ncols = len(t.types)
        for c in range(len(t.types)):
            ob.col[c] = t.ptr(c)
            ob.type[c] = t.types[c]
            ob.size[c] = t.sizes[c]
    res = _Res()
    lib.oracle_open(ctypes.byref(ctx))
    lib.oracle_run(ctypes.byref(ctx), ctypes.byref(res))
    try:
        cols = []
        for q in range(res.ncols):
            w = int(res.width[q])
            n = int(res.rows)
            raw = np.ctypeslib.as_array(ctypes.cast(res.col[q], ctypes.POINTER(ctypes.c_uint8)),
                                        shape=(max(n * w, 1),))[: n * w].copy()
            cols.append(raw)
        kinds = {0: "select", 1: "group", 2: "scalar"}
        out_cols = []
        if kind == 0:
            out = _cols(d.output_info)
            for q, raw in enumerate(cols):
                t, size = out[q]
                if t == STRING:
                    out_cols.append(raw.reshape(-1, size))
                else:
                    out_cols.append(raw.view(NP_DTYPE[t]))
        else:
            for q, raw in enumerate(cols):
                out_cols.append(raw.view(np.int32 if (kind == 1 and q < nk) else np.int64))
        dropped = {o: int(ctx.dropped[o]) for (_, _), (o, _) in zip(d.join_lkey, d.join_rkey)}
        inserted = {o: int(ctx.inserted[o]) for (_, _), (o, _) in zip(d.join_lkey, d.join_rkey)}
        result = OResult(kinds[kind], out_cols, int(res.rows), dropped, inserted,
                         ctx.open_ms, ctx.run_ms, ctx.merge_ms)
        if kind == 0:
            out = _cols(d.output_info)
            tcols = out_cols + [None] * (len(out) - len(out_cols))
            result.table = OTable([t for t, _ in out], [s for _, s in out], tcols,
                                  max_tuple_count=tuple_per_bulk, rows=int(res.rows),
                                  bulk0_rows=int(res.bulk0_rows))
        return result
    finally:
        lib.oracle_free(ctypes.byref(ctx), ctypes.byref(res))
