"""Host-side operator surface of the data plane, mirroring the reference's
``PipelineTest`` harness (test/test.hpp:78-149) and ``CreateSchema``
(test/test.hpp:17-27): build a ``PipelineInfo`` from Source / Filter / Map /
JoinWith / ReduceBy / Sink operators, then run it against objects held in the
device store.  Execution happens in libsvm.so on the GPU; this module only
assembles plans and moves columns.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .proto import serialize

# ValType (com/svm.proto:62-72)
BYTE, INT16, INT32, INT64, FLOAT16, FLOAT32, FLOAT64, STRING, TENSOR = range(9)
TInt32, TString = INT32, STRING
# OperatorInfo.Type (com/svm.proto:102-111)
SOURCE, LOADER, FILTER, MAP, REDUCEBY, SORTBY, JOINWITH, SINK = range(8)

NP_DTYPE = {BYTE: np.int8, INT16: np.int16, INT32: np.int32, INT64: np.int64}
HashRate = 5  # com/cpp/define.hpp:29


def create_schema(types, sizes):
    """CreateSchema (test/test.hpp:17-27): an ObjectInfo with the given columns."""
    if len(types) != len(sizes):
        raise ValueError("types and sizes differ in length")
    return {"column_list": [{"type": int(t), "val_size": int(s)} for t, s in zip(types, sizes)],
            "storage": 1}


# ---------------------------------------------------------------- engine
def init(device=0, tuple_per_bulk=3_000_000, hbm_budget=0, num_streams=8, verbose=0):
    cfg = _lib.SvmConfig(device, tuple_per_bulk, hbm_budget, num_streams, verbose)
    _lib.check(_lib.lib().svm_init(ctypes.byref(cfg)))


def shutdown():
    _lib.check(_lib.lib().svm_shutdown())


def set_hybrid(enabled, cpu_ratio=5, cpu_threads=16):
    """Hybrid execution (ResMgr::enable_hybrid_exec_ / CPU_ratio_ / CPU_num_): bulk i of
    a GPU ReduceBy pipeline runs on host cores when i % 10 < cpu_ratio."""
    _lib.check(_lib.lib().svm_set_hybrid(1 if enabled else 0, cpu_ratio, cpu_threads))


def table_create(id, types, sizes, rows):
    n = len(types)
    t = (ctypes.c_int32 * n)(*types)
    s = (ctypes.c_int64 * n)(*sizes)
    _lib.check(_lib.lib().svm_table_create(id.encode(), n, t, s, rows))


def table_put(id, col, array):
    a = np.ascontiguousarray(array)
    _lib.check(_lib.lib().svm_table_put_column(id.encode(), col, a.ctypes.data, a.nbytes))


def table_upload(id, types, sizes, columns):
    """Create table `id` and copy host columns (int arrays, or (rows, size) uint8 for strings)."""
    rows = int(columns[0].shape[0]) if columns else 0
    table_create(id, types, sizes, rows)
    for c, a in enumerate(columns):
        if a is not None:
            table_put(id, c, a)


def table_info(id):
    rows = ctypes.c_uint64()
    ncols = ctypes.c_int()
    _lib.check(_lib.lib().svm_table_info(id.encode(), ctypes.byref(rows), ctypes.byref(ncols)))
    return int(rows.value), int(ncols.value)


def table_column(id, col):
    rows, _ = table_info(id)
    t, s = ctypes.c_int32(), ctypes.c_int64()
    _lib.check(_lib.lib().svm_table_column_type(id.encode(), col, ctypes.byref(t), ctypes.byref(s)))
    buf = np.zeros(rows * s.value, np.uint8)
    _lib.check(_lib.lib().svm_table_get_column(id.encode(), col, buf.ctypes.data, buf.nbytes))
    return _view(buf, t.value, s.value)


def table_view(id, base_id, row0, rows):
    """Zero-copy view of rows [row0, row0+rows) of `base_id`."""
    _lib.check(_lib.lib().svm_table_view(id.encode(), base_id.encode(), row0, rows))


def table_column_device(id, col):
    """Device address of column `col` of table `id`."""
    p = ctypes.c_void_p()
    _lib.check(_lib.lib().svm_table_column_device(id.encode(), col, ctypes.byref(p)))
    return p.value


def run_noalloc(plan, out_device_ptrs, max_rows):
    """svm_pipeline_run_noalloc: run a ReduceBy pipeline without allocating, writing its
    K key columns (int32) and M sums (int64) into the given device buffers of
    `max_rows` rows each; returns the number of result rows."""
    data = _bytes(plan)
    buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
    n = len(out_device_ptrs)
    ptrs = (ctypes.c_void_p * n)(*out_device_ptrs)
    rows = ctypes.c_uint64()
    _lib.check(_lib.lib().svm_pipeline_run_noalloc(buf, len(data), ptrs, n, max_rows, ctypes.byref(rows)))
    return int(rows.value)


def batch_run(plans, want_result=True):
    """svm_batch_run over plan dicts; returns (results list, number fused).
    Absorbed dimension pipelines give None."""
    arr, lens, keep = _plan_array(plans)
    n = len(keep[2])
    outs = (ctypes.c_void_p * n)()
    fused = _lib.check(_lib.lib().svm_batch_run(n, arr, lens, outs if want_result else None))
    return (_take_results(outs, n) if want_result else None), fused


def _take_results(outs, n):
    """Results of a batch as numpy arrays (None where outs[i] is NULL), read with one
    describe, one copy and one free call for the whole batch."""
    L = _lib.lib()
    try:
        descs = (_lib.SvmResultDesc * n)()
        _lib.check(L.svm_results_describe(n, outs, descs))
        a = np.frombuffer(descs, dtype=np.dtype(_lib.SvmResultDesc))
        idx = np.flatnonzero(a["kind"] >= 0)
        d = a[idx]
        mc = int(d["ncols"].max()) if len(idx) else 0
        # bytes of each column (0 past a result's last column); each starts at a
        # multiple of 8 in the copy, results in order, columns in order
        nb = np.where(np.arange(mc)[None, :] < d["ncols"][:, None], d["rows"][:, None] * d["val_size"][:, :mc], 0)
        padded = (nb + 7) & ~7
        offs = (np.cumsum(padded.ravel()) - padded.ravel()).reshape(padded.shape).tolist()
        total = int(padded.sum())
        buf = np.empty(total + 8, np.uint8)
        _lib.check(L.svm_results_copy(n, outs, buf.ctypes.data, total + 8))
        kinds, ncols, rows = d["kind"].tolist(), d["ncols"].tolist(), d["rows"].tolist()
        types, sizes, nb = d["type"][:, :mc].tolist(), d["val_size"][:, :mc].tolist(), nb.tolist()
        bms, sms, fms = d["build_ms"].tolist(), d["scan_ms"].tolist(), d["finish_ms"].tolist()
        res = [None] * n
        for j, i in enumerate(idx.tolist()):
            cols = [_view(buf[offs[j][c]:offs[j][c] + nb[j][c]], types[j][c], sizes[j][c]) for c in range(ncols[j])]
            res[i] = Result(_KINDS[kinds[j]], rows[j], cols, bms[j], sms[j], fms[j])
        return res
    finally:
        L.svm_results_free(n, outs)


def _bytes(plan):
    """PipelineInfo bytes of a plan dict (bytes pass through unchanged)."""
    return bytes(plan) if isinstance(plan, (bytes, bytearray)) else serialize(plan)


def _plan_array(plans):
    """(pointer array, length array, objects to keep alive) for a list of plans: the
    bytes go into one buffer, and the pointers are computed with numpy."""
    datas = [_bytes(p) for p in plans]
    n = len(datas)
    blob = np.frombuffer(b"".join(datas) + b"\0", np.uint8)
    lens = np.fromiter((len(d) for d in datas), np.uint64, n)
    addrs = np.zeros(n, np.uint64)
    if n > 1:
        np.cumsum(lens[:-1], out=addrs[1:])
    addrs += np.uint64(blob.ctypes.data)
    arr = addrs.ctypes.data_as(ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)))
    return arr, lens.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)), (blob, addrs, datas, lens)


def server_start(workers=2, budget=0, max_batch=64, mode=0):
    """Serving loop (mode 0 continuous batching, mode 1 one query per plain launch)."""
    _lib.check(_lib.lib().svm_server_start(workers, budget, max_batch, mode))


def server_submit(plans):
    """Submit one query (its dimension pipelines, then the fact pipeline); returns a ticket."""
    arr, lens, keep = _plan_array(plans)
    t = ctypes.c_uint64()
    _lib.check(_lib.lib().svm_server_submit(arr, lens, len(plans), ctypes.byref(t)))
    return int(t.value)


def server_wait(ticket, timeout_ms=-1):
    out = ctypes.c_void_p()
    _lib.check(_lib.lib().svm_server_wait(ticket, timeout_ms, ctypes.byref(out)))
    return _take_result(out) if out.value else None


def server_wait_timed(ticket, timeout_ms=-1, want_result=True):
    """svm_server_wait_timed: (result or None, queued ms, total ms), the latencies
    measured by the server from the query's submission."""
    out = ctypes.c_void_p()
    q, t = ctypes.c_double(), ctypes.c_double()
    _lib.check(_lib.lib().svm_server_wait_timed(ticket, timeout_ms, ctypes.byref(out) if want_result else None,
                                                ctypes.byref(q), ctypes.byref(t)))
    res = _take_result(out) if (want_result and out.value) else None
    return res, q.value, t.value


def server_stats():
    v = [ctypes.c_uint64() for _ in range(4)]
    _lib.check(_lib.lib().svm_server_stats(*[ctypes.byref(x) for x in v]))
    return dict(zip(("queued", "running", "done", "bytes_in_use"), (x.value for x in v)))


def server_stop():
    _lib.check(_lib.lib().svm_server_stop())


def hash_drops(id, col, slots=0):
    """Inserts the reference's join hash table drops over column `col` of table `id`
    inserted in row order into `slots` slots (0: HashRate x bulk capacity)."""
    d = ctypes.c_uint64()
    _lib.check(_lib.lib().svm_debug_hash_drops(id.encode(), col, slots, ctypes.byref(d)))
    return int(d.value)


def table_drop(id):
    _lib.check(_lib.lib().svm_table_drop(id.encode()))


def table_exists(id):
    return bool(_lib.lib().svm_table_exists(id.encode()))


def gen_params(n_cust=0, n_supp=0, n_part=0, zipf_s=1.1):
    return _lib.SvmGenParams(n_cust, n_supp, n_part, zipf_s)


def gen_schema(kind):
    t = (ctypes.c_int32 * 8)()
    s = (ctypes.c_int64 * 8)()
    n = _lib.check(_lib.lib().svm_gen_schema(kind, t, s))
    return list(t[:n]), list(s[:n])


def gen_table(id, kind, rows, seed, params=None):
    p = params or gen_params()
    _lib.check(_lib.lib().svm_gen_table(id.encode(), kind, rows, seed, ctypes.byref(p)))


def gen_host(kind, seed, row0, n, params=None, threads=8):
    """The same rows a device table of `kind` holds, generated on the host."""
    types, sizes = gen_schema(kind)
    cols = [np.zeros(n * s + 1, np.uint8) for s in sizes]
    ptrs = (ctypes.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
    p = params or gen_params()
    _lib.check(_lib.lib().svm_gen_host(kind, seed, ctypes.byref(p), row0, n, ptrs, threads))
    return types, sizes, [_view(c[: n * s], t, s) for c, t, s in zip(cols, types, sizes)]


def _view(buf, t, s):
    if t == STRING or t == TENSOR:
        return buf.reshape(-1, s)
    return buf.view(NP_DTYPE[t])


# ---------------------------------------------------------------- results
@dataclass
class Result:
    kind: str          # "select" | "group" | "scalar"
    rows: int
    columns: list      # numpy arrays (group-by: K int32 keys then M int64 sums)
    build_ms: float
    scan_ms: float
    finish_ms: float


_KINDS = {0: "select", 1: "group", 2: "scalar"}


def _take_result(ptr) -> Result:
    """Copy a result to numpy arrays and free it (svm_result_describe: one call)."""
    L = _lib.lib()
    try:
        d = _lib.SvmResultDesc()
        _lib.check(L.svm_result_describe(ptr, ctypes.byref(d)))
        rows = int(d.rows)
        cols = []
        for c in range(d.ncols):
            t, s = int(d.type[c]), int(d.val_size[c])
            nbytes = rows * s
            buf = np.empty(nbytes, np.uint8)
            if nbytes and d.col[c]:
                ctypes.memmove(buf.ctypes.data, d.col[c], nbytes)
            cols.append(_view(buf, t, s))
        return Result(_KINDS[int(d.kind)], rows, cols, d.build_ms, d.scan_ms, d.finish_ms)
    finally:
        L.svm_result_free(ptr)


def run_plan(plan: dict, want_result=True) -> Result | None:
    """Serialize a plan dict to PipelineInfo bytes (or take the bytes) and run it
    (svm_pipeline_run)."""
    data = _bytes(plan)
    buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
    out = ctypes.c_void_p()
    _lib.check(_lib.lib().svm_pipeline_run(buf, len(data), ctypes.byref(out) if want_result else None))
    return _take_result(out) if want_result else None


class PipelineTest:
    """test/test.hpp:78-149: AddOperator / SetMode / Run against cached objects."""

    def __init__(self, id):
        self.pipe_info = {"id": id, "workflow": [], "pre_pipeline_id": [],
                          "codegen": True, "debug": False, "isGPU": True}

    def AddOperator(self, type, args, output):
        self.pipe_info["workflow"].append(
            {"type": int(type), "id": self.pipe_info["id"], "args": dict(args), "output": output})
        return self

    def SetMode(self, codegen, debug, isGPU=True):
        self.pipe_info.update(codegen=codegen, debug=debug, isGPU=isGPU)
        return self

    def plan(self, input_obj, ref_obj=()):
        """The PipelineInfo dict; a Loader pipeline (file Source) takes input_obj None."""
        p = dict(self.pipe_info)
        p["pre_pipeline_id"] = ([] if input_obj is None else [input_obj]) + list(ref_obj)
        return p

    def Run(self, input_obj, ref_obj=(), codegen=True, print_=False, isGPU=True, want_result=True):
        """Run against the objects named `input_obj` (source) and `ref_obj` (refs);
        the output is cached in the store under this pipeline's id."""
        self.pipe_info.update(codegen=codegen, debug=print_, isGPU=isGPU)
        return run_plan(self.plan(input_obj, ref_obj), want_result)
