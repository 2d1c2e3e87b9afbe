"""Seeded synthetic star-schema workloads: tables, dimension pipelines and the
fact pipeline of each configuration, written with the reference's operator
surface (the shapes of test/test-SSB-Q11.cu:13-44, test-SSB-Q21.cu:17-74 and
test-SSB-Q41.cu:17-85).

Configurations
  single_join  : date 2,556 rows; fact 6M x (orderdate, quantity, discount, extendedprice).
                 Filter discount in [d, d+2] and quantity < q, join date filtered to one year,
                 SUM(extendedprice*discount) as one int64 scalar.
  three_way    : part 800K (category 1..25, brand 1..1000), supplier 20K (region/nation/city codes),
                 fact 60M x (orderdate, partkey, suppkey, revenue); part.category == c,
                 supplier.region == r; reduceBy (d_year, p_brand), SUM(revenue).
  four_way     : customer 3M / supplier 200K with 10-byte region/nation/city strings, part 1.4M,
                 date; fact 600M x (orderdate, custkey, suppkey, partkey, revenue, supplycost)
                 with Zipf(1.1) foreign keys; c_nation == x, s_nation == y, d_year in [Y, Y+1];
                 reduceBy (d_year, c_city, s_city), SUM(revenue - supplycost).
"""
from __future__ import annotations

from dataclasses import dataclass, field

from . import pipeline as P

# svm_gen_kind (shadowvm_amd/csrc/datagen.h)
GEN_DATE, GEN_FACT1, GEN_PART2, GEN_SUPP2, GEN_FACT2, GEN_CUST3, GEN_SUPP3, GEN_PART3, GEN_FACT3 = range(9)

I, S = P.INT32, P.STRING


def _mix(seed, salt):
    z = (seed * 0x9E3779B97F4A7C15 + salt * 0xD1B54A32D192ED03) & (2**64 - 1)
    z ^= z >> 31
    z = (z * 0xBF58476D1CE4E5B9) & (2**64 - 1)
    return z ^ (z >> 29)


@dataclass
class TableSpec:
    id: str
    kind: int
    rows: int
    seed: int


@dataclass
class Workload:
    name: str
    params: tuple           # (n_cust, n_supp, n_part, zipf_s)
    tables: list            # TableSpec
    dims: list              # (PipelineTest, source id, ref ids) run before the fact pipeline
    fact: tuple             # (PipelineTest, source id, ref ids)
    constants: dict = field(default_factory=dict)

    def gen_params(self):
        return P.gen_params(*self.params)

    @property
    def fact_rows(self):
        return next(t.rows for t in self.tables if t.kind in (GEN_FACT1, GEN_FACT2, GEN_FACT3))

    def plans(self):
        """All pipeline plans in execution order (dimension pipelines, then the fact pipeline)."""
        return [pt.plan(src, refs) for pt, src, refs in self.dims] + [self.fact[0].plan(self.fact[1], self.fact[2])]


def _op(pt, t, args, out):
    pt.AddOperator(t, args, out)


def single_join(seed=1, fact_rows=6_000_000, prefix="sj"):
    r = _mix(seed, 1)
    year = 1992 + r % 7
    d = (r >> 8) % 9
    q = 20 + (r >> 16) % 11
    date_id, fact_id = f"{prefix}_date", f"{prefix}_lineorder"
    tables = [TableSpec(date_id, GEN_DATE, 2556, 0), TableSpec(fact_id, GEN_FACT1, fact_rows, _mix(seed, 11))]
    dsch = P.create_schema([I, I, S], [4, 4, 40])
    dq = P.PipelineTest(f"{prefix}_query_date")
    _op(dq, P.SOURCE, {"type": "object"}, P.create_schema([I, I, I], [4, 4, 4]))
    _op(dq, P.FILTER, {"predication": f"$(0.1)=={year}"}, dsch)
    _op(dq, P.MAP, {"select": "$(0.0), $(0.1)"}, dsch)
    _op(dq, P.SINK, {}, dsch)
    res = P.create_schema([I] * 9, [4] * 9)
    fq = P.PipelineTest(f"{prefix}_query")
    _op(fq, P.SOURCE, {"type": "object"}, P.create_schema([I] * 4, [4] * 4))
    _op(fq, P.FILTER, {"predication": f" $(0.2)>={d} && $(0.2)<={d + 2} && $(0.1)<{q} "}, res)
    _op(fq, P.JOINWITH, {"lkey": "$(0.0)", "rkey": "$(1.0)"}, res)
    _op(fq, P.REDUCEBY, {"dim": "", "metric": "$(0.3)*$(0.2)"}, res)
    _op(fq, P.SINK, {}, res)
    return Workload("single_join", (0, 0, 0, 1.1), tables, [(dq, date_id, [])],
                    (fq, fact_id, [dq.pipe_info["id"]]), {"year": year, "discount": d, "quantity": q})


def three_way(seed=1, fact_rows=60_000_000, n_part=800_000, n_supp=20_000, prefix="tw"):
    r = _mix(seed, 2)
    cat = 1 + r % 25
    region = (r >> 8) % 5
    ids = {k: f"{prefix}_{k}" for k in ("date", "part", "supplier", "lineorder")}
    tables = [TableSpec(ids["date"], GEN_DATE, 2556, 0),
              TableSpec(ids["part"], GEN_PART2, n_part, _mix(seed, 21)),
              TableSpec(ids["supplier"], GEN_SUPP2, n_supp, _mix(seed, 22)),
              TableSpec(ids["lineorder"], GEN_FACT2, fact_rows, _mix(seed, 23))]
    ht = P.HashRate * (4 + 32)
    dsch = P.create_schema([I, I, S], [4, 4, ht])
    dq = P.PipelineTest(f"{prefix}_query_date")
    _op(dq, P.SOURCE, {"type": "object"}, P.create_schema([I] * 3, [4] * 3))
    _op(dq, P.MAP, {"select": "$(0.0), $(0.1)"}, dsch)
    _op(dq, P.SINK, {}, dsch)
    psch = P.create_schema([I, I, S], [4, 4, ht])
    pq = P.PipelineTest(f"{prefix}_query_part")
    _op(pq, P.SOURCE, {"type": "object"}, P.create_schema([I] * 3, [4] * 3))
    _op(pq, P.FILTER, {"predication": f"$(0.1)=={cat} "}, psch)
    _op(pq, P.MAP, {"select": "$(0.0), $(0.2)"}, psch)
    _op(pq, P.SINK, {}, psch)
    ssch = P.create_schema([I, S], [4, ht])
    sq = P.PipelineTest(f"{prefix}_query_supplier")
    _op(sq, P.SOURCE, {"type": "object"}, P.create_schema([I] * 4, [4] * 4))
    _op(sq, P.FILTER, {"predication": f"$(0.1)=={region}"}, ssch)
    _op(sq, P.MAP, {"select": "$(0.0) "}, ssch)
    _op(sq, P.SINK, {}, ssch)
    res = P.create_schema([I] * 9, [4] * 9)
    fq = P.PipelineTest(f"{prefix}_query")
    _op(fq, P.SOURCE, {"type": "object"}, P.create_schema([I] * 4, [4] * 4))
    _op(fq, P.JOINWITH, {"lkey": "$(0.0)", "rkey": "$(1.0)"}, res)
    _op(fq, P.JOINWITH, {"lkey": "$(0.1)", "rkey": "$(2.0)"}, res)
    _op(fq, P.JOINWITH, {"lkey": "$(0.2)", "rkey": "$(3.0)"}, res)
    _op(fq, P.REDUCEBY, {"dim": "$(1.1), $(2.1)", "metric": "$(0.3)"}, res)
    _op(fq, P.SINK, {}, res)
    dims = [(dq, ids["date"], []), (pq, ids["part"], []), (sq, ids["supplier"], [])]
    fact = (fq, ids["lineorder"], [dq.pipe_info["id"], pq.pipe_info["id"], sq.pipe_info["id"]])
    return Workload("three_way", (0, n_supp, n_part, 1.1), tables, dims, fact,
                    {"category": cat, "region": region})


def four_way(seed=1, fact_rows=600_000_000, n_cust=3_000_000, n_supp=200_000, n_part=1_400_000,
             zipf_s=1.1, prefix="fw"):
    r = _mix(seed, 3)
    year = 1992 + r % 6
    cn = (r >> 8) % 25
    sn = (r >> 16) % 25
    ids = {k: f"{prefix}_{k}" for k in ("date", "customer", "supplier", "part", "lineorder")}
    tables = [TableSpec(ids["date"], GEN_DATE, 2556, 0),
              TableSpec(ids["customer"], GEN_CUST3, n_cust, _mix(seed, 31)),
              TableSpec(ids["supplier"], GEN_SUPP3, n_supp, _mix(seed, 32)),
              TableSpec(ids["part"], GEN_PART3, n_part, _mix(seed, 33)),
              TableSpec(ids["lineorder"], GEN_FACT3, fact_rows, _mix(seed, 34))]
    dsch = P.create_schema([I, I, S], [4, 4, 40])
    dq = P.PipelineTest(f"{prefix}_query_date")
    _op(dq, P.SOURCE, {"type": "object"}, P.create_schema([I] * 3, [4] * 3))
    _op(dq, P.FILTER, {"predication": f"$(0.1)>={year} && $(0.1)<={year + 1}"}, dsch)
    _op(dq, P.MAP, {"select": "$(0.0), $(0.1)"}, dsch)
    _op(dq, P.SINK, {}, dsch)
    geo = P.create_schema([I, S, S, S], [4, 10, 10, 10])
    csch = P.create_schema([I, S, S], [4, 10, 40])
    cq = P.PipelineTest(f"{prefix}_query_customer")
    _op(cq, P.SOURCE, {"type": "object"}, geo)
    _op(cq, P.FILTER, {"predication": f'StrCmp($(0.2),"NATION_{cn:02d}",10)==0 '}, csch)
    _op(cq, P.MAP, {"select": "$(0.0),$(0.1)"}, csch)
    _op(cq, P.SINK, {}, csch)
    sq = P.PipelineTest(f"{prefix}_query_supplier")
    _op(sq, P.SOURCE, {"type": "object"}, geo)
    _op(sq, P.FILTER, {"predication": f'StrCmp($(0.2),"NATION_{sn:02d}",10)==0 '}, csch)
    _op(sq, P.MAP, {"select": "$(0.0),$(0.1)"}, csch)
    _op(sq, P.SINK, {}, csch)
    psch = P.create_schema([I, S], [4, 40])
    pq = P.PipelineTest(f"{prefix}_query_part")
    _op(pq, P.SOURCE, {"type": "object"}, P.create_schema([I] * 3, [4] * 3))
    _op(pq, P.MAP, {"select": "$(0.0)"}, psch)
    _op(pq, P.SINK, {}, psch)
    res = P.create_schema([I] * 12, [4] * 12)
    fq = P.PipelineTest(f"{prefix}_query")
    _op(fq, P.SOURCE, {"type": "object"}, P.create_schema([I] * 6, [4] * 6))
    _op(fq, P.JOINWITH, {"lkey": "$(0.0)", "rkey": "$(1.0)"}, res)
    _op(fq, P.JOINWITH, {"lkey": "$(0.1)", "rkey": "$(2.0)"}, res)
    _op(fq, P.JOINWITH, {"lkey": "$(0.2)", "rkey": "$(3.0)"}, res)
    _op(fq, P.JOINWITH, {"lkey": "$(0.3)", "rkey": "$(4.0)"}, res)
    _op(fq, P.REDUCEBY, {"dim": "$(1.1), $(2.1), $(3.1)", "metric": "$(0.4)-$(0.5)"}, res)
    _op(fq, P.SINK, {}, res)
    dims = [(dq, ids["date"], []), (cq, ids["customer"], []), (sq, ids["supplier"], []), (pq, ids["part"], [])]
    fact = (fq, ids["lineorder"], [dq.pipe_info["id"], cq.pipe_info["id"], sq.pipe_info["id"], pq.pipe_info["id"]])
    return Workload("four_way", (n_cust, n_supp, n_part, zipf_s), tables, dims, fact,
                    {"year": year, "c_nation": cn, "s_nation": sn})


def load_device(wl: Workload):
    """Generate every table of the workload directly in HBM."""
    for t in wl.tables:
        if not P.table_exists(t.id):
            P.gen_table(t.id, t.kind, t.rows, t.seed, wl.gen_params())


def host_table(wl: Workload, t: TableSpec, threads=8, row0=0, n=None):
    """(types, sizes, columns) of table `t` generated on the host (same bytes as HBM)."""
    n = t.rows if n is None else n
    return P.gen_host(t.kind, t.seed, row0, n, wl.gen_params(), threads)


def run_dims(wl: Workload):
    for pt, src, refs in wl.dims:
        pt.Run(src, refs, want_result=False)


def run_fact(wl: Workload, want_result=True):
    pt, src, refs = wl.fact
    return pt.Run(src, refs, want_result=want_result)


def run_query(wl: Workload):
    """Dimension pipelines then the fact pipeline; returns the fact pipeline's result."""
    run_dims(wl)
    return run_fact(wl)
