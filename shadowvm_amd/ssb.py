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


@dataclass
class Query:
    """One star-join query of a batch: its dimension pipelines, the fact pipeline
    over a zero-copy slice (view) of the fact table, and the slice itself."""
    qid: str
    view: tuple           # (view id, base fact id, row0, rows)
    dims: list            # (PipelineTest, source id, ref ids)
    fact: tuple           # (PipelineTest, source id, ref ids)
    shape: dict

    def plans(self):
        return [pt.plan(src, refs) for pt, src, refs in self.dims] + [self.fact[0].plan(self.fact[1], self.fact[2])]


# Dimension choices of the template generator over the four-way tables:
# (filter template, approximate selectivity, projected key column, key cardinality)
_GEO = [("nation", 2, 25), ("city", 1, 250), ("region", 3, 5)]


def batch_queries(wl: Workload, n=64, seed=1, window=60_000_000, min_rows=1_000_000, prefix="bq"):
    """Seeded template generator of batched star-join queries over the four-way
    tables: 1-4 joins, selectivities about 1e-4..0.5, group cardinality 1..100K,
    fact slices of U[min_rows, window] rows inside the first `window` rows."""
    import random
    rng = random.Random(_mix(seed, 77))
    ids = {t.kind: t.id for t in wl.tables}
    fact_id = ids[GEN_FACT3]
    fact_rows = wl.fact_rows
    window = min(window, fact_rows)
    min_rows = min(min_rows, window)
    out = []
    for qn in range(n):
        qid = f"{prefix}{qn}"
        rows = rng.randint(min_rows, window)
        row0 = rng.randint(0, window - rows)
        view = (f"{qid}_slice", fact_id, row0, rows)
        nj = rng.randint(1, 4)
        dims = rng.sample(["date", "customer", "supplier", "part"], nj)
        order = ["date", "customer", "supplier", "part"]
        dims.sort(key=order.index)
        max_groups = rng.choice([1, 10, 100, 1000, 10_000, 100_000])
        groups = 1
        dim_specs, keys = [], []
        for d in dims:
            if d == "date":
                span = rng.choice([0, 1, 2, 6])
                y = rng.randint(1992, 1998 - span)
                filt = f"$(0.1)>={y} && $(0.1)<={y + span}" if span else f"$(0.1)=={y}"
                card, keycol = span + 1, 1
                dim_specs.append((d, ids[GEN_DATE], [I, I, I], [4, 4, 4], filt, "$(0.0), $(0.1)",
                                  [I, I, S], [4, 4, 40]))
            elif d in ("customer", "supplier"):
                attr, col, card = rng.choice(_GEO[:2])
                pick = rng.randint(0, card - 1)
                lit = f"NATION_{pick:02d}" if attr == "nation" else f"CITY_{(pick // 50) * 50 + pick % 50:03d}"
                filt = f'StrCmp($(0.{col}),"{lit}",10)==0'
                key_attr = rng.choice([1, 2, 3])  # city / nation / region of the passing rows
                card = {1: 250, 2: 25, 3: 5}[key_attr]
                keycol = 1
                src_id = ids[GEN_CUST3] if d == "customer" else ids[GEN_SUPP3]
                dim_specs.append((d, src_id, [I, S, S, S], [4, 10, 10, 10], filt, f"$(0.0), $(0.{key_attr})",
                                  [I, S, S], [4, 10, 40]))
            else:
                cat = rng.randint(1, 25)
                filt = f"$(0.1)=={cat}" if rng.random() < 0.7 else f"$(0.1)<={cat}"
                card, keycol = 1000, 1
                dim_specs.append((d, ids[GEN_PART3], [I, I, I], [4, 4, 4], filt, "$(0.0), $(0.2)",
                                  [I, I, S], [4, 4, 40]))
            if groups * card <= max_groups and len(keys) < 3:
                keys.append((len(dim_specs), keycol))
                groups *= card
        # optional fact predicate on revenue / supplycost to spread selectivity
        fact_filters = []
        r = rng.random()
        if r < 0.3:
            hi = rng.randint(10_000, 10_000_000)
            fact_filters.append(f"$(0.4)<={hi}")
        elif r < 0.45:
            lo = rng.randint(1_000, 119_000)
            fact_filters.append(f"$(0.5)>={lo} && $(0.5)<{lo + rng.randint(100, 60_000)}")
        metric = rng.choice(["$(0.4)", "$(0.4)-$(0.5)", "$(0.5)"])
        dpts, refs = [], []
        for k, (d, src_id, st, ss, filt, sel, ot, os_) in enumerate(dim_specs):
            pt = P.PipelineTest(f"{qid}_{d}")
            osch = P.create_schema(ot, os_)
            pt.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema(st, ss))
            pt.AddOperator(P.FILTER, {"predication": filt}, osch)
            pt.AddOperator(P.MAP, {"select": sel}, osch)
            pt.AddOperator(P.SINK, {}, osch)
            dpts.append((pt, src_id, []))
            refs.append(pt.pipe_info["id"])
        lk = {"date": 0, "customer": 1, "supplier": 2, "part": 3}
        res = P.create_schema([I] * 12, [4] * 12)
        fq = P.PipelineTest(qid)
        fq.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema([I] * 6, [4] * 6))
        for f in fact_filters:
            fq.AddOperator(P.FILTER, {"predication": f}, res)
        for k, d in enumerate(dims):
            fq.AddOperator(P.JOINWITH, {"lkey": f"$(0.{lk[d]})", "rkey": f"$({k + 1}.0)"}, res)
        dim_s = ", ".join(f"$({o}.{c})" for o, c in keys)
        fq.AddOperator(P.REDUCEBY, {"dim": dim_s, "metric": metric} if keys else {"dim": "", "metric": metric}, res)
        fq.AddOperator(P.SINK, {}, res)
        out.append(Query(qid, view, dpts, (fq, view[0], refs),
                         {"joins": nj, "keys": len(keys), "groups_max": groups, "rows": rows}))
    return out


def serve_queries(wl: Workload, n, seed=1, light_frac=0.9, light_rows=60_000_000, prefix="sq", id_digits=0):
    """Serving mix over the four-way tables: light queries (one join, <= 0.1% of
    rows pass, <= 100 groups, a slice of at most `light_rows` rows anywhere in the
    fact table) and heavy queries (the full four-way star join over the whole
    fact table, thousands of groups)."""
    import random
    rng = random.Random(_mix(seed, 99))
    ids = {t.kind: t.id for t in wl.tables}
    fact_id, fact_rows = ids[GEN_FACT3], wl.fact_rows
    out = []
    for qn in range(n):
        qid = f"{prefix}{qn:0{id_digits}d}" if id_digits else f"{prefix}{qn}"
        if rng.random() >= light_frac:
            # heavy: the four-way star join with seeded constants, whole table
            h = four_way(seed=rng.randint(1, 1 << 30), fact_rows=fact_rows, prefix=qid)
            year, cn, sn = h.constants["year"], h.constants["c_nation"], h.constants["s_nation"]
            fq, _, refs = h.fact
            dims = []
            for (pt, _, _), src in zip(h.dims, [ids[GEN_DATE], ids[GEN_CUST3], ids[GEN_SUPP3], ids[GEN_PART3]]):
                dims.append((pt, src, []))
            view = (f"{qid}_slice", fact_id, 0, fact_rows)
            out.append(Query(qid, view, dims, (fq, view[0], refs), {"kind": "heavy", "rows": fact_rows}))
            continue
        rows = rng.randint(min(1_000_000, light_rows), min(light_rows, fact_rows))
        row0 = rng.randint(0, fact_rows - rows)
        view = (f"{qid}_slice", fact_id, row0, rows)
        which = rng.choice(["date", "customer", "supplier"])
        res = P.create_schema([I] * 12, [4] * 12)
        pt = P.PipelineTest(f"{qid}_{which}")
        if which == "date":
            y = rng.randint(1992, 1998)
            dsch = P.create_schema([I, I, S], [4, 4, 40])
            pt.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema([I] * 3, [4] * 3))
            pt.AddOperator(P.FILTER, {"predication": f"$(0.1)=={y}"}, dsch)
            pt.AddOperator(P.MAP, {"select": "$(0.0), $(0.2)"}, dsch)  # key + month (12 groups)
            lkey, src = 0, ids[GEN_DATE]
            fact_filter = f"$(0.4)<={rng.randint(1_000, 60_000)}"   # ~0.1..0.6% of rows, times 1/7
        else:
            c = rng.randint(0, 249)
            dsch = P.create_schema([I, S, S], [4, 10, 40])
            pt.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema([I, S, S, S], [4, 10, 10, 10]))
            pt.AddOperator(P.FILTER, {"predication": f'StrCmp($(0.1),"CITY_{c:03d}",10)==0'}, dsch)
            pt.AddOperator(P.MAP, {"select": "$(0.0), $(0.2)"}, dsch)  # key + nation
            lkey = 1 if which == "customer" else 2
            src = ids[GEN_CUST3] if which == "customer" else ids[GEN_SUPP3]
            fact_filter = f"$(0.4)<={rng.randint(100_000, 2_000_000)}"  # 1..20% of rows, times ~1/250
        pt.AddOperator(P.SINK, {}, dsch)
        fq = P.PipelineTest(qid)
        fq.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema([I] * 6, [4] * 6))
        fq.AddOperator(P.FILTER, {"predication": fact_filter}, res)
        fq.AddOperator(P.JOINWITH, {"lkey": f"$(0.{lkey})", "rkey": "$(1.0)"}, res)
        fq.AddOperator(P.REDUCEBY, {"dim": "$(1.1)", "metric": rng.choice(["$(0.4)", "$(0.4)-$(0.5)"])}, res)
        fq.AddOperator(P.SINK, {}, res)
        out.append(Query(qid, view, [(pt, src, [])], (fq, view[0], [pt.pipe_info["id"]]),
                         {"kind": "light", "rows": rows}))
    return out


def make_views(queries):
    for q in queries:
        vid, base, row0, rows = q.view
        if P.table_exists(vid):
            P.table_drop(vid)
        P.table_view(vid, base, row0, rows)


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
