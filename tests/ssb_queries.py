"""The thirteen Star Schema Benchmark queries as the reference's tests write them
(test/test-SSB-Q11.cu ... test-SSB-Q43.cu), on seeded SSB-shaped tables.

Tables follow the reference's load schemas (test/test.hpp:29-46 for lineorder,
test/test-load.cu:79-122 for date, customer, part and supplier): Int32 columns and
fixed-width NUL-padded strings, with SSB dbgen's value domains (25 nations in 5
regions, cities as the nation's first nine characters plus a digit, MFGR#m / #mc /
#mcbb part hierarchies, date attributes of the days from 1992-01-01).

Each query is (dimension pipelines, fact pipeline), transcribed operator by operator
from the reference test, quirks included: the constant-true `(0.8)<35` of Q1.2/Q1.3,
the trailing `)` of some reduce-key lists, projections wider than their source column
(Q2.x's 8-byte p_mfgr into 10 bytes, Q3.1's 12-byte s_region into 15, Q4.3's 10-byte
s_city into 25), Q2.2's prefix-equal StrCmp range, Q3.x's joins listed out of ref order,
and Q4.x's joins, which pair lineorder keys with the ref objects in the order the
reference test passes them (lo_suppkey with part, lo_custkey with supplier, lo_partkey
with customer).
"""
from __future__ import annotations

import datetime

import numpy as np

from shadowvm_amd import pipeline as P

I, S = P.INT32, P.STRING

LINEORDER = ([I] * 6 + [S] + [I] * 9 + [S], [4] * 6 + [15] + [4] * 9 + [10])
DATE = ([I, S, S, S, I, I, S, I, I, I, I, I, S, I, I, I, I],
        [4, 18, 9, 9, 4, 4, 8, 4, 4, 4, 4, 4, 12, 4, 4, 4, 4])
CUSTOMER = ([I, S, S, S, S, S, S, S], [4, 25, 25, 10, 15, 15, 15, 15])
PART = ([I, S, S, S, S, S, S, I, S], [4, 22, 8, 8, 10, 11, 25, 4, 10])
SUPPLIER = ([I, S, S, S, S, S, S], [4, 25, 25, 10, 15, 12, 15])

NATIONS = [("ALGERIA", 0), ("ARGENTINA", 1), ("BRAZIL", 1), ("CANADA", 1), ("EGYPT", 4), ("ETHIOPIA", 0),
           ("FRANCE", 3), ("GERMANY", 3), ("INDIA", 2), ("INDONESIA", 2), ("IRAN", 4), ("IRAQ", 4),
           ("JAPAN", 2), ("JORDAN", 4), ("KENYA", 0), ("MOROCCO", 0), ("MOZAMBIQUE", 0), ("PERU", 1),
           ("CHINA", 2), ("ROMANIA", 3), ("SAUDI ARABIA", 4), ("VIETNAM", 2), ("RUSSIA", 3),
           ("UNITED KINGDOM", 3), ("UNITED STATES", 1)]
REGIONS = ["AFRICA", "AMERICA", "ASIA", "EUROPE", "MIDDLE EAST"]
MONTHS = ["January", "February", "March", "April", "May", "June", "July", "August", "September",
          "October", "November", "December"]
DAYS = ["Monday", "Tuesday", "Wednesday", "Thursday", "Friday", "Saturday", "Sunday"]


def _fixed(strings, width):
    """NUL-padded fixed-width byte rows (truncated to the width)."""
    out = np.zeros((len(strings), width), np.uint8)
    for r, s in enumerate(strings):
        b = s.encode()[:width]
        out[r, :len(b)] = np.frombuffer(b, np.uint8)
    return out


def _city(nation, digit):
    return nation[:9].ljust(9) + str(digit)


def gen_tables(seed=1, lineorder_rows=1_000_000, n_cust=30_000, n_supp=2_000, n_part=200_000):
    """{name: (types, sizes, columns)} of the five SSB tables."""
    rng = np.random.default_rng(seed)
    tabs = {}
    # date: 2,556 consecutive days from 1992-01-01
    d0 = datetime.date(1992, 1, 1)
    days = [d0 + datetime.timedelta(days=k) for k in range(2556)]
    dk = np.array([d.year * 10000 + d.month * 100 + d.day for d in days], np.int32)
    tabs["date"] = (DATE[0], DATE[1], [
        dk,
        _fixed([f"{MONTHS[d.month - 1]} {d.day}, {d.year}" for d in days], 18),
        _fixed([DAYS[d.weekday()] for d in days], 9),
        _fixed([MONTHS[d.month - 1] for d in days], 9),
        np.array([d.year for d in days], np.int32),
        np.array([d.year * 100 + d.month for d in days], np.int32),
        _fixed([MONTHS[d.month - 1][:3] + str(d.year) for d in days], 8),
        np.array([d.weekday() + 1 for d in days], np.int32),
        np.array([d.day for d in days], np.int32),
        np.array([d.timetuple().tm_yday for d in days], np.int32),
        np.array([d.month for d in days], np.int32),
        np.array([d.isocalendar()[1] for d in days], np.int32),
        _fixed([["Winter", "Spring", "Summer", "Fall"][(d.month % 12) // 3] for d in days], 12),
        np.array([int(d.weekday() == 6) for d in days], np.int32),
        np.array([int((d + datetime.timedelta(days=1)).month != d.month) for d in days], np.int32),
        (rng.random(len(days)) < 0.05).astype(np.int32),
        np.array([int(d.weekday() < 5) for d in days], np.int32),
    ])

    def geo(n):
        nat = rng.integers(0, 25, n)
        dig = rng.integers(0, 10, n)
        names = [NATIONS[k][0] for k in nat]
        return (_fixed([_city(NATIONS[k][0], g) for k, g in zip(nat, dig)], 10), names,
                [REGIONS[NATIONS[k][1]] for k in nat])

    city, nat, reg = geo(n_cust)
    ck = np.arange(1, n_cust + 1, dtype=np.int32)
    tabs["customer"] = (CUSTOMER[0], CUSTOMER[1], [
        ck, _fixed([f"Customer#{k:09d}" for k in ck], 25), _fixed([f"addr{k * 7919 % 100000}" for k in ck], 25),
        city, _fixed(nat, 15), _fixed(reg, 15), _fixed([f"{10 + k % 90}-{k % 1000:03d}" for k in ck], 15),
        _fixed([["AUTOMOBILE", "BUILDING", "FURNITURE", "HOUSEHOLD", "MACHINERY"][k % 5] for k in ck], 15)])
    city, nat, reg = geo(n_supp)
    sk = np.arange(1, n_supp + 1, dtype=np.int32)
    tabs["supplier"] = (SUPPLIER[0], SUPPLIER[1], [
        sk, _fixed([f"Supplier#{k:09d}" for k in sk], 25), _fixed([f"saddr{k * 104729 % 100000}" for k in sk], 25),
        city, _fixed(nat, 15), _fixed(reg, 12), _fixed([f"{10 + k % 90}-{k % 997:03d}" for k in sk], 15)])
    pk = np.arange(1, n_part + 1, dtype=np.int32)
    m = rng.integers(1, 6, n_part)
    c = rng.integers(1, 6, n_part)
    b = rng.integers(1, 41, n_part)
    tabs["part"] = (PART[0], PART[1], [
        pk, _fixed([f"part {k}" for k in pk], 22), _fixed([f"MFGR#{x}" for x in m], 8),
        _fixed([f"MFGR#{x}{y}" for x, y in zip(m, c)], 8),
        _fixed([f"MFGR#{x}{y}{z:02d}" for x, y, z in zip(m, c, b)], 10),
        _fixed([["red", "green", "blue", "ivory", "khaki"][k % 5] for k in pk], 11),
        _fixed([f"STANDARD BRUSHED {k % 7}" for k in pk], 25),
        rng.integers(1, 51, n_part).astype(np.int32),
        _fixed([["SM BOX", "LG CASE", "MED BAG", "JUMBO PKG"][k % 4] for k in pk], 10)])
    n = lineorder_rows
    qty = rng.integers(1, 51, n).astype(np.int32)
    price = rng.integers(90_000, 10_500_000, n).astype(np.int32)
    disc = rng.integers(0, 11, n).astype(np.int32)
    tabs["lineorder"] = (LINEORDER[0], LINEORDER[1], [
        (np.arange(n) // 4 + 1).astype(np.int32), (np.arange(n) % 4 + 1).astype(np.int32),
        rng.integers(1, n_cust + 1, n).astype(np.int32), rng.integers(1, n_part + 1, n).astype(np.int32),
        rng.integers(1, n_supp + 1, n).astype(np.int32), dk[rng.integers(0, len(dk), n)],
        _fixed([["1-URGENT", "2-HIGH", "3-MEDIUM", "4-NOT SPEC", "5-LOW"][k] for k in rng.integers(0, 5, n)], 15),
        np.zeros(n, np.int32), qty, price, rng.integers(1_000_000, 50_000_000, n).astype(np.int32), disc,
        (price.astype(np.int64) * (100 - disc) // 100).astype(np.int32),
        rng.integers(1_000, 120_000, n).astype(np.int32), rng.integers(0, 9, n).astype(np.int32),
        dk[rng.integers(0, len(dk), n)],
        _fixed([["REG AIR", "AIR", "RAIL", "SHIP", "TRUCK", "MAIL", "FOB"][k] for k in rng.integers(0, 7, n)], 10)])
    return tabs


def _dim(pid, source, filt, select, types, sizes):
    pt = P.PipelineTest(pid)
    osch = P.create_schema(types, sizes)
    pt.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema(*source))
    if filt is not None:
        pt.AddOperator(P.FILTER, {"predication": filt}, osch)
    pt.AddOperator(P.MAP, {"select": select}, osch)
    pt.AddOperator(P.SINK, {}, osch)
    return pt


def _fact(pid, filt, joins, dim, metric, ncols):
    pt = P.PipelineTest(pid)
    res = P.create_schema([I] * ncols, [4] * ncols)
    pt.AddOperator(P.SOURCE, {"type": "object"}, P.create_schema(*LINEORDER))
    if filt is not None:
        pt.AddOperator(P.FILTER, {"predication": filt}, res)
    for lk, rk in joins:
        pt.AddOperator(P.JOINWITH, {"lkey": lk, "rkey": rk}, res)
    pt.AddOperator(P.REDUCEBY, {"dim": dim, "metric": metric}, res)
    pt.AddOperator(P.SINK, {}, res)
    return pt


HR36 = P.HashRate * (4 + 32)


def queries(prefix, t):
    """{name: ([(dimension PipelineTest, source table id)], fact PipelineTest, ref ids)};
    t maps "lineorder"/"date"/... to the resident table ids."""
    q = {}
    # Q1.x: one date join, scalar revenue (test-SSB-Q11.cu:15-42, Q12, Q13)
    for name, dfilt, ffilt in [
            ("Q11", "$(0.4)==1993", " $(0.8)<25 && $(0.11)>=1 && $(0.11)<=3 "),
            ("Q12", "$(0.5)==199401", " $(0.8)>26 && (0.8)<35 && $(0.11)>=4 && $(0.11)<=6 "),
            ("Q13", "$(0.4)==1994 && $(0.11)==6", " $(0.8)>26 && (0.8)<35 && $(0.11)>=5 && $(0.11)<=7 ")]:
        d = _dim(f"{prefix}{name}_date", DATE, dfilt, "$(0.0), $(0.4)", [I, I, S], [4, 4, 40])
        f = _fact(f"{prefix}{name}", ffilt, [("$(0.5)", "$(1.0)")], "", "$(0.9)*$(0.11)/1000", 9)
        q[name] = ([(d, t["date"])], f)
    # Q2.x: date, part, supplier; group (d_year, p_mfgr-as-projected) (test-SSB-Q21.cu:17-70, Q22, Q23)
    for name, pfilt, sfilt, pw, dw in [
            ("Q21", 'StrCmp($(0.3),"MFGR#12",8)==0 ', 'StrCmp($(0.5),"AMERICA",15)==0', HR36, HR36),
            ("Q22", 'StrCmp($(0.3),"MFGR#2221",8)>=0 && StrCmp($(0.3),"MFGR#2228",8)<=0',
             'StrCmp($(0.5),"ASIA",15)==0', 40, 40),
            ("Q23", 'StrCmp($(0.3),"MFGR#2221",8)==0', 'StrCmp($(0.5),"EUROPE",15)==0', 40, 40)]:
        d = _dim(f"{prefix}{name}_date", DATE, None, "$(0.0), $(0.4)", [I, I, S], [4, 4, dw])
        p = _dim(f"{prefix}{name}_part", PART, pfilt, "$(0.0), $(0.2)", [I, S, S], [4, 10, pw])
        s = _dim(f"{prefix}{name}_supplier", SUPPLIER, sfilt, "$(0.0) ", [I, S], [4, pw])
        f = _fact(f"{prefix}{name}", None, [("$(0.5)", "$(1.0)"), ("$(0.3)", "$(2.0)"), ("$(0.4)", "$(3.0)")],
                  "$(1.1), $(2.1)", "$(0.12)/100", 9)
        q[name] = ([(d, t["date"]), (p, t["part"]), (s, t["supplier"])], f)
    # Q3.x: date, supplier, customer; group (d_year, s_geo) (test-SSB-Q31.cu:15-85, Q32, Q33, Q34)
    ki = 'StrCmp($(0.3),"UNITED KI1",10)==0 || StrCmp($(0.3),"UNITED KI5",10)==0'
    for name, dfilt, sfilt, ssel, ssz, cfilt, csel, csz, dim in [
            ("Q31", "$(0.4)>1995 && $(0.4)<=1997", 'StrCmp($(0.5),"ASIA",15)==0', "$(0.0), $(0.5)", 15,
             'StrCmp($(0.5),"ASIA",15)==0', "$(0.0),$(0.4)", 15, "$(1.1), $(2.1)"),
            ("Q32", "$(0.4)>=1992 && $(0.4)<=1997", 'StrCmp($(0.4),"UNITED STATES",15)==0', "$(0.0), $(0.3)",
             10, 'StrCmp($(0.4),"UNITED STATES",15)==0', "$(0.0),$(0.3)", 10, "$(1.1), $(2.1))"),
            ("Q33", "$(0.4)>=1992 && $(0.4)<=1997", ki, "$(0.0), $(0.3)", 10, ki, "$(0.0),$(0.3)", 10,
             "$(1.1), $(2.1))"),
            ("Q34", 'StrCmp($(0.6),"Dec1997",8)==0', ki, "$(0.0), $(0.3)", 10, ki, "$(0.0),$(0.3)", 10,
             "$(1.1), $(2.1))")]:
        d = _dim(f"{prefix}{name}_date", DATE, dfilt, "$(0.0), $(0.4)", [I, I, S], [4, 4, 40])
        s = _dim(f"{prefix}{name}_supplier", SUPPLIER, sfilt, ssel, [I, S, S], [4, ssz, 40])
        c = _dim(f"{prefix}{name}_customer", CUSTOMER, cfilt, csel, [I, S, S], [4, csz, 40])
        f = _fact(f"{prefix}{name}", None, [("$(0.4)", "$(2.0)"), ("$(0.5)", "$(1.0)"), ("$(0.2)", "$(3.0)")],
                  dim, "$(0.12)/100", 12)
        q[name] = ([(d, t["date"]), (s, t["supplier"]), (c, t["customer"])], f)
    # Q4.x: date, part, supplier, customer; group (d_year, c_nation), profit (test-SSB-Q41.cu..Q43)
    am = 'StrCmp($(0.5),"AMERICA",15)==0 '
    for name, dfilt, pfilt, psel, ptypes, psizes, sfilt, ssel, stypes, ssizes in [
            ("Q41", None, 'StrCmp($(0.2),"MFGR#1",8)==0 || StrCmp($(0.2),"MFGR#2",8)==0', "$(0.0)", [I, S],
             [4, 40], am, "$(0.0)", [I, S], [4, 40]),
            ("Q42", "$(0.4)==1997 || $(0.4)==1998", 'StrCmp($(0.2),"MFGR#1",8)==0 || StrCmp($(0.2),"MFGR#2",8)==0',
             "$(0.0)", [I, S], [4, 40], am, "$(0.0)", [I, S], [4, 40]),
            ("Q43", "$(0.4)==1997 || $(0.4)==1998", 'StrCmp($(0.3),"MFGR#14",8)==0', "$(0.0),$(0.4)",
             [I, S, S], [4, 10, 40], 'StrCmp($(0.4),"UNITED STATES",15)==0 ', "$(0.0), $(0.3)", [I, S, S],
             [4, 25, 40])]:
        d = _dim(f"{prefix}{name}_date", DATE, dfilt, "$(0.0), $(0.4)", [I, I, S], [4, 4, 40])
        p = _dim(f"{prefix}{name}_part", PART, pfilt, psel, ptypes, psizes)
        s = _dim(f"{prefix}{name}_supplier", SUPPLIER, sfilt, ssel, stypes, ssizes)
        c = _dim(f"{prefix}{name}_customer", CUSTOMER, am, "$(0.0),$(0.4)", [I, S, S], [4, 15, 40])
        f = _fact(f"{prefix}{name}", None, [("$(0.4)", "$(2.0)"), ("$(0.5)", "$(1.0)"), ("$(0.2)", "$(3.0)"),
                                            ("$(0.3)", "$(4.0)")], "$(1.1), $(4.1))", "($(0.12)-$(0.13))/100", 12)
        q[name] = ([(d, t["date"]), (p, t["part"]), (s, t["supplier"]), (c, t["customer"])], f)
    # test-query.cu:30-95 (test_query_SSB_Q2_1): a fact filter with a string predicate on
    # lo_shipmode; the test passes only the date object, so its second join names a ref
    # that does not exist -- here the part object of the same file is passed as that ref
    d = _dim(f"{prefix}QRY_date", DATE, "$(0.4)>=1992 && $(0.4)<=1997", "$(0.0), $(0.4)", [I, I, S], [4, 4, 40])
    p = _dim(f"{prefix}QRY_part", PART, 'StrCmp($(0.3),"MFGR#2221",8)==0 ', "$(0.0), $(0.2)", [I, S, S], [4, 10, 40])
    f = _fact(f"{prefix}QRY", "$(0.5)>=19930000 && $(0.5)<19940000 && $(0.8)<25 && $(0.11)>=1 && $(0.11)<=3 && "
              'StrCmp($(0.16),"AIR", 10)==0', [("$(0.5)", "$(1.0)"), ("$(0.3)", "$(2.0)")], "$(1.1), $(2.1)",
              "$(0.12)/100", 9)
    q["QRY"] = ([(d, t["date"]), (p, t["part"])], f)
    return q


NAMES = ["Q11", "Q12", "Q13", "Q21", "Q22", "Q23", "Q31", "Q32", "Q33", "Q34", "Q41", "Q42", "Q43", "QRY"]
