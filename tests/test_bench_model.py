"""The benchmark's model of the bytes the staged fact scan has to read, on a table
small enough to count the 128-byte lines by hand."""
from types import SimpleNamespace

import numpy as np

import bench


def _case():
    rows = 1024
    cols = [np.zeros(rows, np.int32) for _ in range(6)]
    # join 1 (fact column 1) passes rows 0..31 (line 0) and 512..543 (line 16);
    # join 2 (fact column 2) passes rows 0..31 and 600..609
    cols[1][:] = 1
    cols[1][0:32] = 4
    cols[1][512:544] = 8
    cols[2][:] = 1
    cols[2][0:32] = 6
    cols[2][600:610] = 6
    dims = {"d0": SimpleNamespace(columns=[np.array([0], np.int32)]),
            "d1": SimpleNamespace(columns=[np.array([4, 8], np.int32)]),
            "d2": SimpleNamespace(columns=[np.array([6], np.int32)]),
            "d3": SimpleNamespace(columns=[np.array([0], np.int32)])}
    wl = SimpleNamespace(fact=(None, "f", ["d0", "d1", "d2", "d3"]))
    return cols, dims, wl


def test_staged_bytes_by_lines():
    cols, dims, wl = _case()
    # join 1 at tier 0, join 2 at tier 1, the rest (columns 0, 3, 4, 5) after both
    total, detail = bench.scan_bytes_model("flat compacted-tiers(-1,0,1,-1)", cols, dims, wl)
    # column 1 streamed (32 lines); column 2 in the 2 lines with a row passing join 1;
    # columns 0, 3, 4, 5 in the 1 line with a row passing joins 1 and 2
    assert total == 32 * 128 + 2 * 128 + 4 * 1 * 128
    assert [c["tier"] for c in detail["columns"]] == [2, 0, 1, 2, 2, 2]


def test_packed_tier0_column_costs_three_bytes_a_row():
    cols, dims, wl = _case()
    total, detail = bench.scan_bytes_model("flat compacted-tiers(-1,0,1,-1) packed24(1)", cols, dims, wl)
    # two blocks of 512 rows of 1,536 bytes replace the 32 streamed lines
    assert total == 2 * 1536 + 2 * 128 + 4 * 128
    assert detail["columns"][1] == {"column": 1, "tier": 0, "bytes": 3072, "copy": "24-bit"}


def test_narrower_copies_cost_their_bits():
    cols, dims, wl = _case()
    # 18 bits a row: 1,024 + 128 bytes per block of 512 rows; 16 bits: 1,024
    total, detail = bench.scan_bytes_model("flat compacted-tiers(-1,0,1,-1) packed18(1)", cols, dims, wl)
    assert total == 2 * 1152 + 2 * 128 + 4 * 128
    assert detail["columns"][1]["copy"] == "18-bit"
    total, _ = bench.scan_bytes_model("flat compacted-tiers(-1,0,1,-1) packed16(1)", cols, dims, wl)
    assert total == 2 * 1024 + 2 * 128 + 4 * 128


def test_unstaged_streams_every_column():
    cols, dims, wl = _case()
    total, _ = bench.scan_bytes_model("256x1 hash", cols, dims, wl)
    assert total == 6 * 32 * 128
