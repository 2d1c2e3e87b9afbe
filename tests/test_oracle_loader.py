"""The Loader oracle (oracle/loader.py) against values derived by hand from
core/operator.cpp:24-48 and com/cpp/com.cu:33-44: field splitting, the drop of
bytes after a row's last column separator, the uncounted trailing partial row,
atoi/atol conversion (the C library's own functions), Int16 truncation, Byte,
and String truncation at the declared width."""
import numpy as np

from oracle import loader as L

TYPES = [L.INT32, L.INT16, L.INT32, L.INT64, L.STRING, L.BYTE]
SIZES = [4, 2, 4, 8, 10, 1]

TEXT = (b"1|  +42|-7x|99999999999|CITY_001|A|\n"
        b"|||||||\n"
        b"3|70000|abc|-9223372036854775809|0123456789ABC|xyz|trailing\n"
        b"5|1|2|3\n"
        b"\n"
        b"9|9|")  # partial row: no row separator


def test_hand_derived_rows(tmp_path):
    f = tmp_path / "t.tbl"
    f.write_bytes(TEXT)
    cols = L.load(str(f), TYPES, SIZES)
    assert [len(c) for c in cols] == [5] * 6
    assert cols[0].tolist() == [1, 0, 3, 5, 0]
    assert cols[1].tolist() == [42, 0, 70000 - 65536, 1, 0]
    assert cols[2].tolist() == [-7, 0, 0, 2, 0]
    assert cols[3].tolist() == [99999999999, 0, -(2 ** 63), 0, 0]
    assert bytes(cols[4][0]) == b"CITY_001\0\0"
    assert bytes(cols[4][2]) == b"012345678\0"  # width 10: 9 characters and the NUL
    assert bytes(cols[4][1]) == bytes(cols[4][3]) == b"\0" * 10
    assert cols[5].tolist() == [ord("A"), 0, ord("x"), 0, 0]


def test_atoi_is_the_c_library():
    # strtol saturation then int truncation, as atoi on this platform
    assert L.atoi(b"99999999999") == 99999999999 % 2 ** 32 - 2 ** 32 * (99999999999 % 2 ** 32 >= 2 ** 31)
    assert L.atoi(b"9223372036854775808") == -1
    assert L.atol(b"9223372036854775808") == 2 ** 63 - 1
    assert L.atoi(b"\t\r 12") == 12 and L.atoi(b"+-3") == 0 and L.atoi(b"0x10") == 0


def test_directory_is_read_in_name_order(tmp_path):
    d = tmp_path / "dir"
    d.mkdir()
    (d / "b.tbl").write_bytes(b"2|\n")
    (d / "a.tbl").write_bytes(b"1|\n")
    (d / ".hidden").write_bytes(b"9|\n")
    cols = L.load(str(d), [L.INT32], [4])
    assert cols[0].tolist() == [1, 2]
