"""CPU oracle of the Loader operator.  TEST INFRASTRUCTURE ONLY.

Restates ``Loader::Host`` (core/operator.cpp:24-48) with
``GetValueOfItemBySwitch`` (com/cpp/com.cu:33-44) over the files of a Source
path (object/object.cpp:122-169):

* bytes are scanned in order; a column separator ends a field, which is stored
  into column ``col`` of row ``row`` and ``col`` advances; a row separator
  starts the next row (``col = 0``) and drops the bytes after the row's last
  column separator; a trailing partial row is not counted (``tuple_count_ =
  row``);
* Byte stores the field's first character (NUL for an empty field); Int16 /
  Int32 store ``atoi(field)`` and Int64 ``atol(field)``, here computed by the C
  library's own ``atoi`` / ``atol`` through ctypes, the very functions the
  reference calls; String copies the field and NUL-terminates at the declared
  width.

Deviations (the product documents the same):
  L1 a String field shorter than its width is NUL-padded (the reference copies
     ``size`` bytes of a reused 4 KB item buffer, so the bytes after the NUL are
     left over from earlier, longer fields);
  L2 files of a directory are read in file-name order (the reference takes
     readdir order) and concatenated into one table;
  L3 fields past the schema's last column are ignored and a row without enough
     fields keeps 0 in the missing columns (the reference writes out of bounds /
     leaves the sink bulk's prior contents);
  L4 a file may hold more rows than one bulk (the reference writes past its
     ``tuple_per_bulk`` sink bulk).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

BYTE, INT16, INT32, INT64, STRING = 0, 1, 2, 3, 7

_libc = ctypes.CDLL("libc.so.6")
_libc.atoi.restype = ctypes.c_int
_libc.atoi.argtypes = [ctypes.c_char_p]
_libc.atol.restype = ctypes.c_long
_libc.atol.argtypes = [ctypes.c_char_p]


def atoi(field: bytes) -> int:
    return _libc.atoi(field)


def atol(field: bytes) -> int:
    return _libc.atol(field)


def source_files(path: str) -> list[str]:
    if os.path.isdir(path):
        return sorted(os.path.join(path, n) for n in os.listdir(path)
                      if not n.startswith(".") and os.path.isfile(os.path.join(path, n)))
    return [path]


def _iter_rows(data: bytes, row_sep: bytes):
    start = 0
    while True:
        k = data.find(row_sep, start)
        if k < 0:
            return
        yield data[start:k]
        start = k + 1


def load(path: str, types, sizes, col_sep: str = "|", row_sep: str = "\n"):
    """Columns of the loaded table: int columns as numpy arrays, String columns
    as (rows, size) uint8 arrays."""
    cs, rs = col_sep.encode()[:1], row_sep.encode()[:1]
    rows = []
    for f in source_files(path):
        with open(f, "rb") as fh:
            data = fh.read()
        rows.extend(_iter_rows(data, rs))
    n = len(rows)
    np_t = {BYTE: np.int8, INT16: np.int16, INT32: np.int32, INT64: np.int64}
    cols = [np.zeros((n, s), np.uint8) if t == STRING else np.zeros(n, np_t[t]) for t, s in zip(types, sizes)]
    for r, line in enumerate(rows):
        parts = line.split(cs)
        for c, field in enumerate(parts[:-1]):  # bytes after the last column separator are dropped
            if c >= len(types):
                break
            item = field.split(b"\0", 1)[0]  # the item is a C string
            t = types[c]
            if t == BYTE:
                cols[c][r] = np.frombuffer(field[:1] or b"\0", np.int8)[0]
            elif t == INT16:
                cols[c][r] = np.array(atoi(item), np.int32).astype(np.int16)
            elif t == INT32:
                cols[c][r] = atoi(item)
            elif t == INT64:
                cols[c][r] = atol(item)
            else:
                size = sizes[c]
                b = field[:size - 1]
                cols[c][r, :len(b)] = np.frombuffer(b, np.uint8)
    return cols
