"""codeare's raw `.cod` matrix files (src/matrix/io/CODFile.hpp:45-190).

Layout, little-endian, as CODFile::Write emits it:
    int32   dtype    (IOFile.hpp:20: RLFL 0, RLDB 1, CXFL 2, CXDB 3, LONG 4, SHRT 5)
    uint64  ndim
    uint64  dims[ndim]
    float32 res[ndim]
    uint64  name length, then that many bytes of name
    data    prod(dims) elements, column-major (codeare Matrix storage)
    32 bytes delimiter "543f562189f1e82beb9c177f89f67822" (written, never read)
Several matrices may follow each other in one file only in the sense that
CODFile reads one matrix from the current position; this module reads and
writes one matrix per file, like codeare's IOContext does.
"""
from __future__ import annotations

import struct

import numpy as np

DELIM = b"543f562189f1e82beb9c177f89f67822"
DTYPES = {0: np.float32, 1: np.float64, 2: np.complex64, 3: np.complex128, 4: np.int64, 5: np.int16}
CODES = {np.dtype(v): k for k, v in DTYPES.items()}


def read_cod(path: str):
    """Returns (array in Fortran order with codeare's dims, name, res)."""
    with open(path, "rb") as f:
        buf = f.read()
    off = 0
    (dt,) = struct.unpack_from("<i", buf, off)
    off += 4
    if dt not in DTYPES:
        raise ValueError(f"{path}: unknown data type code {dt}")
    (nd,) = struct.unpack_from("<Q", buf, off)
    off += 8
    if nd < 1 or nd > 16:
        raise ValueError(f"{path}: implausible dimension count {nd}")
    dims = struct.unpack_from(f"<{nd}Q", buf, off)
    off += 8 * nd
    res = struct.unpack_from(f"<{nd}f", buf, off)
    off += 4 * nd
    (nl,) = struct.unpack_from("<Q", buf, off)
    off += 8
    name = buf[off:off + nl].decode("latin-1")
    off += nl
    dtype = np.dtype(DTYPES[dt])
    count = int(np.prod(dims))
    need = count * dtype.itemsize
    if off + need > len(buf):
        raise ValueError(f"{path}: truncated data ({len(buf) - off} of {need} bytes)")
    data = np.frombuffer(buf, dtype=dtype, count=count, offset=off).copy()
    return data.reshape(dims, order="F"), name, tuple(res)


def write_cod(path: str, a, name: str = "", res=None) -> None:
    a = np.asarray(a)
    code = CODES.get(a.dtype)
    if code is None:
        raise ValueError(f"dtype {a.dtype} has no .cod code")
    dims = a.shape if a.ndim else (1,)
    res = tuple(res) if res is not None else (1.0,) * len(dims)
    nb = name.encode("latin-1")
    with open(path, "wb") as f:
        f.write(struct.pack("<i", code))
        f.write(struct.pack("<Q", len(dims)))
        f.write(struct.pack(f"<{len(dims)}Q", *dims))
        f.write(struct.pack(f"<{len(dims)}f", *res))
        f.write(struct.pack("<Q", len(nb)))
        f.write(nb)
        f.write(np.asfortranarray(a).tobytes(order="F"))
        f.write(DELIM)
