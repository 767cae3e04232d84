"""Reader/writer for the .xfb named-array files exchanged with oracle/_ref/ref_harness.

Test infrastructure only.  Record layout: u32 name_len, name, u8 dtype
('d' float64 | 'i' int32), u32 ndim, i64 dims[ndim], raw little-endian data.
"""
import struct

import numpy as np


def write_xfb(path, arrays):
    with open(path, "wb") as fp:
        for name, a in arrays.items():
            a = np.asarray(a)
            if a.dtype.kind == "f":
                a = np.ascontiguousarray(a, dtype="<f8")
                dt = b"d"
            else:
                a = np.ascontiguousarray(a, dtype="<i4")
                dt = b"i"
            nb = name.encode()
            fp.write(struct.pack("<I", len(nb)))
            fp.write(nb)
            fp.write(dt)
            fp.write(struct.pack("<I", a.ndim))
            fp.write(struct.pack("<%dq" % a.ndim, *a.shape))
            fp.write(a.tobytes())


def read_xfb(path):
    out = {}
    with open(path, "rb") as fp:
        data = fp.read()
    pos = 0
    while pos < len(data):
        (nl,) = struct.unpack_from("<I", data, pos)
        pos += 4
        name = data[pos:pos + nl].decode()
        pos += nl
        dt = data[pos:pos + 1]
        pos += 1
        (nd,) = struct.unpack_from("<I", data, pos)
        pos += 4
        dims = struct.unpack_from("<%dq" % nd, data, pos)
        pos += 8 * nd
        cnt = int(np.prod(dims)) if nd else 1
        if dt == b"d":
            arr = np.frombuffer(data, dtype="<f8", count=cnt, offset=pos).reshape(dims)
            pos += 8 * cnt
        else:
            arr = np.frombuffer(data, dtype="<i4", count=cnt, offset=pos).reshape(dims)
            pos += 4 * cnt
        out[name] = arr.copy()
    return out
