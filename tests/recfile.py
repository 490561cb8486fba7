"""Record files exchanged with the plain C++ host harness (tests/cpp/abi_harness.cpp): a
sequence of (u32 name length, name, u8 type 'd' f64 / 'i' i32 / 'l' i64, i64 count, data)."""
import struct

import numpy as np

_TYPES = {"d": np.float64, "i": np.int32, "l": np.int64}


def write(path, recs):
    with open(path, "wb") as f:
        for name, a in recs.items():
            if a is None:
                continue
            a = np.ascontiguousarray(np.atleast_1d(a))
            t = {np.dtype(np.float64): "d", np.dtype(np.int32): "i", np.dtype(np.int64): "l"}[a.dtype]
            nb = name.encode()
            f.write(struct.pack("<I", len(nb)) + nb + t.encode() + struct.pack("<q", a.size))
            f.write(a.tobytes())


def read(path):
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    i = 0
    while i < len(data):
        (ln,) = struct.unpack_from("<I", data, i); i += 4
        name = data[i:i + ln].decode(); i += ln
        t = chr(data[i]); i += 1
        (n,) = struct.unpack_from("<q", data, i); i += 8
        dt = np.dtype(_TYPES[t])
        out[name] = np.frombuffer(data, dt, n, i).copy(); i += n * dt.itemsize
    return out
