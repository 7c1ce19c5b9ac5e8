"""IMXBIN container: the flat record format libimx writes meshes in.

Layout: the 8-byte magic ``IMXBIN01`` followed by records, each a 24-byte
space-padded ASCII name, a 4-byte dtype tag (``i4  `` or ``f8  ``), an int64
element count and the little-endian data.
"""
import numpy as np

_DT = {b"i4  ": np.int32, b"f8  ": np.float64}
_TAG = {np.dtype(np.int32): b"i4  ", np.dtype(np.float64): b"f8  "}


def read(path):
    """Return {name: ndarray} for every record of an IMXBIN file."""
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:8] != b"IMXBIN01":
        raise ValueError(f"{path}: not an IMXBIN file")
    out, pos = {}, 8
    while pos < len(buf):
        name = buf[pos:pos + 24].decode().strip()
        tag = buf[pos + 24:pos + 28]
        n = int(np.frombuffer(buf[pos + 28:pos + 36], dtype=np.int64)[0])
        dt = np.dtype(_DT[tag])
        pos += 36
        out[name] = np.frombuffer(buf[pos:pos + n * dt.itemsize], dtype=dt).copy()
        pos += n * dt.itemsize
    return out


def write(path, records):
    """Write {name: array} (int32 or float64) as an IMXBIN file."""
    with open(path, "wb") as f:
        f.write(b"IMXBIN01")
        for name, a in records.items():
            a = np.ascontiguousarray(a)
            if a.dtype not in _TAG:
                a = a.astype(np.float64 if a.dtype.kind == "f" else np.int32)
            f.write(name.encode().ljust(24)[:24])
            f.write(_TAG[a.dtype])
            f.write(np.int64(a.size).tobytes())
            f.write(a.tobytes())
