"""Reader for ZDF diagnostic files (the format the reference writes with
openmp/em2d/zdf.c and reads back with openmp/em2d/zdf.py).

    info = zdf.read("EMF/Ex-000010.zdf")
    info.type            # "grid" or "particles"
    info.data            # numpy array [count[1], count[0]] for grids,
                         # {quant: 1-D array} for particle files
    info.grid, info.iteration, info.particles   # metadata dicts

A file is the magic b"ZDF1" followed by records {u32 id, string name, u64
length, payload}; strings are u32 length + bytes padded to 4; everything is
little-endian.
"""
import struct

import numpy as np

MAGIC = b"ZDF1"

ID_INT32 = 0x00010000
ID_DOUBLE = 0x00020000
ID_STRING = 0x00030000
ID_DATASET = 0x00100002
ID_CDSET_START = 0x00110000
ID_CDSET_CHUNK = 0x00120000
ID_CDSET_END = 0x00130000
ID_ITERATION = 0x00200001
ID_GRID_INFO = 0x00210001
ID_PART_INFO = 0x00220002
ID_TRACK_INFO = 0x00230001

DTYPES = {1: np.int8, 2: np.uint8, 3: np.int16, 4: np.uint16, 5: np.int32, 6: np.uint32, 7: np.int64,
          8: np.uint64, 9: np.float32, 10: np.float64, 11: np.complex64, 12: np.complex128}


class ZDFError(ValueError):
    pass


class _Cursor:
    def __init__(self, buf):
        self.b = buf
        self.p = 0

    def take(self, n):
        if self.p + n > len(self.b):
            raise ZDFError("truncated ZDF file")
        v = self.b[self.p:self.p + n]
        self.p += n
        return v

    def u32(self):
        return struct.unpack("<I", self.take(4))[0]

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def u64(self):
        return struct.unpack("<Q", self.take(8))[0]

    def f64(self):
        return struct.unpack("<d", self.take(8))[0]

    def string(self):
        n = self.u32()
        if n == 0:
            return ""
        raw = self.take((n + 3) // 4 * 4)
        return bytes(raw[:n]).decode("latin-1")


class ZDFFile:
    """Contents of one ZDF file."""

    def __init__(self):
        self.records = []        # (id, name, payload) in file order
        self.type = None
        self.grid = None
        self.iteration = None
        self.particles = None
        self.datasets = {}       # name -> numpy array (C order, slowest axis first)
        self.data = None

    def __repr__(self):
        return f"ZDFFile(type={self.type!r}, datasets={list(self.datasets)})"


def _grid_info(c):
    nd = c.u32()
    count = [c.u64() for _ in range(nd)]
    g = {"ndims": nd, "count": count, "label": c.string(), "units": c.string()}
    has_axis = c.i32()
    g["axis"] = None
    if has_axis:
        g["axis"] = []
        for _ in range(nd):
            g["axis"].append({"name": c.string(), "type": c.i32(), "min": c.f64(), "max": c.f64(),
                              "label": c.string(), "units": c.string()})
    return g


def _part_info(c):
    p = {"label": c.string(), "np": c.u64()}
    nq = c.u32()
    p["nquants"] = nq
    p["quants"] = [c.string() for _ in range(nq)]
    p["qlabels"] = [c.string() for _ in range(nq)]
    p["qunits"] = [c.string() for _ in range(nq)]
    return p


def parse(buf):
    """Parses the bytes of a ZDF file."""
    if bytes(buf[:4]) != MAGIC:
        raise ZDFError("not a ZDF file (bad magic)")
    c = _Cursor(memoryview(buf))
    c.take(4)
    f = ZDFFile()
    while c.p < len(buf):
        rid = c.u32()
        name = c.string()
        length = c.u64()
        start = c.p
        if rid == ID_STRING:
            val = c.string()
            if name == "TYPE":
                f.type = val
            f.records.append((rid, name, val))
        elif rid == ID_INT32:
            f.records.append((rid, name, c.i32()))
        elif rid == ID_DOUBLE:
            f.records.append((rid, name, c.f64()))
        elif rid == ID_ITERATION:
            it = {"name": name, "n": c.i32(), "t": c.f64(), "time_units": c.string()}
            f.iteration = it
            f.records.append((rid, name, it))
        elif rid == ID_GRID_INFO:
            g = _grid_info(c)
            g["name"] = name
            f.grid = g
            f.records.append((rid, name, g))
        elif rid == ID_PART_INFO:
            p = _part_info(c)
            p["name"] = name
            f.particles = p
            f.records.append((rid, name, p))
        elif rid == ID_DATASET:
            ds_id = c.u32()
            dtype = DTYPES.get(c.i32())
            if dtype is None:
                raise ZDFError("unsupported dataset type")
            nd = c.u32()
            count = [c.u64() for _ in range(nd)]
            n = int(np.prod(count)) if nd else 0
            raw = c.take(n * np.dtype(dtype).itemsize)
            arr = np.frombuffer(bytes(raw), dtype=dtype).reshape(list(reversed(count)) if nd else [0])
            f.datasets[name] = arr
            f.records.append((rid, name, {"id": ds_id, "count": count}))
        else:
            c.take(length)
            f.records.append((rid, name, None))
        if c.p - start != length:
            raise ZDFError(f"record {name!r}: payload {c.p - start} bytes, header says {length}")
    if f.type == "grid" and f.datasets:
        f.data = next(iter(f.datasets.values()))
    elif f.type == "particles":
        f.data = dict(f.datasets)
    return f


def read(path):
    with open(path, "rb") as fh:
        return parse(fh.read())
