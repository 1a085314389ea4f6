"""Minimal numpy reader for PCD v0.7 files (ascii and binary), used by the
tests to load the reference's data files independently of the product's
C++ reader.  Returns a dict of field name -> numpy array."""
import numpy as np

_TYPES = {("F", 4): np.float32, ("F", 8): np.float64, ("U", 4): np.uint32, ("U", 1): np.uint8,
          ("U", 2): np.uint16, ("I", 4): np.int32, ("I", 1): np.int8, ("I", 2): np.int16}


def read_pcd(path):
    with open(path, "rb") as f:
        raw = f.read()
    hdr = {}
    pos = 0
    while True:
        end = raw.index(b"\n", pos)
        line = raw[pos:end].decode("ascii", "replace").strip()
        pos = end + 1
        if not line or line.startswith("#"):
            continue
        key, _, rest = line.partition(" ")
        hdr[key.upper()] = rest.split()
        if key.upper() == "DATA":
            break
    fields = hdr["FIELDS"]
    sizes = [int(s) for s in hdr["SIZE"]]
    types = hdr["TYPE"]
    counts = [int(c) for c in hdr.get("COUNT", ["1"] * len(fields))]
    n = int(hdr["POINTS"][0])
    mode = hdr["DATA"][0]
    dt = np.dtype([(f if f != "_" else f"_pad{i}", _TYPES[(t, s)], (c,) if c > 1 else ())
                   for i, (f, s, t, c) in enumerate(zip(fields, sizes, types, counts))])
    if mode == "binary":
        arr = np.frombuffer(raw[pos:pos + n * dt.itemsize], dtype=dt, count=n)
    elif mode == "ascii":
        rows = raw[pos:].decode("ascii").split("\n")
        rows = [r for r in rows if r.strip()][:n]
        vals = [r.split() for r in rows]
        arr = np.zeros(n, dtype=dt)
        col = 0
        for f, c, t in zip(dt.names, counts, types):
            colvals = [v[col:col + c] for v in vals]
            if t == "F":
                arr[f] = np.array(colvals, dtype=np.float64).reshape(arr[f].shape)
            else:
                arr[f] = np.array(colvals, dtype=np.float64).astype(arr[f].dtype).reshape(arr[f].shape)
            col += c
    else:
        raise ValueError(f"unsupported PCD DATA mode {mode}")
    return {"fields": fields, "points": n, "data": arr}


def read_xyz(path):
    d = read_pcd(path)["data"]
    return np.stack([d["x"], d["y"], d["z"]], axis=1).astype(np.float32)
