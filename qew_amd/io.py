"""Reader for the qew tools' output archives ("qew-kv 1").

The tools write the reference's HDF5 dataset keys (src/qew_io.h:24-63), one
dataset per line: `key<TAB>count<TAB>values` with values printed as %.17g.
"""
import numpy as np


def read_kv(path):
    """Returns {key: numpy array (or float for single values)}."""
    out = {}
    with open(path) as f:
        header = f.readline().strip()
        if header != "# qew-kv 1":
            raise ValueError(f"{path}: not a qew-kv archive (header {header!r})")
        for line in f:
            line = line.rstrip("\n")
            if not line:
                continue
            key, count, values = line.split("\t")
            arr = np.array([float(v) for v in values.split()]) if values else np.empty(0)
            if arr.size != int(count):
                raise ValueError(f"{path}: {key} has {arr.size} values, header says {count}")
            out[key] = float(arr[0]) if arr.size == 1 and not key.endswith("/h") else arr
    return out
