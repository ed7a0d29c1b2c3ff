"""Pack the compiled reference's dumps into compressed golden fixtures.

Run by ``make -C oracle golden`` in the build container (the only place the
reference tree exists). Reads the flat record files written by
oracle/ref_dump.cpp and writes one ``<case>.npz`` per case here. The
fixtures hold numbers only: per-atom inputs the reference computed for a
built peptide (positions, charges, radii, GB terms, bond orders, neighbour
lists) and its outputs (Born radii, energies, forces).
"""
import os
import struct
import sys

import numpy as np

_DT = {b"d": np.float64, b"i": np.int32, b"u": np.uint32, b"b": np.int8}


def read_records(path):
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    off = 0
    while off < len(data):
        (nl,) = struct.unpack_from("<I", data, off)
        off += 4
        name = data[off:off + nl].decode()
        off += nl
        dt = data[off:off + 1]
        off += 1
        (cnt,) = struct.unpack_from("<Q", data, off)
        off += 8
        dtype = np.dtype(_DT[dt])
        arr = np.frombuffer(data, dtype=dtype, count=cnt, offset=off).copy()
        off += cnt * dtype.itemsize
        out[name] = arr
    return out


def main(src, dst):
    os.makedirs(dst, exist_ok=True)
    for fn in sorted(os.listdir(src)):
        if not fn.endswith(".bin"):
            continue
        rec = read_records(os.path.join(src, fn))
        np.savez_compressed(os.path.join(dst, fn[:-4] + ".npz"), **rec)
        print(fn, len(rec["charge"]), "atoms")


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
