"""Reader for the tagged binary dumps written by oracle/ref_harness.cpp (test
infrastructure: used by tests/golden/make_golden.py and bench.py's CPU-baseline
leg)."""
import numpy as np


def read_dump(path):
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        nl = data.index(b"\n", pos)
        name, dtype, count = data[pos:nl].decode().split()
        count = int(count)
        pos = nl + 1
        dt = np.dtype("<f8") if dtype == "f8" else np.dtype("<i4")
        nbytes = count * dt.itemsize
        out[name] = np.frombuffer(data[pos:pos + nbytes], dtype=dt).copy()
        pos += nbytes
    return out
