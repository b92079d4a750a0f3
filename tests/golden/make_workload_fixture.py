"""Writes workload_pte2.json and workload_uneven.json: SHA-256 of every array
of the 65,536-cell two-material workload (mt19937_64, seed 0x5EED0000) and of
the 2^20-cell uneven workload (seed 0x5EED0002), plus the values of their
first cells. Regenerate only when a generator's specification changes."""
import hashlib
import importlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
wl = importlib.import_module("singularity-eos_amd.workloads")

n = 65536
a = wl.pte2(n)
keys = ("rho", "vfrac", "sie", "temp", "press", "sie_tot")
out = {
    "ncell": n,
    "seed": "0x5EED0000",
    "sha256": {k: hashlib.sha256(np.ascontiguousarray(a[k]).tobytes()).hexdigest() for k in keys},
    "head": {k: [float(x) for x in a[k][: 32 if k != "sie_tot" else 16]] for k in keys},
}
json.dump(out, open(os.path.join(HERE, "workload_pte2.json"), "w"), indent=1)

ctx, _ = wl.pool_context(host_only=True)
n = 1 << 20
a = wl.uneven(ctx, n)
ctx.close()
keys = ("offsets", "mat_ids", "rho", "vfrac", "sie", "temp", "press", "sie_tot")
out = {
    "ncell": n,
    "seed": "0x5EED0002",
    "nnz": int(a["nnz"]),
    "sha256": {k: hashlib.sha256(np.ascontiguousarray(a[k]).tobytes()).hexdigest() for k in keys},
    "head": {k: [float(x) if a[k].dtype.kind == "f" else int(x) for x in a[k][:24]]
             for k in keys},
}
json.dump(out, open(os.path.join(HERE, "workload_uneven.json"), "w"), indent=1)
