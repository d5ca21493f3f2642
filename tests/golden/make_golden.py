"""Writes tests/golden/mt19937_ref.json from the reference's own MT19937
(sqaodc/common/Random.cpp compiled from the reference tree into oracle/_ref
by oracle/Makefile).  Run in the build container, where the reference tree
exists:  python tests/golden/make_golden.py
"""
import ctypes
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

SEEDS = [0, 1, 5489, 12345, 1133557, (1 << 32) + 7, (1 << 64) - 1]


def main():
    O.build()
    R = O.ref_random()
    if R is None:
        raise SystemExit("oracle/_ref not built (reference tree missing)")
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    out = {"source": "sqaodc/common/Random.cpp (reference, compiled from source)", "cases": []}
    for s in SEEDS:
        w = np.empty(24, np.uint32)
        R.sqref_mt_stream(s, 0, p(w), 24)
        wskip = np.empty(8, np.uint32)
        R.sqref_mt_stream(s, 1000, p(wskip), 8)
        f32 = np.empty(8, np.float32)
        R.sqref_mt_f32(s, p(f32), 8)
        f64 = np.empty(8, np.float64)
        R.sqref_mt_f64(s, p(f64), 8)
        ri = {}
        for n in (2, 7, 128, 4096):
            r = np.empty(12, np.uint32)
            R.sqref_mt_randint(s, n, p(r), 12)
            ri[str(n)] = r.tolist()
        out["cases"].append({
            "seed": s,
            "u32_first24": w.tolist(),
            "u32_skip1000_8": wskip.tolist(),
            "f32_first8_bits": f32.view(np.uint32).tolist(),
            "f64_first8_bits": [int(v) for v in f64.view(np.uint64)],
            "randint_first12": ri,
        })
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mt19937_ref.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
