"""Write Hilbert-order golden vectors from the reference's own ordering code.

Run only where the reference's sources are present (the tests read the
committed vectors, never the reference):

    python tests/golden/make_hilbert_golden.py

It loads ``covertreex/algo/order/strategy.py`` from the reference as a
standalone module (that file imports numpy alone, so no other part of the
reference is imported) and records, for point sets in D in {2, 3, 5, 8, 21}
with duplicate rows and a constant column, the Skilling Hilbert codes
(``_axes_to_transpose`` + ``_transpose_to_hilbert``) and the permutation
``compute_batch_order(points, strategy="hilbert")`` returns.  The output,
``hilbert_golden.npz``, holds only those inputs and outputs.
"""

from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np

REFERENCE = "/root/reference/covertreex/algo/order/strategy.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hilbert_golden.npz")


def load_strategy():
    spec = importlib.util.spec_from_file_location("reference_order_strategy", REFERENCE)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod  # its dataclass looks itself up there
    spec.loader.exec_module(mod)
    return mod


def cases():
    rng = np.random.default_rng(20261204)
    for d, n in ((2, 300), (3, 300), (5, 240), (8, 200), (21, 120)):
        pts = rng.standard_normal((n, d))
        pts[rng.choice(n, n // 10, replace=False)] = pts[:n // 10]  # duplicate rows
        pts[:, d - 1] = np.round(pts[:, d - 1], 1)                 # coarse ties on one axis
        yield f"d{d}", pts
    const = rng.uniform(-2.0, 2.0, (64, 3))
    const[:, 1] = 0.75                                             # zero span on one axis
    yield "d3_const_axis", const


def main():
    mod = load_strategy()
    data = {}
    for name, pts in cases():
        n, d = pts.shape
        bits = max(1, min(16, 64 // max(1, d)))
        lo, hi = pts.min(axis=0), pts.max(axis=0)
        scaled = (pts - lo) / np.maximum(hi - lo, 1e-9)
        ints = np.clip(np.round(scaled * ((1 << bits) - 1)), 0, (1 << bits) - 1).astype(np.uint64)
        codes = np.array([mod._transpose_to_hilbert(mod._axes_to_transpose(r.tolist(), bits),
                                                    bits, d) for r in ints], dtype=np.uint64)
        perm = mod.compute_batch_order(pts, strategy="hilbert").permutation
        data[f"{name}_points"] = pts
        data[f"{name}_codes"] = codes
        data[f"{name}_perm"] = np.asarray(perm, np.int64)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT}: {sorted(k for k in data if k.endswith('_perm'))}")


if __name__ == "__main__":
    main()
