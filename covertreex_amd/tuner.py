"""Tuner for the mixed engine's screen tile shape (candidates per segment).

The screen runs one workgroup per (64-query tile, candidate segment); an XCD
streams a segment's BF16 hi/lo V rows for every query tile while they sit in
its L2, so the segment length trades L2 reuse against the number of
workgroups and the histogram/code traffic per workgroup.  For each shape
(N, D, m, kernel, k, predecessor) :func:`tune_screen_segment` builds a
synthetic dataset of that shape on the GPU (N(0,1) or U[-3,3] coordinates,
the device Nystrom factor), runs the all-points mixed kNN once per candidate
segment length, times the screen (pass 1, the threshold and the filter) with
the tree's own HIP events, checks that every segment length returns the same
neighbours and distances, and reports the fastest.

``python -m covertreex_amd.tuner`` tunes the workload shapes and prints one
JSON line per shape; the library's default segment rule follows its picks.
"""

from __future__ import annotations

import argparse
import json
import sys
from dataclasses import dataclass

import numpy as np

CANDIDATES = (256, 512, 1024, 2048, 4096, 8192)


@dataclass(frozen=True)
class Shape:
    name: str
    n: int
    d: int
    m: int
    k: int
    kernel_type: int = 1        # 0 RBF, 1 Matern-5/2
    predecessor: bool = True
    uniform: bool = False       # U[-3, 3] coordinates instead of N(0, 1)
    ard: bool = False           # per-dimension lengthscales log-uniform in [0.5, 2]


WORKLOAD_SHAPES = (
    Shape("c1", 4096, 3, 64, 16, 1, predecessor=False),
    Shape("c2", 32768, 3, 256, 50, 1),
    Shape("c4", 262144, 5, 512, 30, 0, uniform=True, ard=True),
    Shape("serve-m64", 65536, 3, 64, 30, 1),
    Shape("serve-m128", 65536, 2, 128, 30, 0),
    Shape("serve-m256", 131072, 5, 256, 50, 1),
)


def _tree(shape: Shape):
    from .native import HipTree
    from .residual import build_residual_backend

    rng = np.random.default_rng(0)
    if shape.uniform:
        x = rng.uniform(-3, 3, (shape.n, shape.d)).astype(np.float32)
    else:
        x = rng.standard_normal((shape.n, shape.d)).astype(np.float32)
    ls = (np.exp(np.random.default_rng(0).uniform(np.log(0.5), np.log(2.0), shape.d))
          if shape.ard else np.ones(1))
    b = build_residual_backend(x, seed=0, inducing_count=shape.m, variance=1.0,
                               lengthscale=ls, kernel_type=shape.kernel_type)
    return HipTree.build_residual(b.v_matrix, b.p_diag, x, 1.0, np.asarray(ls, np.float32),
                                  kernel_type=shape.kernel_type)


def tune_screen_segment(shape: Shape, candidates=CANDIDATES, reps: int = 3) -> dict:
    """Screen time (ms, median of ``reps`` calls) per segment length for one
    shape, and the pick."""
    import torch

    from .native import HipTree

    tree = _tree(shape)
    dev = torch.device("cuda:0")
    n, k = shape.n, shape.k
    q = torch.arange(n, dtype=torch.int64, device=dev)
    oi = torch.empty((n, k), dtype=torch.int64, device=dev)
    od = torch.empty((n, k), dtype=torch.float32, device=dev)
    oe = torch.empty((n, k), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    names = HipTree.PHASES["mixed"]
    screen = [names.index(p) for p in ("screen_pass1", "hist_tau", "screen_pass2")]
    times, ref = {}, None
    try:
        for seg in candidates:
            tree.set_screen_segment(seg)
            call = lambda: tree.knn_residual_device(  # noqa: E731
                q.data_ptr(), n, k, oi.data_ptr(), od.data_ptr(), oe.data_ptr(),
                predecessor_mode=shape.predecessor, mode="mixed")
            call()
            tree.set_profiling(True)
            ms = []
            for _ in range(reps):
                call()
                prof = tree.last_profile()
                ms.append(sum(prof[i] for i in screen))
            tree.set_profiling(False)
            torch.cuda.synchronize()
            out = (oi.cpu().numpy(), od.cpu().numpy().view(np.uint32))
            if ref is None:
                ref = out
            elif not (np.array_equal(ref[0], out[0]) and np.array_equal(ref[1], out[1])):
                raise RuntimeError(f"{shape.name}: segment {seg} changed the results")
            times[seg] = float(np.median(ms))
    finally:
        tree.close()
    pick = min(times, key=times.get)
    return {"shape": shape.name, "n": shape.n, "d": shape.d, "m": shape.m, "k": shape.k,
            "screen_ms": {str(s): round(t, 4) for s, t in times.items()}, "pick": pick}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--shapes", default=",".join(s.name for s in WORKLOAD_SHAPES))
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)
    import torch  # noqa: F401  (loads the HIP runtime before the library)

    want = set(args.shapes.split(","))
    for shape in WORKLOAD_SHAPES:
        if shape.name in want:
            print(json.dumps(tune_screen_segment(shape, reps=args.reps)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
