"""FP64 NumPy oracle (TEST INFRASTRUCTURE, see ``oracle/__init__``).

- :func:`nystrom` restates ``build_residual_backend``
  (covertreex/metrics/residual/host_backend.py:165-246) in FP64.
- :func:`residual_dist_fp64` / :func:`brute_knn_fp64` evaluate the residual
  distance d = sqrt(1 - |rho|) in FP64 on the *stored FP32 inputs* (coords, V,
  p cast up exactly) with the consistent per-dimension lengthscale, and take the
  exact top-k with the Vecchia mask j < i.
"""

from __future__ import annotations

import numpy as np


def kernel_fp64(x: np.ndarray, y: np.ndarray, *, variance: float, lengthscale, kernel_type: int):
    """RBF / Matern-5/2 kernel in FP64 (host_backend.py:14-43; kernels.py:59-132).

    ``lengthscale`` may be a scalar or a per-dimension vector (ARD)."""
    ls = np.maximum(np.asarray(lengthscale, dtype=np.float64), 1e-12)
    xs = np.asarray(x, np.float64) / ls
    ys = np.asarray(y, np.float64) / ls
    diff = xs[:, None, :] - ys[None, :, :]
    sq = np.sum(diff * diff, axis=2)
    if kernel_type == 1:
        a = np.sqrt(5.0) * np.sqrt(sq)
        return float(variance) * (1.0 + a + a * a / 3.0) * np.exp(-a)
    return float(variance) * np.exp(-0.5 * sq)


def nystrom(points, *, inducing_count: int, seed: int, variance: float = 1.0,
            lengthscale=1.0, kernel_type: int = 0, jitter: float = 1e-6):
    """Nystrom factor V = (L^-1 K_xm^T)^T, p = max(var - |V|^2, 1e-9).

    Returns FP32 V, p, k_diag (as stored by the reference) plus the FP64 V and
    the sorted inducing indices."""
    pts = np.asarray(points, np.float64)
    n = pts.shape[0]
    m = min(int(inducing_count), n)
    if m <= 0:
        m = min(32, n)
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(n, size=m, replace=False)) if m < n else np.arange(n)
    u = pts[idx]
    kmm = kernel_fp64(u, u, variance=variance, lengthscale=lengthscale, kernel_type=kernel_type)
    kmm = kmm + np.eye(m) * (jitter * variance)
    lmm = np.linalg.cholesky(kmm)
    kxm = kernel_fp64(pts, u, variance=variance, lengthscale=lengthscale, kernel_type=kernel_type)
    v64 = np.linalg.solve(lmm, kxm.T).T
    kdiag = np.full(n, variance, np.float64)
    p64 = np.maximum(kdiag - np.sum(v64 * v64, axis=1), 1e-9)
    return (v64.astype(np.float32), p64.astype(np.float32), kdiag.astype(np.float32), v64, p64,
            idx)


def residual_dist_fp64(coords, v, p, qi, cj, *, variance=1.0, lengthscale=1.0, kernel_type=0):
    """FP64 residual distance matrix between rows ``qi`` and ``cj``."""
    coords = np.asarray(coords, np.float64)
    v = np.asarray(v, np.float64)
    p = np.asarray(p, np.float64)
    qi = np.asarray(qi, np.int64)
    cj = np.asarray(cj, np.int64)
    k = kernel_fp64(coords[qi], coords[cj], variance=variance, lengthscale=lengthscale,
                    kernel_type=kernel_type)
    dot = v[qi] @ v[cj].T
    denom = np.sqrt(np.maximum(p[qi][:, None] * p[cj][None, :], 1e-18))
    rho = np.clip((k - dot) / denom, -1.0, 1.0)
    d = np.sqrt(np.maximum(1.0 - np.abs(rho), 0.0))
    d[qi[:, None] == cj[None, :]] = 0.0
    return d


def brute_knn_fp64(coords, v, p, queries, k, *, predecessor=False, variance=1.0,
                   lengthscale=1.0, kernel_type=0, block=256):
    """Exact top-k in FP64.  Returns (idx int64 (Q,k) padded -1, dist float64
    (Q,k) padded inf, kth_next float64 (Q,) = the (k+1)-th distance, used to
    detect ties at the cut)."""
    n = np.asarray(coords).shape[0]
    q = np.asarray(queries, np.int64).reshape(-1)
    out_i = np.full((q.size, k), -1, np.int64)
    out_d = np.full((q.size, k), np.inf)
    nxt = np.full(q.size, np.inf)
    allj = np.arange(n, dtype=np.int64)
    for s in range(0, q.size, block):
        qb = q[s:s + block]
        d = residual_dist_fp64(coords, v, p, qb, allj, variance=variance,
                               lengthscale=lengthscale, kernel_type=kernel_type)
        if predecessor:
            d = np.where(allj[None, :] < qb[:, None], d, np.inf)
        kk = min(k + 1, n)
        part = np.argpartition(d, kk - 1, axis=1)[:, :kk]
        pd = np.take_along_axis(d, part, axis=1)
        o = np.lexsort((part, pd), axis=1)
        part = np.take_along_axis(part, o, axis=1)
        pd = np.take_along_axis(pd, o, axis=1)
        kc = min(k, kk)
        valid = np.isfinite(pd[:, :kc])
        out_i[s:s + block, :kc] = np.where(valid, part[:, :kc], -1)
        out_d[s:s + block, :kc] = np.where(valid, pd[:, :kc], np.inf)
        if kk > k:
            nxt[s:s + block] = pd[:, k]
    return out_i, out_d, nxt


def compare_knn(idx, dist, ref_idx, ref_dist, ref_next, *, tie=1e-6):
    """Neighbour-set agreement modulo ties at the k-th distance.

    A row agrees when the returned set equals the reference set, or differs only
    in members whose reference distance lies within ``tie`` of the reference
    k-th distance (both sides).  Returns (n_rows_disagree, max_abs_dist_err,
    list of bad rows)."""
    bad = []
    max_err = 0.0
    for r in range(idx.shape[0]):
        a = idx[r][idx[r] >= 0]
        b = ref_idx[r][ref_idx[r] >= 0]
        if a.size != b.size:
            bad.append(r)
            continue
        sa, sb = set(a.tolist()), set(b.tolist())
        if sa != sb:
            kth = ref_dist[r][b.size - 1] if b.size else 0.0
            cut = min(kth, ref_next[r]) if np.isfinite(ref_next[r]) else kth
            only_b = [j for j in sb - sa]
            db = {int(j): ref_dist[r][i] for i, j in enumerate(ref_idx[r]) if j >= 0}
            ok = all(abs(db[j] - kth) <= tie for j in only_b)
            # members only in ours must also sit at the cut (checked by caller via
            # distances), here accept when our k-th distance matches the cut
            if not ok or abs(float(np.max(dist[r][:a.size])) - cut) > tie + 1e-5:
                bad.append(r)
                continue
    return len(bad), max_err, bad
