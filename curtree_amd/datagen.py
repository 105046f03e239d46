"""Deterministic synthetic box and window sets for the benchmark workloads.

Every generator is seeded (numpy PCG64) and builds boxes in FP32 as
center -/+ fl(size/2).  Sizes and distributions:

  uniform_boxes      centers U[0,1)^2, widths/heights U[0, wmax]
                     (1e6 boxes, wmax 1e-3: the CPU-runnable plain workload)
  uniform_windows    centers U[0,1)^2, square side U[0, smax]
  mixture_boxes      64-component Gaussian mixture, means U[0,1]^2,
                     sigma U[0.01, 0.05], equal weights, rejected to [0,1]^2;
                     widths/heights log-normal(mu=-9, sigma=1)
  loguniform_windows centers U[0,1)^2, square side log-uniform [lo, hi]
  clustered_boxes    (1-frac) uniform + frac in k Gaussian clusters
                     (sigma 1e-3); widths/heights U[0, wmax]
  skewed_boxes       frac_hot in hot cells of side 1e-3 chosen with
                     Zipf(1.2) weights, rest uniform; a share are points
  skewed_boxes_large the same distribution, generated in chunks (1e9 boxes)
  mixed_queries      centers U[0,1)^2; half points, half squares of side
                     log-uniform [lo, hi]
  roadlike_boxes     line-segment MBRs: uniform start, Pareto length,
                     random angle
"""
import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _boxes(cx, cy, w, h):
    cx = cx.astype(np.float32, copy=False)
    cy = cy.astype(np.float32, copy=False)
    hw = (w.astype(np.float32, copy=False) / np.float32(2))
    hh = (h.astype(np.float32, copy=False) / np.float32(2))
    return (cx - hw, cy - hh, cx + hw, cy + hh)


def uniform_boxes(n, seed=1, wmax=1e-3):
    r = _rng(seed)
    cx = r.random(n, dtype=np.float32)
    cy = r.random(n, dtype=np.float32)
    w = r.random(n, dtype=np.float32) * np.float32(wmax)
    h = r.random(n, dtype=np.float32) * np.float32(wmax)
    return _boxes(cx, cy, w, h)


def uniform_windows(q, seed=2, smax=0.01):
    r = _rng(seed)
    cx = r.random(q, dtype=np.float32)
    cy = r.random(q, dtype=np.float32)
    s = r.random(q, dtype=np.float32) * np.float32(smax)
    return _boxes(cx, cy, s, s)


def loguniform_windows(q, seed=4, lo=1e-4, hi=1e-2):
    r = _rng(seed)
    cx = r.random(q, dtype=np.float32)
    cy = r.random(q, dtype=np.float32)
    s = np.exp(r.uniform(np.log(lo), np.log(hi), q)).astype(np.float32)
    return _boxes(cx, cy, s, s)


def _mixture_centers(r, n, k, sig_lo, sig_hi):
    means = r.random((k, 2))
    sig = r.uniform(sig_lo, sig_hi, k)
    cx = np.empty(n, dtype=np.float32)
    cy = np.empty(n, dtype=np.float32)
    todo = np.arange(n)
    while todo.size:
        comp = r.integers(0, k, todo.size)
        x = means[comp, 0] + sig[comp] * r.standard_normal(todo.size)
        y = means[comp, 1] + sig[comp] * r.standard_normal(todo.size)
        ok = (x >= 0) & (x <= 1) & (y >= 0) & (y <= 1)
        cx[todo[ok]] = x[ok]
        cy[todo[ok]] = y[ok]
        todo = todo[~ok]
    return cx, cy


def mixture_boxes(n, seed=3, k=64):
    r = _rng(seed)
    cx, cy = _mixture_centers(r, n, k, 0.01, 0.05)
    w = np.exp(-9.0 + r.standard_normal(n)).astype(np.float32)
    h = np.exp(-9.0 + r.standard_normal(n)).astype(np.float32)
    return _boxes(cx, cy, w, h)


def clustered_boxes(n, seed=5, frac=0.1, k=5, sigma=1e-3, wmax=1e-3):
    r = _rng(seed)
    nc = int(n * frac)
    nu = n - nc
    cx = np.empty(n, dtype=np.float32)
    cy = np.empty(n, dtype=np.float32)
    cx[:nu] = r.random(nu, dtype=np.float32)
    cy[:nu] = r.random(nu, dtype=np.float32)
    means = r.uniform(0.1, 0.9, (k, 2))
    comp = r.integers(0, k, nc)
    cx[nu:] = np.clip(means[comp, 0] + sigma * r.standard_normal(nc), 0, 1)
    cy[nu:] = np.clip(means[comp, 1] + sigma * r.standard_normal(nc), 0, 1)
    perm = r.permutation(n)
    cx, cy = cx[perm], cy[perm]
    w = r.random(n, dtype=np.float32) * np.float32(wmax)
    h = r.random(n, dtype=np.float32) * np.float32(wmax)
    return _boxes(cx, cy, w, h)


def skewed_boxes(n, seed=9, cells=1000, frac_hot=0.7, cell=1e-3, zipf=1.2, points=0.2,
                 wmax=1e-4):
    r = _rng(seed)
    nh = int(n * frac_hot)
    cx = np.empty(n, dtype=np.float32)
    cy = np.empty(n, dtype=np.float32)
    corner = r.random((cells, 2)) * (1 - cell)
    wts = 1.0 / np.arange(1, cells + 1) ** zipf
    wts /= wts.sum()
    c = r.choice(cells, nh, p=wts)
    cx[:nh] = corner[c, 0] + cell * r.random(nh)
    cy[:nh] = corner[c, 1] + cell * r.random(nh)
    cx[nh:] = r.random(n - nh, dtype=np.float32)
    cy[nh:] = r.random(n - nh, dtype=np.float32)
    perm = r.permutation(n)
    cx, cy = cx[perm], cy[perm]
    w = r.random(n, dtype=np.float32) * np.float32(wmax)
    h = r.random(n, dtype=np.float32) * np.float32(wmax)
    pt = r.random(n) < points
    w[pt] = 0
    h[pt] = 0
    return _boxes(cx, cy, w, h)


def mixed_queries(q, seed=10, lo=1e-5, hi=1e-3, points=0.5):
    """Query boxes: centers U[0,1)^2; a share `points` are points (side 0),
    the others squares of side log-uniform [lo, hi]."""
    r = _rng(seed)
    cx = r.random(q, dtype=np.float32)
    cy = r.random(q, dtype=np.float32)
    s = np.exp(r.uniform(np.log(lo), np.log(hi), q)).astype(np.float32)
    s[r.random(q) < points] = 0
    return _boxes(cx, cy, s, s)


def skewed_boxes_large(n, seed=9, chunk=100_000_000, progress=None, cells=1000, frac_hot=0.7,
                       cell=1e-3, zipf=1.2, points=0.2, wmax=1e-4):
    """skewed_boxes' distribution (same hot cells for every chunk) generated in
    chunks of `chunk` boxes, each with its own stream, so that 1e9 boxes fit
    in host memory with few temporaries; `progress(done)` after each chunk.
    The values differ from skewed_boxes(n, seed)."""
    r0 = _rng(seed)
    corner = r0.random((cells, 2)) * (1 - cell)
    wts = 1.0 / np.arange(1, cells + 1) ** zipf
    wts /= wts.sum()
    out = tuple(np.empty(n, dtype=np.float32) for _ in range(4))
    for k, s in enumerate(range(0, n, chunk)):
        m = min(chunk, n - s)
        r = _rng(seed * 1000 + k + 1)
        nh = int(m * frac_hot)
        cx = np.empty(m, dtype=np.float32)
        cy = np.empty(m, dtype=np.float32)
        c = r.choice(cells, nh, p=wts)
        cx[:nh] = corner[c, 0] + cell * r.random(nh)
        cy[:nh] = corner[c, 1] + cell * r.random(nh)
        cx[nh:] = r.random(m - nh, dtype=np.float32)
        cy[nh:] = r.random(m - nh, dtype=np.float32)
        perm = r.permutation(m)
        w = r.random(m, dtype=np.float32) * np.float32(wmax)
        h = r.random(m, dtype=np.float32) * np.float32(wmax)
        pt = r.random(m) < points
        w[pt] = 0
        h[pt] = 0
        for o, v in zip(out, _boxes(cx[perm], cy[perm], w, h)):
            o[s:s + m] = v
        if progress:
            progress(s + m)
    return out


def roadlike_boxes(n, seed=7, lmin=1e-5, alpha=1.5):
    r = _rng(seed)
    x0 = r.random(n)
    y0 = r.random(n)
    length = lmin * (1.0 + r.pareto(alpha, n))
    length = np.minimum(length, 0.05)
    ang = r.uniform(0, 2 * np.pi, n)
    x1 = np.clip(x0 + length * np.cos(ang), 0, 1)
    y1 = np.clip(y0 + length * np.sin(ang), 0, 1)
    return (np.minimum(x0, x1).astype(np.float32), np.minimum(y0, y1).astype(np.float32),
            np.maximum(x0, x1).astype(np.float32), np.maximum(y0, y1).astype(np.float32))
