"""Seeded synthetic instances for the five benchmark configurations.

Each generator returns float64 coordinate arrays (x, y) of the points after
exact-duplicate removal (first occurrence kept, as the TSPLIB reader does,
FXY.hpp:262-270), the cost mode and K.  Seeds are 0xF2A0 + config number;
numpy's PCG64 makes the points identical on every machine with this numpy.

  C1  n=1000 integer points uniform in [0,1000]^2, NINT costs, K=20
  C2  n=100000 fp32 points, 50-component isotropic Gaussian mixture
      (centres U[0,1e6]^2, sigma U[5e3,3e4]), NINT costs, K=20
  C3  n=1000000 lattice points, pitch 10: the square [0,L)^2 with
      L = 10*ceil(1.05*sqrt(1e6/0.42)), rows kept with p=0.7, cells with
      p=0.6, first 1e6 points in row-major order, NINT costs, K=10
  C4  4096 instances, n = floor(exp(U[ln 100, ln 20000])), fp32 points
      U[0,1)^2 scaled by 1e6, NINT costs, K=16
  C5  256 instances of n=2000 fp32 points U[0,1)^2, RAW_FP32 costs, K=20
"""
from __future__ import annotations

import math

import numpy as np

COST_RAW_FP64 = 0
COST_NINT = 1
COST_RAW_FP32 = 2

SEED0 = 0xF2A0


def dedup(x: np.ndarray, y: np.ndarray):
    """Drop exact duplicate points, keeping the first occurrence."""
    xy = np.stack([np.asarray(x, np.float64), np.asarray(y, np.float64)], 1)
    _, first = np.unique(xy, axis=0, return_index=True)
    keep = np.sort(first)
    return xy[keep, 0].copy(), xy[keep, 1].copy()


def c1(n: int = 1000, seed: int = SEED0 + 1):
    rng = np.random.default_rng(seed)
    p = rng.integers(0, 1001, size=(n, 2))
    x, y = dedup(p[:, 0], p[:, 1])
    return x, y, COST_NINT, 20


def c2(n: int = 100_000, seed: int = SEED0 + 2, comps: int = 50):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1e6, size=(comps, 2))
    sig = rng.uniform(5e3, 3e4, size=comps)
    which = rng.integers(0, comps, size=n)
    z = rng.standard_normal(size=(n, 2))
    p = (centres[which] + sig[which, None] * z).astype(np.float32)
    x, y = dedup(p[:, 0], p[:, 1])
    return x, y, COST_NINT, 20


def c3(n: int = 1_000_000, seed: int = SEED0 + 3):
    rng = np.random.default_rng(seed)
    side = int(math.ceil(1.05 * math.sqrt(n / 0.42)))  # lattice columns/rows
    rows = rng.random(side) < 0.7
    cells = rng.random((side, side)) < 0.6
    keep = cells & rows[:, None]
    iy, ix = np.nonzero(keep)  # row-major order
    iy, ix = iy[:n], ix[:n]
    return (10.0 * ix).astype(np.float64), (10.0 * iy).astype(np.float64), COST_NINT, 10


def c4_sizes(count: int = 4096, seed: int = SEED0 + 4):
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(math.log(100), math.log(20000), size=count))).astype(np.int64)


def c4(count: int = 4096, seed: int = SEED0 + 4, only=None):
    """List of (x, y) per instance; cost mode NINT, K=16.  With `only` (instance
    indices), just those instances, in that order: every instance's points are
    still drawn in sequence, so each equals its copy in the full workload."""
    sizes = c4_sizes(count, seed)
    rng = np.random.default_rng(seed + 1000)
    pts = []
    for n in sizes:
        pts.append((rng.random((int(n), 2), dtype=np.float32) * np.float32(1e6)).astype(np.float32))
    idx = range(count) if only is None else [int(i) for i in only]
    out = [dedup(pts[i][:, 0], pts[i][:, 1]) for i in idx]
    return out, COST_NINT, 16


def c5(count: int = 256, n: int = 2000, seed: int = SEED0 + 5):
    """List of (x, y) per instance (fp32 values held in float64); RAW_FP32, K=20."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        p = rng.random((n, 2), dtype=np.float32)
        out.append(dedup(p[:, 0], p[:, 1]))
    return out, COST_RAW_FP32, 20


def uniform_int(n: int, span: int, seed: int):
    """Small test instances: integer points uniform in [0, span]^2."""
    rng = np.random.default_rng(seed)
    p = rng.integers(0, span + 1, size=(n, 2))
    return dedup(p[:, 0], p[:, 1])
