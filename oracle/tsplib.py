"""TSPLIB reader restated in numpy (FXY.hpp:186-309) -- test infrastructure only.

Lines before NODE_COORD_SECTION are ignored, reading stops at a line that
starts with EOF; each data line is `label x y`.  Coordinates are rounded to
1e-3 with std::round semantics (half away from zero) on x*1000, exact
duplicate points are dropped keeping the first label (FXY.hpp:252-270).
Returns labels (int64), x, y (float64) and the duplicate map {kept: [dropped]}.
"""
from __future__ import annotations

import numpy as np


def _round_away(v):
    return np.sign(v) * np.floor(np.abs(v) + 0.5)


def read(path_or_text, is_text=False):
    text = path_or_text if is_text else open(path_or_text, "r", errors="replace").read()
    toks = []
    started = False
    for line in text.splitlines():
        if line.startswith("EOF"):
            break
        if started:
            toks.extend(line.split())
            continue
        if line.startswith("NODE_COORD_SECTION"):
            started = True
    if not started or not toks:
        raise ValueError("not a TSPLIB coordinate file")
    toks = toks[: len(toks) // 3 * 3]
    a = np.array(toks, dtype=object).reshape(-1, 3)
    lab = a[:, 0].astype(np.int64)
    x = _round_away(a[:, 1].astype(np.float64) * 1000) / 1000
    y = _round_away(a[:, 2].astype(np.float64) * 1000) / 1000
    seen = {}
    keep = []
    dups = {}
    for i in range(len(lab)):
        k = (x[i], y[i])
        if k in seen:
            dups.setdefault(int(lab[seen[k]]), []).append(int(lab[i]))
            continue
        seen[k] = i
        keep.append(i)
    keep = np.array(keep, np.int64)
    return lab[keep], x[keep], y[keep], dups
