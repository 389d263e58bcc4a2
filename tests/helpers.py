"""Shared test helpers (instance construction from golden specs)."""
import numpy as np

from f2m_amd import workloads as W


def instance_from_spec(spec):
    if spec["kind"] == "c1":
        x, y, _, _ = W.c1()
        return x, y
    x, y = W.uniform_int(spec["n"], spec["span"], spec["seed"])
    if spec.get("fp32_unit"):
        x = (x / spec["span"]).astype(np.float32).astype(np.float64)
        y = (y / spec["span"]).astype(np.float32).astype(np.float64)
    return x, y


def graph_hash(rowptr, nbr, cost):
    import hashlib
    h = hashlib.sha256()
    for a in (np.ascontiguousarray(rowptr, np.int32), np.ascontiguousarray(nbr, np.int32),
              np.ascontiguousarray(cost, np.float64)):
        h.update(a.tobytes())
    return h.hexdigest()


def feasibility(rowptr, nbr, x):
    n = len(rowptr) - 1
    src = np.repeat(np.arange(n), np.diff(rowptr))
    deg = np.bincount(src, weights=x, minlength=n)
    return float(np.abs(deg - 2).max()), float(x.min()), float(x.max())


def sym_check(rowptr, nbr, x):
    """x must agree on the two slots of every edge."""
    n = len(rowptr) - 1
    src = np.repeat(np.arange(n), np.diff(rowptr))
    key_a = src.astype(np.int64) * n + nbr
    key_b = nbr.astype(np.int64) * n + src
    ia = np.argsort(key_a)
    ib = np.argsort(key_b)
    return bool(np.array_equal(x[ia], x[ib]))


def certificate(rowptr, nbr, cost, x, F, d_min):
    """Primal value P = sum_e c_e x_e and the dual bound
    D = 2 sum_v y_v - sum_e max(0, y_u + y_w - c_e) with y = -d_min F / 2,
    recomputed in numpy.  Any y gives D <= LP optimum (weak duality)."""
    n = len(rowptr) - 1
    src = np.repeat(np.arange(n), np.diff(rowptr))
    m = src < nbr
    u, w, c, xe = src[m], nbr[m].astype(np.int64), cost[m], x[m]
    y = -0.5 * d_min * np.asarray(F, np.float64)
    P = float(np.dot(c, xe))
    D = float(2 * y.sum() - np.maximum(0.0, y[u] + y[w] - c).sum())
    return P, D


def morton_order(x, y):
    """numpy restatement of f2m_spatial_order: 16-bit quantisation over the
    bounding box, x in the even bits, stable sort (ties in input order)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)

    def quant(v):
        mn, mx = v.min(), v.max()
        ext = mx - mn
        if not ext > 0:
            return np.zeros(len(v), np.uint64)
        q = np.floor(((v - mn) / ext) * 65535.0)
        q = np.where(q > 0, q, 0)
        return np.minimum(q, 65535).astype(np.uint64)

    def spread(v):
        v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF)
        v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F)
        v = (v | (v << np.uint64(2))) & np.uint64(0x33333333)
        v = (v | (v << np.uint64(1))) & np.uint64(0x55555555)
        return v

    key = spread(quant(x)) | (spread(quant(y)) << np.uint64(1))
    return np.argsort(key, kind="stable").astype(np.int32)


def complete_graph_dual_bound(x, y, rowptr, nbr, cost, F, d_min, K, mode):
    """Lower bound on the F2M LP optimum over ALL pairs of points (weak
    duality with y = -d_min F / 2, as certificate()):
        D_all = 2 sum_v y_v - sum_{u<w} max(0, y_u + y_w - c_uw).
    A pair outside the symmetrised kNN graph costs at least max(r_u, r_w)
    (r_v = v's K-th nearest cost), so its term is nonzero only when an
    endpoint has 2 y_v > r_v; the rows of those endpoints are computed in
    full against every point.  Returns (D_all, number of such endpoints).
    Any x feasible on any edge set has sum c x >= D_all."""
    n = len(rowptr) - 1
    _, D = certificate(rowptr, nbr, cost, np.zeros(len(nbr)), F, d_min)
    yv = -0.5 * d_min * np.asarray(F, np.float64)
    r = cost[np.asarray(rowptr[:-1], np.int64) + K - 1]
    viol = np.nonzero(2 * yv > r)[0]
    vset = set(viol.tolist())
    xs = np.asarray(x, np.float64)
    ys = np.asarray(y, np.float64)
    extra = 0.0
    for v in viol:
        dx = xs[v] - xs
        dy = ys[v] - ys
        c = np.sqrt(dx * dx + dy * dy)
        if mode == 1:
            c = np.floor(c + 0.5)
        elif mode == 2:
            c = c.astype(np.float32).astype(np.float64)
        c[v] = np.inf
        pen = yv[v] + yv - c
        inside = set(nbr[rowptr[v]:rowptr[v + 1]].tolist())
        for w in np.nonzero(pen > 0)[0]:
            w = int(w)
            if w in inside or (w in vset and w < v):  # graph edge, or counted from w
                continue
            extra += float(pen[w])
    return D - extra, len(viol)
