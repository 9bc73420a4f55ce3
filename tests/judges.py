"""Surface judges used by the accuracy tests.

surface_points: zero crossings of phi on grid edges between defined
neighbours, linearly interpolated -- the node set of the marching-cubes
surface that CompareChamfer extracts (lsCompareChamfer.hpp:104-308 via
lsToSurfaceMesh.hpp), without the minimum-node-distance merging.
chamfer: (sum of forward + backward nearest-neighbour distances) / (nT + nS)
(lsCompareChamfer.hpp:283-285).
"""
import numpy as np
from scipy.spatial import cKDTree

POS = np.finfo(np.float64).max


def surface_points(idx, val, dx):
    idx = np.asarray(idx, dtype=np.int64)
    val = np.asarray(val)
    m = np.abs(val) != POS
    I, V = idx[m], val[m]
    D = I.shape[1]
    off = I.min(0) - 1
    span = I.max(0) - off + 2
    key = np.zeros(len(I), np.int64)
    mul = 1
    for d in range(D):
        key += (I[:, d] - off[d]) * mul
        mul *= int(span[d])
    order = np.argsort(key)
    ks = key[order]
    pts = []
    mul = 1
    for d in range(D):
        nk = key + mul
        pos = np.searchsorted(ks, nk)
        pos = np.minimum(pos, len(ks) - 1)
        hit = ks[pos] == nk
        a = np.where(hit)[0]
        b = order[pos[hit]]
        va, vb = V[a], V[b]
        s = (va > 0) != (vb > 0)
        a, b, va, vb = a[s], b[s], va[s], vb[s]
        t = va / (va - vb)
        p = I[a].astype(float) + t[:, None] * (I[b] - I[a])
        pts.append(p * dx)
        mul *= int(span[d])
    return np.concatenate(pts) if pts else np.zeros((0, D))


def chamfer(P, Q):
    dPQ, _ = cKDTree(Q).query(P)
    dQP, _ = cKDTree(P).query(Q)
    return (dPQ.sum() + dQP.sum()) / (len(P) + len(Q))


def volume_mismatch(defA, defB, dx):
    """Cells whose sign differs (negative = inside) times dx^D, over the union
    of the two defined-point sets (lsCompareVolume.hpp:176-178,271-289 uses
    corner averages; this is the grid-point version)."""
    ia, va = defA
    ib, vb = defB
    D = ia.shape[1]
    ka = {tuple(p): v for p, v in zip(map(tuple, ia), va)}
    kb = {tuple(p): v for p, v in zip(map(tuple, ib), vb)}
    diff = 0
    for k in set(ka) | set(kb):
        a = ka.get(k)
        b = kb.get(k)
        if a is None or b is None:
            continue
        diff += (a < 0) != (b < 0)
    return diff * dx ** D
