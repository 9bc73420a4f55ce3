"""Conversion of a canonical run-start list to ViennaHRLE's run structure.

A band is stored canonically: every defined point plus every undefined point
whose value differs from the previous grid point's value or follows a
defined point.  ViennaHRLE (the external dependency behind Domain,
lsDomain.hpp:8-9) encodes runs hierarchically: every row (all dimensions but
x) that holds a defined point starts its own run list at the grid's first x
index, and the row after it starts a new run as well; the same holds one
level up for z-slabs in 3D.  Iterating a ViennaHRLE domain therefore visits
these extra run starts (tests/BooleanOperationExactZero/result.hpp shows
them, e.g. {-10, 3} and {-10, 11}).  They carry the value of the run they
split, so the level-set value field is the same in both forms.
"""
import numpy as np

POS_VALUE = float(np.finfo(np.float64).max)


def _keys(idx, D):
    # lexicographic key with the highest dimension most significant
    return [idx[:, d] for d in range(D)]


def to_hrle_runs(idx, val, grid_lo, grid_hi):
    """idx (n, D) int32 canonical run starts (entry 0 = grid minimum), val (n,).
    Returns (idx, val) with ViennaHRLE's additional row/slab run starts."""
    idx = np.asarray(idx)
    val = np.asarray(val)
    n, D = idx.shape
    lo = np.asarray(grid_lo[:D])
    hi = np.asarray(grid_hi[:D])
    defined = np.abs(val) != POS_VALUE
    extra = []
    rows = np.unique(idx[defined][:, 1:], axis=0) if defined.any() else np.zeros((0, D - 1), np.int64)
    for r in rows:
        p = np.concatenate([[lo[0]], r])
        extra.append(p)
        q = p.copy()
        for d in range(1, D):  # next row in lexicographic order
            if q[d] < hi[d]:
                q[d] += 1
                extra.append(q)
                break
            q[d] = lo[d]
    if D == 3:
        slabs = np.unique(idx[defined][:, 2]) if defined.any() else []
        for z in slabs:
            extra.append(np.array([lo[0], lo[1], z]))
            if z < hi[2]:
                extra.append(np.array([lo[0], lo[1], z + 1]))
    if not extra:
        return idx, val
    ex = np.unique(np.array(extra, dtype=np.int64), axis=0)
    # value of the run containing each extra point: last canonical entry <= p
    order = np.lexsort(_keys(idx.astype(np.int64), D))
    assert np.all(order == np.arange(n)), "canonical list must be sorted"
    allk = np.concatenate([idx.astype(np.int64), ex])
    src = np.concatenate([np.arange(n), -np.ones(len(ex), np.int64)])
    o = np.lexsort(tuple(reversed([allk[:, D - 1 - d] for d in range(D)])))
    o = np.lexsort(_keys(allk, D))
    allk, src = allk[o], src[o]
    out_i, out_v = [], []
    last = -1
    prev_key = None
    for k, s in zip(allk, src):
        key = tuple(k)
        if s >= 0:
            last = s
            if prev_key == key and out_i:  # an extra point equal to a canonical entry: keep the entry
                out_i[-1], out_v[-1] = k, val[s]
            else:
                out_i.append(k)
                out_v.append(val[s])
        else:
            if prev_key == key:
                continue
            out_i.append(k)
            out_v.append(val[last])
        prev_key = key
    return np.array(out_i, dtype=np.int32), np.array(out_v)
