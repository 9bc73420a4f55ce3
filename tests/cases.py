"""Seeded level-set inputs shared by the parity tests (built with the oracle's
restatement of MakeGeometry / insertPoints and fed identically to every
implementation)."""
import numpy as np

import pyoracle as po

REF, INF, PER = 0, 1, 2


def plane_points(D, gmin, gmax, bc, dx, z0, axis=None):
    """Band |phi| <= 1 of the plane x_axis*dx = z0 (normal +e_axis) over the
    bounded dimensions."""
    axis = D - 1 if axis is None else axis
    rng = []
    for d in range(D):
        if d == axis:
            continue
        hi = gmax[d] - (1 if bc[d] == PER else 0)
        rng.append(np.arange(gmin[d], hi + 1))
    ks = np.arange(int(np.floor(z0 / dx)) - 2, int(np.ceil(z0 / dx)) + 3)
    grids = np.meshgrid(*rng, ks, indexing="ij")
    coords = [g.ravel() for g in grids]
    phi = (coords[-1] * dx - z0) / dx
    m = np.abs(phi) <= 1.0
    idx = np.zeros((m.sum(), D), np.int32)
    j = 0
    for d in range(D):
        if d == axis:
            idx[:, d] = coords[-1][m]
        else:
            idx[:, d] = coords[j][m]
            j += 1
    return idx, phi[m]


def box_points(D, mn, mx, dx, gmin=None, gmax=None):
    """Band |phi| <= 1 of the signed Euclidean distance to an axis-aligned box."""
    lo = [int(np.floor(mn[d] / dx)) - 2 for d in range(D)]
    hi = [int(np.ceil(mx[d] / dx)) + 2 for d in range(D)]
    if gmin is not None:
        lo = [max(lo[d], gmin[d]) for d in range(D)]
        hi = [min(hi[d], gmax[d]) for d in range(D)]
    grids = np.meshgrid(*[np.arange(lo[d], hi[d] + 1) for d in range(D)], indexing="ij")
    c = np.stack([g.ravel() for g in grids], 1)
    p = c * dx
    mn_, mx_ = np.array(mn, float), np.array(mx, float)
    e = np.maximum(np.maximum(mn_ - p, p - mx_), 0)
    inside = np.all((p >= mn_) & (p <= mx_), 1)
    din = np.min(np.minimum(p - mn_, mx_ - p), 1)
    phi = np.where(inside, -din / dx, np.sqrt((e ** 2).sum(1)) / dx)
    m = np.abs(phi) <= 1.0
    return c[m].astype(np.int32), phi[m]


def sphere(D=3, r=10.0, dx=0.25, origin=None, ext=None):
    origin = origin or [0.0] * D
    ext = ext or int(np.ceil((r * 1.8) / dx)) + 8
    d = po.Domain(D, [-ext] * D, [ext] * D, [INF] * D, dx)
    d.make_sphere(origin, r)
    return d


def trench(L=32, width=16, depth=16, dx=1.0):
    """Plane z=0 (material below) minus a centred box trench, reflective x/y."""
    gmin, gmax, bc = [-L, -L, -L], [L, L, L], [REF, REF, INF]
    d = po.Domain(3, gmin, gmax, bc, dx)
    ip, vp = plane_points(3, gmin, gmax, bc, dx, 0.0)
    d.from_points(ip, vp, 2)
    b = po.Domain(3, gmin, gmax, bc, dx)
    h = width / 2
    ib, vb = box_points(3, [-h, -h, -depth], [h, h, 1.0], dx, gmin, gmax)
    b.from_points(ib, vb, 2)
    d.boolean(b, 2)
    return d


def wavy_periodic(N=32, amp=2.0, dx=1.0):
    """Periodic x/y sinusoidal surface z = h(x, y), infinite z."""
    gmin, gmax, bc = [0, 0, -20], [N, N, 20], [PER, PER, INF]
    d = po.Domain(3, gmin, gmax, bc, dx)
    x, y = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    h = amp * np.sin(2 * np.pi * x / N) * np.cos(2 * np.pi * y / N) + 0.3
    pts, vals = [], []
    g = np.sqrt(1 + (2 * np.pi / N * amp) ** 2)
    for z in range(-int(amp) - 4, int(amp) + 5):
        phi = (z - h) / g
        m = np.abs(phi) <= 1
        pts.append(np.stack([x[m], y[m], np.full(m.sum(), z)], 1))
        vals.append(phi[m])
    d.from_points(np.concatenate(pts).astype(np.int32), np.concatenate(vals), 2)
    return d


def sphere2d(r=5.0, dx=0.5):
    d = po.Domain(2, [-20, -20], [20, 20], [REF, INF], dx)
    d.make_sphere([0.1, 0.2], r)
    return d


# velocity configurations (the built-in functor parameters)
VEL_GROW = dict(kind=0, b=1.0)
VEL_ETCH_DIRECTIONAL = dict(kind=1, dir=2, k=1, b=-0.1, a_pos=-1.0, a_neg=0.0)   # -(max(0,n_z) + 0.1)
VEL_NORMAL_SQ = dict(kind=1, dir=2, k=2, b=0.05, a_pos=1.0, a_neg=0.0)           # max(0,n_z)^2 + 0.05
VEL_EGG = dict(kind=1, dir=0, k=2, b=1.0, a_pos=2.3, a_neg=0.5)                  # tests/Advection/Advection.cpp:30-32
VEL_TRANSLATE_X = dict(kind=0, b=0.0, vec=(1.0, 0.0, 0.0))                      # tests/TimeIntegration


def bit_equal(ia, va, ib, vb):
    return ia.shape == ib.shape and np.array_equal(ia, ib) and np.array_equal(va.view(np.int64), vb.view(np.int64))


def first_diff(ia, va, ib, vb):
    n = min(len(va), len(vb))
    for i in range(n):
        if not (np.array_equal(ia[i], ib[i]) and va.view(np.int64)[i] == vb.view(np.int64)[i]):
            return i, ia[i], va[i], ib[i], vb[i]
    return n, None, None, None, None
