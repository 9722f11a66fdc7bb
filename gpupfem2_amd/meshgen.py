"""Mesh generators for the benchmark workloads (host side, numpy).

The reference only reads SALOME DAT files (src/mesh_2d.cu:36-96); the workloads here need generated
meshes. Every generator is deterministic for a given seed (numpy's default_rng) and returns
(vertices, cells) as float64 / int32 arrays with 0-based vertex indices. ``write_dat`` writes the
reference's DAT format (without a trailing newline, see the reader's end-of-file rule).
"""
import math

import numpy as np


def unit_square(n=64, jitter=0.0, seed=0):
    """n x n squares on [0,1]^2, each split by its "/" diagonal; node (i, j) -> j*(n+1)+i."""
    return rectangle(0.0, 1.0, 0.0, 1.0, n, n, jitter=jitter, seed=seed)


def rectangle(x0, x1, y0, y1, nx, ny, jitter=0.0, seed=0):
    """Structured nx x ny quads split into 2 triangles each; interior nodes jittered by
    U(-jitter*h, jitter*h) in each direction (boundary nodes stay on their lines)."""
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    X, Y = np.meshgrid(xs, ys)
    xy = np.stack([X.ravel(), Y.ravel()], axis=1).astype(np.float64)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        hx, hy = (x1 - x0) / nx, (y1 - y0) / ny
        d = rng.uniform(-jitter, jitter, size=(xy.shape[0], 2)) * np.array([hx, hy])
        ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
        interior = ((ii > 0) & (ii < nx) & (jj > 0) & (jj < ny)).ravel()
        xy[interior] += d[interior]
    i, j = np.meshgrid(np.arange(nx), np.arange(ny))
    p00 = (j * (nx + 1) + i).ravel()
    p10 = p00 + 1
    p01 = p00 + nx + 1
    p11 = p01 + 1
    t1 = np.stack([p00, p10, p11], axis=1)
    t2 = np.stack([p00, p11, p01], axis=1)
    tri = np.empty((2 * p00.size, 3), np.int32)
    tri[0::2] = t1
    tri[1::2] = t2
    return xy, tri


def _geometric(n, first, length):
    """n segment lengths growing geometrically from `first`, summing to `length`."""
    if first * n >= length:
        return np.full(n, length / n)
    lo, hi = 1.0 + 1e-12, 2.0
    f = lambda q: first * (q ** n - 1.0) / (q - 1.0) - length
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
    q = 0.5 * (lo + hi)
    seg = first * q ** np.arange(n)
    return seg * (length / seg.sum())


def cylinder_channel(h=1.0 / 112, D=1.0, L=20.0, H=4.0, cx=4.0, cy=2.1, box=1.0, radial_layers=None,
                     first_layer=None):
    """Channel [0,L]x[0,H] around a cylinder of diameter D centred at (cx, cy).

    A tensor-product background grid of spacing ~h (grid lines through the edges of the square
    [cx-box, cx+box] x [cy-box, cy+box]) with that square cut out, plus an O-grid from the cylinder
    surface to the square: rays from the centre to every square-boundary node, graded layers from the
    first-layer thickness (default: the angular spacing on the cylinder) out to the square. Quads are
    split into two triangles. Returns (xy, tri, info) with the node index sets of the cylinder.
    """
    R = 0.5 * D
    xa, xb, ya, yb = cx - box, cx + box, cy - box, cy + box

    def seg(a, b, hh):
        n = max(1, int(round((b - a) / hh)))
        return np.linspace(a, b, n + 1)

    xs = np.concatenate([seg(0.0, xa, h)[:-1], seg(xa, xb, h)[:-1], seg(xb, L, h)])
    ys = np.concatenate([seg(0.0, ya, h)[:-1], seg(ya, yb, h)[:-1], seg(yb, H, h)])
    nx, ny = xs.size - 1, ys.size - 1
    ia = int(np.argmin(abs(xs - xa)))
    ib = int(np.argmin(abs(xs - xb)))
    ja = int(np.argmin(abs(ys - ya)))
    jb = int(np.argmin(abs(ys - yb)))
    # background nodes: all grid nodes except those strictly inside the square
    I, J = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    inside = (I > ia) & (I < ib) & (J > ja) & (J < jb)
    keep = ~inside
    gid = -np.ones((ny + 1, nx + 1), np.int64)
    gid[keep] = np.arange(keep.sum())
    X, Y = np.meshgrid(xs, ys)
    xy_bg = np.stack([X[keep], Y[keep]], axis=1)
    # background quads outside the square
    qi, qj = np.meshgrid(np.arange(nx), np.arange(ny))
    qin = (qi >= ia) & (qi < ib) & (qj >= ja) & (qj < jb)
    qi, qj = qi[~qin], qj[~qin]
    p00 = gid[qj, qi]
    p10 = gid[qj, qi + 1]
    p01 = gid[qj + 1, qi]
    p11 = gid[qj + 1, qi + 1]
    # square boundary, counter-clockwise from (xa, ya)
    ring = [(ja, i) for i in range(ia, ib)] + [(j, ib) for j in range(ja, jb)] + \
           [(jb, i) for i in range(ib, ia, -1)] + [(j, ia) for j in range(jb, ja, -1)]
    ring_ids = np.array([gid[j, i] for (j, i) in ring], np.int64)
    S = np.array([[xs[i], ys[j]] for (j, i) in ring])
    ctr = np.array([cx, cy])
    dvec = S - ctr
    dist = np.linalg.norm(dvec, axis=1)
    unit = dvec / dist[:, None]
    Cp = ctr + R * unit
    nring = ring_ids.size
    if first_layer is None:
        first_layer = 2 * math.pi * R / nring
    if radial_layers is None:
        radial_layers = max(2, int(round((box - R) / h)) + 8)
    # per-ray grading (rays have different lengths): same layer count, geometric from first_layer
    t = np.zeros((nring, radial_layers + 1))
    for k in range(nring):
        lk = dist[k] - R
        segs = _geometric(radial_layers, first_layer, lk)
        t[k, 1:] = np.cumsum(segs) / lk
    t[:, -1] = 1.0
    nbg = xy_bg.shape[0]
    # O-grid nodes of layers 0..radial_layers-1 (layer radial_layers is the square boundary)
    P = Cp[:, None, :] + t[:, :radial_layers, None] * (S - Cp)[:, None, :]
    oid = nbg + np.arange(nring * radial_layers).reshape(nring, radial_layers)
    xy = np.concatenate([xy_bg, P.reshape(-1, 2)], axis=0)
    node = np.concatenate([oid, ring_ids[:, None]], axis=1)   # (nring, radial_layers + 1)
    k0 = np.arange(nring)
    k1 = (k0 + 1) % nring
    o00 = node[k0][:, :-1].ravel()
    o10 = node[k1][:, :-1].ravel()
    o01 = node[k0][:, 1:].ravel()
    o11 = node[k1][:, 1:].ravel()
    t_bg = np.concatenate([np.stack([p00, p10, p11], 1), np.stack([p00, p11, p01], 1)], axis=0)
    # O-grid quads (tangential k -> k+1, outward j -> j+1) split along the shorter diagonal,
    # counter-clockwise
    d1 = np.linalg.norm(xy[o00] - xy[o11], axis=1)
    d2 = np.linalg.norm(xy[o10] - xy[o01], axis=1)
    s = d1 <= d2
    ta = np.where(s[:, None], np.stack([o00, o01, o11], 1), np.stack([o00, o01, o10], 1))
    tb = np.where(s[:, None], np.stack([o00, o11, o10], 1), np.stack([o10, o01, o11], 1))
    t_og = np.concatenate([ta, tb], axis=0)
    tri = np.concatenate([t_bg, t_og], axis=0).astype(np.int32)
    info = dict(cylinder_nodes=oid[:, 0].copy(), R=R, center=(cx, cy), L=L, H=H, D=D, h=h,
                radial_layers=radial_layers, ring=nring)
    return xy.astype(np.float64), tri, info


def box3d(n=128):
    """Unit cube, n^3 cubes, each split into 6 Kuhn tetrahedra sharing the main diagonal;
    node (i, j, k) -> (k*(n+1)+j)*(n+1)+i."""
    g = np.arange(n + 1, dtype=np.float64) / n
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = ((k * (n + 1) + j) * (n + 1) + i).ravel()
    dx, dy, dz = 1, n + 1, (n + 1) * (n + 1)
    v = {}
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                v[(a, b, c)] = base + a * dx + b * dy + c * dz
    # the 6 monotone paths from (0,0,0) to (1,1,1)
    paths = [((1, 0, 0), (1, 1, 0)), ((1, 0, 0), (1, 0, 1)), ((0, 1, 0), (1, 1, 0)),
             ((0, 1, 0), (0, 1, 1)), ((0, 0, 1), (1, 0, 1)), ((0, 0, 1), (0, 1, 1))]
    tets = [np.stack([v[(0, 0, 0)], v[p1], v[p2], v[(1, 1, 1)]], 1) for p1, p2 in paths]
    tet = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    return xyz, tet


def write_dat(path, xy, cells):
    """SALOME DAT: 'nV nE', vertex lines 'id x y z', element lines 'id type n1..nk' (1-based)."""
    dim3 = xy.shape[1] == 3
    typ = 304 if cells.shape[1] == 4 else 203
    lines = [f"{xy.shape[0]} {cells.shape[0]}"]
    for k, p in enumerate(xy):
        z = p[2] if dim3 else 0.0
        lines.append(f"{k + 1} {p[0]:.17e} {p[1]:.17e} {z:.17e}")
    for k, c in enumerate(cells):
        lines.append(f"{k + 1} {typ} " + " ".join(str(int(v) + 1) for v in c))
    with open(path, "w") as f:
        f.write("\n".join(lines))
