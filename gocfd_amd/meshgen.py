"""Deterministic synthetic meshes for the benchmark and test workloads.

jittered_quads(): an nx x ny grid of quads on [x0,x1] x [y0,y1]; interior
nodes are moved by up to +-jitter*h in each coordinate with a SplitMix64
stream from `seed`; every quad is split on its (v00, v11) diagonal into two
counter-clockwise triangles. Boundary markers follow the reference's SU2
conventions (readfiles/readSU2Grid.go:27-52): "far" on y=y0 and y=y1, and one
tag "periodic-1" given as two contiguous chains (x=x0 and x=x1) listed in the
same (upward) direction, so that Curve.ReOrder and the pairing of
DG2D/triangulation.go:63-99 are deterministic.
"""
from __future__ import annotations

import numpy as np

_MASK = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n uniform doubles in [0,1) from the SplitMix64 sequence of `seed`."""
    out = np.empty(n)
    s = seed & _MASK
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & _MASK
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        z ^= z >> 31
        out[i] = (z >> 11) * (1.0 / (1 << 53))
    return out


def jittered_quads(nx: int, ny: int, x0=0.0, x1=10.0, y0=-5.0, y1=5.0, jitter=0.1, seed=1,
                   periodic_x=True, periodic_y=False):
    """Returns (VX, VY, EToV [K,3], tags) with K = 2*nx*ny."""
    hx, hy = (x1 - x0) / nx, (y1 - y0) / ny
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    X = x0 + ii * hx
    Y = y0 + jj * hy
    interior = (ii > 0) & (ii < nx) & (jj > 0) & (jj < ny)
    nint = int(interior.sum())
    u = splitmix64(seed, 2 * nint)
    X = X.astype(np.float64)
    Y = Y.astype(np.float64)
    X[interior] += (2 * u[:nint] - 1) * jitter * hx
    Y[interior] += (2 * u[nint:] - 1) * jitter * hy
    VX, VY = X.ravel(), Y.ravel()

    def v(i, j):
        return j * (nx + 1) + i

    tris = []
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = v(i, j), v(i + 1, j), v(i + 1, j + 1), v(i, j + 1)
            tris.append((a, b, c))
            tris.append((a, c, d))
    EToV = np.array(tris, dtype=np.int32)
    bottom = [(v(i, 0), v(i + 1, 0)) for i in range(nx)]
    top = [(v(i, ny), v(i + 1, ny)) for i in range(nx)]
    far = [] if periodic_y else bottom + top
    tags = {"far": far} if far else {}
    if periodic_y:
        tags["periodic-2"] = bottom + top
    left = [(v(0, j), v(0, j + 1)) for j in range(ny)]
    right = [(v(nx, j), v(nx, j + 1)) for j in range(ny)]
    if periodic_x:
        tags["periodic-1"] = left + right
    else:
        tags["far"] = far + left + right
    return VX, VY, EToV, tags


def periodic_mesh(path: str, n: int, seed: int = 1, jitter: float = 0.1) -> str:
    """n x n jittered quads on [0,10]x[-5,5], periodic in x and y (no boundary)."""
    VX, VY, EToV, tags = jittered_quads(n, n, seed=seed, jitter=jitter, periodic_y=True)
    write_su2(path, VX, VY, EToV, tags)
    return path


def write_su2(path: str, VX, VY, EToV, tags, split_periodic=True) -> None:
    """SU2 text mesh (NDIME, NELEM type 5, NPOIN, NMARK type 3)."""
    lines = ["NDIME= 2", f"NELEM= {len(EToV)}"]
    lines += [f"5 {a} {b} {c} {k}" for k, (a, b, c) in enumerate(EToV)]
    lines.append(f"NPOIN= {len(VX)}")
    lines += [f"{float(x)!r} {float(y)!r} {i}" for i, (x, y) in enumerate(zip(VX, VY))]
    markers = []
    for tag, edges in tags.items():
        if split_periodic and tag.startswith("periodic"):
            h = len(edges) // 2
            markers.append((tag, edges[:h]))
            markers.append((tag, edges[h:]))
        else:
            markers.append((tag, edges))
    lines.append(f"NMARK= {len(markers)}")
    for tag, edges in markers:
        lines.append(f"MARKER_TAG= {tag}")
        lines.append(f"MARKER_ELEMS= {len(edges)}")
        lines += [f"3 {a} {b}" for a, b in edges]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def vortex_mesh(path: str, n: int, seed: int = 1, jitter: float = 0.1) -> str:
    """The isentropic-vortex mesh on [0,10]x[-5,5]: n x n jittered quads -> 2n^2 triangles."""
    VX, VY, EToV, tags = jittered_quads(n, n, seed=seed, jitter=jitter)
    write_su2(path, VX, VY, EToV, tags)
    return path


def shocktube_mesh(path: str, nx: int, ny: int = 2, height: float = 0.1, jitter: float = 0.0,
                   seed: int = 1) -> str:
    """Sod shock-tube channel on [0,1] x [0,height]: nx x ny quads split into
    2 nx ny triangles, "in" on x=0, "out" on x=1 and "wall" on y=0 and
    y=height (the marker names readfiles/readSU2Grid.go maps to BC_In,
    BC_Out, BC_Wall). With an even nx and no jitter the initial jump at
    x = 0.5 lies on element edges."""
    VX, VY, EToV, _ = jittered_quads(nx, ny, x0=0.0, x1=1.0, y0=0.0, y1=height, jitter=jitter, seed=seed)

    def v(i, j):
        return j * (nx + 1) + i

    tags = {
        "in": [(v(0, j), v(0, j + 1)) for j in range(ny)],
        "out": [(v(nx, j), v(nx, j + 1)) for j in range(ny)],
        "wall": [(v(i, 0), v(i + 1, 0)) for i in range(nx)] + [(v(i, ny), v(i + 1, ny)) for i in range(nx)],
    }
    write_su2(path, VX, VY, EToV, tags)
    return path


def naca4_thickness(x, t=0.12):
    """Half-thickness of the symmetric 4-digit section with a closed trailing
    edge: y = 5 t (0.2969 sqrt x - 0.126 x - 0.3516 x^2 + 0.2843 x^3 - 0.1036 x^4)."""
    return 5.0 * t * (0.2969 * np.sqrt(x) - 0.126 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 - 0.1036 * x ** 4)


def airfoil_mesh(path: str, n_around: int = 400, n_radial: int = 125, radius: float = 20.0, seed: int = 1,
                 jitter: float = 0.1, stretch: float = 8.0, points_only: bool = False):
    """Symmetric 4-digit airfoil (chord 1, 12 % thick, closed trailing edge)
    in an O-domain of the given radius about (0.5, 0): n_around surface
    points (cosine-clustered toward the leading edge) by n_radial
    layers, geometrically stretched toward the wall (g = (e^(s u) - 1)/(e^s - 1));
    layer points leave the wall along its normal and bend toward the circle
    point at the surface point's angle about the centre,
    P = S + n R g (1 - g) + (C - S) g^2. Interior nodes move by up to
    +-jitter of the local spacing (SplitMix64 stream from `seed`); each quad is
    split into two counter-clockwise triangles: 2 n_around n_radial
    triangles (100,000 at the defaults, 50,000 points). Tags: "wall" on the
    airfoil, "far" on the circle. points_only: return the points (VX, VY)
    ordered layer by layer (n_around per layer, the wall first) instead."""
    assert n_around % 2 == 0
    m = n_around // 2
    # clustered toward the leading edge only: a cusp-side clustering would
    # pinch the cells between the two rays that leave the trailing edge
    xs = 1.0 - np.cos(0.5 * np.pi * np.linspace(0.0, 1.0, m + 1))  # 0 (LE) .. 1 (TE)
    up_x = xs[::-1][:-1]                       # TE -> LE, without the LE point
    lo_x = xs[:-1]                             # LE -> TE, without the TE point
    sx = np.concatenate([up_x, lo_x])
    sy = np.concatenate([naca4_thickness(up_x), -naca4_thickness(lo_x)])
    sy[0] = 0.0                                # the trailing-edge point
    theta = np.arctan2(sy, sx - 0.5)
    theta[m:] = np.where(theta[m:] <= 0.0, theta[m:] + 2 * np.pi, theta[m:])  # monotone around the body
    cx, cy = 0.5 + radius * np.cos(theta), radius * np.sin(theta)
    # outward wall normals (the surface runs counter-clockwise; at the
    # trailing-edge cusp the central difference gives the bisector)
    tx = np.roll(sx, -1) - np.roll(sx, 1)
    ty = np.roll(sy, -1) - np.roll(sy, 1)
    tn = np.hypot(tx, ty)
    nx, ny = ty / tn, -tx / tn
    u = np.linspace(0.0, 1.0, n_radial + 1)
    g = (np.expm1(stretch * u) / np.expm1(stretch))[:, None]
    # near the wall the layers follow the normals, far out the circle:
    # P = S + n R g (1 - g) + (C - S) g^2
    X = sx[None, :] + nx[None, :] * radius * g * (1 - g) + (cx - sx)[None, :] * g ** 2   # [n_radial+1][n_around]
    Y = sy[None, :] + ny[None, :] * radius * g * (1 - g) + (cy - sy)[None, :] * g ** 2
    g = g[:, 0]
    # jitter of the interior layers, scaled by the local radial and tangential spacing
    dr = np.hypot(np.diff(X, axis=0), np.diff(Y, axis=0))
    ds = np.hypot(np.roll(X, -1, axis=1) - X, np.roll(Y, -1, axis=1) - Y)
    h = np.minimum(dr[1:], np.minimum(ds[1:-1], np.roll(ds[1:-1], 1, axis=1)))
    r = splitmix64(seed, 2 * h.size).reshape(2, *h.shape)
    X[1:-1] += (2 * r[0] - 1) * jitter * h
    Y[1:-1] += (2 * r[1] - 1) * jitter * h
    VX, VY = X.ravel(), Y.ravel()
    if points_only:
        return VX, VY

    def v(j, i):
        return j * n_around + (i % n_around)

    tris = []
    for j in range(n_radial):
        for i in range(n_around):
            a, b, c, d = v(j, i), v(j, i + 1), v(j + 1, i + 1), v(j + 1, i)
            tris.append((a, b, c))
            tris.append((a, c, d))
    EToV = np.array(tris, dtype=np.int64)
    p0, p1, p2 = (np.stack([VX[EToV[:, k]], VY[EToV[:, k]]], axis=1) for k in range(3))
    cross = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0])
    EToV[cross < 0] = EToV[cross < 0][:, [0, 2, 1]]
    tags = {"wall": [(v(0, i), v(0, i + 1)) for i in range(n_around)],
            "far": [(v(n_radial, i), v(n_radial, i + 1)) for i in range(n_around)]}
    write_su2(path, VX, VY, EToV.astype(np.int32), tags)
    return path


def wedge_mesh(path: str, n_points: int = 125_000, seed: int = 1, x_ramp: float = 0.5,
               angle_deg: float = 15.0, jitter: float = 0.2) -> str:
    """Mach-2 wedge channel: [0,3] x [0,1] with a ramp rising at angle_deg from
    x = x_ramp to the outlet. Delaunay triangulation (scipy) of about n_points
    interior points -- a hexagonal lattice of spacing h, each point moved by
    up to +-jitter*h per coordinate with a SplitMix64 stream from `seed` --
    plus boundary points at spacing h; interior points closer than h/2 to the
    boundary are dropped. (Purely uniform random points give sliver
    triangles whose CFL limit stalls the time step.) Tags: "wall" on the
    floor, ramp and roof, "far" on the inlet and outlet (the reference's
    In/Out need FSIn/FSOut, which only the shock-tube case sets). About
    2 n_points triangles."""
    from scipy.spatial import Delaunay
    t = np.tan(np.radians(angle_deg))
    L, H = 3.0, 1.0
    area = L * H - 0.5 * (L - x_ramp) ** 2 * t
    h = np.sqrt(2.0 * area / (np.sqrt(3.0) * n_points))
    corners = [(0.0, 0.0), (x_ramp, 0.0), (L, (L - x_ramp) * t), (L, H), (0.0, H)]
    bpts, segs = [], []
    for a, b in zip(corners, corners[1:] + corners[:1]):
        a, b = np.array(a), np.array(b)
        m = max(1, int(np.ceil(np.linalg.norm(b - a) / h)))
        for i in range(m):
            bpts.append(a + (b - a) * (i / m))
        segs.append(m)
    bpts = np.array(bpts)
    dy = h * np.sqrt(3.0) / 2
    ny, nx = int(H / dy) + 2, int(L / h) + 2
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    x = (ii + 0.5 * (jj % 2)).ravel() * h
    y = jj.ravel() * dy
    u = splitmix64(seed, 2 * x.size)
    x = x + (2 * u[:x.size] - 1) * jitter * h
    y = y + (2 * u[x.size:] - 1) * jitter * h
    c, sn = np.cos(np.radians(angle_deg)), np.sin(np.radians(angle_deg))
    d_ramp = np.where(x >= x_ramp, (y - (x - x_ramp) * t) * c, np.inf)
    d_floor = np.where(x < x_ramp, y, np.hypot(x - x_ramp, y))
    dist = np.minimum.reduce([d_floor, d_ramp, H - y, x, L - x])
    inside = (y > np.where(x < x_ramp, 0.0, (x - x_ramp) * t)) & (y < H) & (x > 0) & (x < L)
    pts = np.stack([x, y], axis=1)[inside & (dist > 0.5 * h)]
    P = np.vstack([bpts, pts])
    tri = Delaunay(P).simplices.astype(np.int64)
    a, b, cc = P[tri[:, 0]], P[tri[:, 1]], P[tri[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (cc[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (cc[:, 0] - a[:, 0])
    tri = tri[np.abs(cross) > 1e-12 * h * h]
    cross = cross[np.abs(cross) > 1e-12 * h * h]
    tri[cross < 0] = tri[cross < 0][:, [0, 2, 1]]
    # boundary edges, walked along the boundary point list
    nb = len(bpts)
    start = np.cumsum([0] + segs)
    walls, far = [], []
    for s_i, (s0, m) in enumerate(zip(start[:-1], segs)):
        edges = [(s0 + i, (s0 + i + 1) % nb) for i in range(m)]
        # segments: floor, ramp, outlet, roof, inlet
        (far if s_i in (2, 4) else walls).extend(edges)
    tags = {"wall": walls, "far": far}
    write_su2(path, P[:, 0], P[:, 1], tri.astype(np.int32), tags)
    return path


def airfoil_delaunay_mesh(path: str, n_around: int = 400, n_radial: int = 125, radius: float = 20.0,
                          seed: int = 1, jitter: float = 0.1, stretch: float = 8.0) -> str:
    """The airfoil workload as a Delaunay mesh: the Delaunay triangulation
    (scipy) of the seeded, wall-clustered point set of airfoil_mesh
    (n_around x (n_radial + 1) points: 50,400 at the defaults), without the
    triangles inside the airfoil. The points form a jittered lattice, as for
    the wedge channel: independent uniform points would leave sliver
    triangles whose CFL limit stalls the time step. Every airfoil surface
    segment and every circle segment must be an edge of the triangulation
    (checked). Tags: "wall" on the airfoil, "far" on the circle."""
    from scipy.spatial import Delaunay
    VX, VY = airfoil_mesh(path, n_around, n_radial, radius, seed, jitter, stretch, points_only=True)
    P = np.stack([VX, VY], axis=1)
    tri = Delaunay(P).simplices.astype(np.int64)
    # drop triangles inside the airfoil (all three vertices on the surface
    # and the centroid within the section)
    on_wall = np.arange(len(VX)) < n_around
    c = P[tri].mean(axis=1)
    xc = np.clip(c[:, 0], 0.0, 1.0)
    inside = on_wall[tri].all(axis=1) & (c[:, 0] > 0) & (c[:, 0] < 1) & (np.abs(c[:, 1]) < naca4_thickness(xc))
    tri = tri[~inside]
    a, b, cc = P[tri[:, 0]], P[tri[:, 1]], P[tri[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (cc[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (cc[:, 0] - a[:, 0])
    keep = np.abs(cross) > 0
    tri, cross = tri[keep], cross[keep]
    tri[cross < 0] = tri[cross < 0][:, [0, 2, 1]]
    wall = [(i, (i + 1) % n_around) for i in range(n_around)]
    o = n_radial * n_around
    far = [(o + i, o + (i + 1) % n_around) for i in range(n_around)]
    edges = set()
    for t in tri:
        for k in range(3):
            u, w = int(t[k]), int(t[(k + 1) % 3])
            edges.add((min(u, w), max(u, w)))
    missing = [e for e in wall + far if (min(e), max(e)) not in edges]
    if missing:
        raise RuntimeError(f"{len(missing)} boundary segments are not Delaunay edges")
    write_su2(path, VX, VY, tri.astype(np.int32), {"wall": wall, "far": far})
    return path
