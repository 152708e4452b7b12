"""Deterministic synthetic meshes for the isentropic-vortex workloads.

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
