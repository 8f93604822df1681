"""Structured mesh generators for the synthetic workloads.

The reference builds its meshes with Gmsh at run time
(examples/uniaxial_extension/uniaxial_mesh.jl:22-93), which is not available
offline.  These generators replicate the reference's own manufactured hexahedron
split (examples/uniaxial_extension/uniaxial_extension.jl:45-73) over a
structured grid; every hexahedron is cut into the same six tetrahedra around its
(x0, y1, z0)-(x1, y0, z1) diagonal, which makes neighbouring faces conform.

All arrays are numpy: coordinates ``(nn, 3) float64``, connectivities
``(ne, k) int32`` with 0-based node ids, node order as in the reference (every
tetrahedron has positive volume, every boundary face normal points out of the
solid).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Hexahedron corner numbering of uniaxial_extension.jl:45-53, as (dx, dy, dz) offsets:
#   n1 (0,0,0) n2 (0,0,1) n3 (0,1,1) n4 (0,1,0) n5 (1,0,0) n6 (1,0,1) n7 (1,1,1) n8 (1,1,0)
_HEX_CORNERS = np.array([(0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0),
                         (1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)], dtype=np.int64)
# The six tetrahedra of uniaxial_extension.jl:66-71 (1-based corner ids there).
_HEX_TETS = np.array([(1, 4, 2, 6), (6, 2, 3, 4), (4, 3, 6, 7),
                      (4, 1, 5, 6), (4, 6, 5, 8), (4, 7, 6, 8)], dtype=np.int64) - 1
# Local faces of a tetrahedron (node triples), any orientation; fixed afterwards.
_TET_FACES = np.array([(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)], dtype=np.int64)


def splitmix64(x):
    """splitmix64 finaliser on a uint64 array (Steele, Lea & Flood 2014)."""
    z = np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform_pm1(seed, ids):
    """Deterministic uniform numbers in [-1, 1) keyed by (seed, integer id)."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) << np.uint64(40)) + np.asarray(ids, dtype=np.uint64)
        h = splitmix64(key)
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0


@dataclass
class StructuredMesh:
    """A structured tetrahedral (or truss) mesh with named node and face sets."""

    coords: np.ndarray                      # (nn, 3) float64
    tets: np.ndarray                        # (ne, 4) int32
    trusses: np.ndarray = None              # (nb, 2) int32
    faces: dict = field(default_factory=dict)       # name -> (nf, 3) int32 outward
    node_sets: dict = field(default_factory=dict)   # name -> (k,) int64
    shape: tuple = ()

    @property
    def nn(self):
        return self.coords.shape[0]


def _grid_ids(nx, ny, nz):
    """node id of grid point (i, j, k), x fastest."""
    return lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)


def _hex_tets(nx, ny, nz, node_id, periodic_y=0):
    """Six tets per hex cell of an (nx, ny, nz) cell grid.  With periodic_y > 0 the
    y index wraps modulo periodic_y (used for the cylinder's theta direction)."""
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).reshape(-1) for a in (i, j, k))  # x fastest cells
    corners = []
    for dx, dy, dz in _HEX_CORNERS:
        jj = j + dy
        if periodic_y:
            jj = jj % periodic_y
        corners.append(node_id(i + dx, jj, k + dz))
    corners = np.stack(corners, axis=1)               # (ncell, 8)
    tets = corners[:, _HEX_TETS]                      # (ncell, 6, 4)
    return tets.reshape(-1, 4).astype(np.int32)


def _orient_outward(coords, tri, outward):
    """Flip triangles whose area vector does not point along `outward` (per face)."""
    x1, x2, x3 = (coords[tri[:, a]] for a in range(3))
    n = np.cross(x2 - x1, x3 - x1)
    flip = np.einsum("ij,ij->i", n, outward) < 0
    out = tri.copy()
    out[flip, 1], out[flip, 2] = tri[flip, 2], tri[flip, 1]
    return out


def _boundary_faces(tets, on_plane):
    """Tet faces whose three nodes all satisfy `on_plane` (bool per node)."""
    fs = tets[:, _TET_FACES]                          # (ne, 4, 3)
    sel = on_plane[fs].all(axis=2)
    return fs[sel].reshape(-1, 3)


def cube(n, L=1.0, jitter=0.0, seed=0, ny=None, nz=None, Ly=None, Lz=None):
    """Structured box of n x ny x nz hexahedra, each split into the reference's six
    tetrahedra (uniaxial_extension.jl:66-71).

    jitter: interior nodes move by jitter * h * U(-1, 1) per component
    (splitmix64 keyed by seed and node id); boundary nodes stay put, so affine
    fields remain exact on the perturbed mesh.

    Face sets (outward): x0, x1, y0, y1, z0, z1.  Node sets: x0 ... z1, all.
    """
    nx = n
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    Lx = L
    Ly = L if Ly is None else Ly
    Lz = L if Lz is None else Lz
    nid = _grid_ids(nx, ny, nz)
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    i, j, k = i.reshape(-1), j.reshape(-1), k.reshape(-1)
    coords = np.stack([i * (Lx / nx), j * (Ly / ny), k * (Lz / nz)], axis=1).astype(np.float64)
    on = {"x0": i == 0, "x1": i == nx, "y0": j == 0, "y1": j == ny, "z0": k == 0, "z1": k == nz}
    if jitter:
        interior = ~(on["x0"] | on["x1"] | on["y0"] | on["y1"] | on["z0"] | on["z1"])
        ids = np.nonzero(interior)[0].astype(np.uint64)
        h = np.array([Lx / nx, Ly / ny, Lz / nz])
        for d in range(3):
            coords[interior, d] += jitter * h[d] * uniform_pm1(seed, 3 * ids + np.uint64(d))
    tets = _hex_tets(nx, ny, nz, nid)
    normals = {"x0": (-1, 0, 0), "x1": (1, 0, 0), "y0": (0, -1, 0), "y1": (0, 1, 0),
               "z0": (0, 0, -1), "z1": (0, 0, 1)}
    faces = {}
    for name, mask in on.items():
        tri = _boundary_faces(tets, mask)
        faces[name] = _orient_outward(coords, tri, np.broadcast_to(normals[name], (len(tri), 3)))
    node_sets = {name: np.nonzero(mask)[0].astype(np.int64) for name, mask in on.items()}
    node_sets["all"] = np.arange(coords.shape[0], dtype=np.int64)
    return StructuredMesh(coords, tets, None, faces, node_sets, (nx, ny, nz))


def cylinder(nr, ntheta, nz, Ri, Re, L):
    """Thick-walled cylinder structured in (r, theta, z) with theta periodic; the
    same six-tet split is applied in index space (r -> x, theta -> y, z -> z, a
    right-handed map, so volumes stay positive).

    Face sets: inner (outward normal of the solid, i.e. pointing into the hole),
    outer, z0, z1.  Node sets: z0, z1, theta0 (theta = 0 line), theta90, inner,
    outer.
    """
    nt = ntheta
    nid = lambda i, j, k: i + (nr + 1) * (j + nt * k)
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(nt), np.arange(nr + 1), indexing="ij")
    i, j, k = i.reshape(-1), j.reshape(-1), k.reshape(-1)
    r = Ri + (Re - Ri) * i / nr
    th = 2 * np.pi * j / nt
    coords = np.stack([r * np.cos(th), r * np.sin(th), L * k / nz], axis=1)
    tets = _hex_tets(nr, nt, nz, nid, periodic_y=nt)
    on = {"inner": i == 0, "outer": i == nr, "z0": k == 0, "z1": k == nz}
    faces = {}
    for name, mask in on.items():
        tri = _boundary_faces(tets, mask)
        cen = coords[tri].mean(axis=1)
        if name in ("inner", "outer"):
            rad = cen.copy()
            rad[:, 2] = 0
            rad /= np.linalg.norm(rad, axis=1, keepdims=True)
            out = -rad if name == "inner" else rad
        else:
            out = np.broadcast_to((0, 0, -1 if name == "z0" else 1), (len(tri), 3))
        faces[name] = _orient_outward(coords, tri, out)
    node_sets = {name: np.nonzero(mask)[0].astype(np.int64) for name, mask in on.items()}
    node_sets["theta0"] = np.nonzero(j == 0)[0].astype(np.int64)
    if nt % 4 == 0:
        node_sets["theta90"] = np.nonzero(j == nt // 4)[0].astype(np.int64)
    node_sets["all"] = np.arange(coords.shape[0], dtype=np.int64)
    return StructuredMesh(coords, tets, None, faces, node_sets, (nr, nt, nz))


def truss_lattice(n, h=0.1, jitter=0.0, seed=1):
    """Space-truss lattice on an (n+1)^3 node grid: bars along every cell edge, one
    diagonal per cell face (the same diagonal family as the six-tet split), and one
    body diagonal per cell.

    jitter: every node moves by jitter * h * U(-1, 1) per component except the
    component normal to a boundary plane the node lies on.
    Node sets: x0 ... z1, all.
    """
    nid = _grid_ids(n, n, n)
    k, j, i = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    i, j, k = i.reshape(-1), j.reshape(-1), k.reshape(-1)
    coords = np.stack([i * h, j * h, k * h], axis=1).astype(np.float64)
    on = {"x0": i == 0, "x1": i == n, "y0": j == 0, "y1": j == n, "z0": k == 0, "z1": k == n}
    if jitter:
        ids = np.arange(coords.shape[0], dtype=np.uint64)
        for d, (lo, hi) in enumerate((("x0", "x1"), ("y0", "y1"), ("z0", "z1"))):
            free = ~(on[lo] | on[hi])
            coords[free, d] += jitter * h * uniform_pm1(seed, 3 * ids[free] + np.uint64(d))
    bars = []
    r = np.arange(n + 1)
    c = np.arange(n)
    # axis edges
    a, b, e = np.meshgrid(c, r, r, indexing="ij")
    bars.append((nid(a, b, e), nid(a + 1, b, e)))
    a, b, e = np.meshgrid(r, c, r, indexing="ij")
    bars.append((nid(a, b, e), nid(a, b + 1, e)))
    a, b, e = np.meshgrid(r, r, c, indexing="ij")
    bars.append((nid(a, b, e), nid(a, b, e + 1)))
    # face diagonals (uniaxial_extension.jl faces: x-normal n2-n4, y-normal n1-n6, z-normal n4-n5)
    a, b, e = np.meshgrid(r, c, c, indexing="ij")
    bars.append((nid(a, b, e + 1), nid(a, b + 1, e)))
    a, b, e = np.meshgrid(c, r, c, indexing="ij")
    bars.append((nid(a, b, e), nid(a + 1, b, e + 1)))
    a, b, e = np.meshgrid(c, c, r, indexing="ij")
    bars.append((nid(a, b + 1, e), nid(a + 1, b, e)))
    # body diagonal n4-n6
    a, b, e = np.meshgrid(c, c, c, indexing="ij")
    bars.append((nid(a, b + 1, e), nid(a + 1, b, e + 1)))
    tr = np.concatenate([np.stack([p.reshape(-1), q.reshape(-1)], axis=1) for p, q in bars])
    node_sets = {name: np.nonzero(mask)[0].astype(np.int64) for name, mask in on.items()}
    node_sets["all"] = np.arange(coords.shape[0], dtype=np.int64)
    return StructuredMesh(coords, np.zeros((0, 4), np.int32), tr.astype(np.int32), {},
                          node_sets, (n, n, n))


def tet_volumes(coords, tets):
    """Reference-configuration volumes, det[X1-X2, X4-X2, X3-X2] / 6
    (src/Entities/Tetrahedrons.jl:129-138)."""
    X = coords[tets]
    J = np.stack([X[:, 0] - X[:, 1], X[:, 3] - X[:, 1], X[:, 2] - X[:, 1]], axis=2)
    return np.linalg.det(J) / 6.0
