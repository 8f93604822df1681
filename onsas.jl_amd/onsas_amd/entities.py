"""Mesh, element sets, face sets and cross sections.

The reference keeps one Julia object per node / element / face
(src/Entities/Nodes.jl:74-84, Tetrahedrons.jl:24-48, Trusses.jl:54-99,
TriangularFaces.jl, src/Meshes/Meshes.jl:158-184).  A GPU engine wants
connectivity arrays, so an element *set* here is a block of elements of one
type: the same node order, the same dof numbering (node i owns dofs 3i..3i+2,
set_dofs!, Meshes.jl:85-98) and the same geometry functions, vectorised.
Node ids are 0-based.
"""
from __future__ import annotations

import numpy as np


class RotatedEngineeringStrain:
    """eps = (l^2 - l0^2) / (l0 (l0 + l))  (src/Entities/Trusses.jl:37, 187-189)."""
    code = 0


class GreenStrain:
    """eps = (l^2 - l0^2) / (2 l0^2)  (src/Entities/Trusses.jl:44, 192-194)."""
    code = 1


DEFAULT_STRAIN_MODEL = RotatedEngineeringStrain


class Circle:
    """Circular cross section of diameter d (src/CrossSections/Circles.jl:14-20)."""

    def __init__(self, d):
        self.d = d

    def area(self):
        return np.pi * np.asarray(self.d) ** 2 / 4


class Square:
    """Square cross section of side a (src/CrossSections/Squares.jl)."""

    def __init__(self, a):
        self.a = a

    def area(self):
        return np.asarray(self.a) ** 2


class GenericCrossSection:
    """Cross section given by its area."""

    def __init__(self, A):
        self.A = A

    def area(self):
        return np.asarray(self.A, dtype=np.float64)


def area(x):
    return x.area()


class _Set:
    nodes_per_entity = 0

    def __init__(self, conn, label=None):
        conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, self.nodes_per_entity)
        self.conn = conn
        self.label = "no_label" if label is None else str(label)

    def __len__(self):
        return self.conn.shape[0]

    def nodes(self):
        return np.unique(self.conn)


class Tetrahedra(_Set):
    """A set of 4-node Tetrahedron elements (src/Entities/Tetrahedrons.jl:24-48).
    Node order must give det[X1-X2, X4-X2, X3-X2] > 0."""
    nodes_per_entity = 4
    kind = "tet"

    def volumes(self, coords):
        X = coords[self.conn]
        J = np.stack([X[:, 0] - X[:, 1], X[:, 3] - X[:, 1], X[:, 2] - X[:, 1]], axis=2)
        return np.linalg.det(J) / 6.0  # _volume (Tetrahedrons.jl:134-138)


class Trusses(_Set):
    """A set of 2-node Truss elements (src/Entities/Trusses.jl:54-99)."""
    nodes_per_entity = 2
    kind = "truss"

    def __init__(self, conn, cross_section, strain_model=DEFAULT_STRAIN_MODEL, label=None):
        super().__init__(conn, label)
        self.cross_section = cross_section
        self.strain_model = strain_model

    def areas(self):
        a = self.cross_section.area() if hasattr(self.cross_section, "area") else self.cross_section
        return np.broadcast_to(np.asarray(a, dtype=np.float64), (len(self),)).copy()

    def lengths(self, coords):
        d = coords[self.conn[:, 1]] - coords[self.conn[:, 0]]
        return np.linalg.norm(d, axis=1)

    def volumes(self, coords):
        return self.areas() * self.lengths(coords)  # volume(::Truss) (Trusses.jl:102-107)


class TriangularFaces(_Set):
    """A set of TriangularFace faces (src/Entities/TriangularFaces.jl)."""
    nodes_per_entity = 3
    kind = "face"

    def area_vectors(self, coords):
        x1, x2, x3 = (coords[self.conn[:, a]] for a in range(3))
        return 0.5 * np.cross(x2 - x1, x3 - x1)  # _area_vec (TriangularFaces.jl:44-47)

    def areas(self, coords):
        return np.linalg.norm(self.area_vectors(coords), axis=1)

    def normal_directions(self, coords):
        av = self.area_vectors(coords)
        return av / np.linalg.norm(av, axis=1, keepdims=True)


class NodeSet:
    """A set of nodes (0-based ids)."""

    def __init__(self, ids, label=None):
        self.ids = np.unique(np.asarray(ids, dtype=np.int64).reshape(-1))
        self.label = "no_label" if label is None else str(label)

    def __len__(self):
        return self.ids.size

    def nodes(self):
        return self.ids


class Mesh:
    """Nodes plus element and face sets (src/Meshes/Meshes.jl:158-184)."""

    def __init__(self, nodes, elements=(), faces=()):
        self.coords = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64).reshape(-1, 3))
        self.elements = list(elements)
        self.faces = list(faces)

    @property
    def num_nodes(self):
        return self.coords.shape[0]

    def num_dofs(self):
        return 3 * self.num_nodes

    @classmethod
    def from_structured(cls, sm):
        """Mesh of a meshes.StructuredMesh: one element set, its named face sets."""
        elements = []
        if sm.tets is not None and len(sm.tets):
            elements.append(Tetrahedra(sm.tets, "tets"))
        faces = [TriangularFaces(f, name) for name, f in sm.faces.items()]
        return cls(sm.coords, elements, faces)

    def face_set(self, label):
        for f in self.faces:
            if f.label == label:
                return f
        raise KeyError(label)
