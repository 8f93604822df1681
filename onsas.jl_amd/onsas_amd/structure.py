"""Structural model: materials assigned to element sets, boundary conditions,
free dofs.

StructuralMaterial -- src/StructuralModel/StructuralMaterials.jl:25-35
Structure          -- src/StructuralModel/Structures.jl:129-142
"""
from __future__ import annotations

import numpy as np

from .boundary_conditions import StructuralBoundaryCondition
from .entities import Mesh


class StructuralMaterial:
    """Materials mapped to the element sets they apply to:
    StructuralMaterial({material: element_set, ...}) or from pairs."""

    def __init__(self, pairs):
        items = list(pairs.items()) if isinstance(pairs, dict) else list(pairs)
        self._pairs = []
        for m, es in items:
            for e in (es if isinstance(es, (list, tuple)) else [es]):
                self._pairs.append((m, e))

    def __iter__(self):
        return iter(self._pairs)

    def __len__(self):
        return len(self._pairs)

    def __getitem__(self, label):
        for m, _ in self._pairs:
            if m.label == str(label):
                return m
        raise KeyError(label)

    def replace(self, old, new):
        """replace! of a material (StructuralMaterials.jl)."""
        self._pairs = [(new if m is old else m, e) for m, e in self._pairs]


class Structure:
    """Mesh + materials + boundary conditions; free dofs are every node dof minus
    the fixed ones, in ascending order (Structures.jl:129-142)."""

    def __init__(self, mesh: Mesh, materials: StructuralMaterial,
                 bcs: StructuralBoundaryCondition):
        self.mesh = mesh
        self.materials = materials
        self.bcs = bcs
        n = mesh.num_dofs()
        fixed = bcs.fixed_dofs()
        mask = np.ones(n, dtype=bool)
        mask[fixed] = False
        self.fixed_dofs = np.unique(fixed)
        self.free_dofs = np.nonzero(mask)[0].astype(np.int64)

    def num_dofs(self):
        return self.mesh.num_dofs()

    def num_free_dofs(self):
        return self.free_dofs.size

    def elements(self):
        return [e for _, e in self.materials]


def free_dofs(s: Structure):
    return s.free_dofs


def mesh(s: Structure):
    return s.mesh


def materials(s: Structure):
    return s.materials


def boundary_conditions(s: Structure):
    return s.bcs
