"""Boundary conditions of the hot path and their `apply` rules.

FixedField  -- src/BoundaryConditions/FixedFieldBoundaryConditions.jl:23-48
GlobalLoad  -- src/BoundaryConditions/GlobalLoadBoundaryConditions.jl:21-89
Pressure    -- src/BoundaryConditions/LocalLoadBoundaryConditions.jl:21-56
StructuralBoundaryCondition -- src/StructuralModel/StructuralBoundaryConditions.jl

Only the displacement field ``:u`` with 3 components per node exists on this
path.  Components are 1-based, as in the reference.
"""
from __future__ import annotations

import numpy as np

from .entities import NodeSet, TriangularFaces, _Set


def _check_field(field):
    if str(field).lstrip(":") != "u":
        raise ValueError("only the displacement field :u is supported on this path")


class FixedField:
    """Null values of the given components of the field (FixedFieldBoundaryConditions.jl:23-34)."""

    def __init__(self, field, components, label=None):
        _check_field(field)
        self.field = "u"
        self.components = [int(c) for c in components]
        if any(c not in (1, 2, 3) for c in self.components):
            raise ValueError("components must be in 1..3")
        self.label = "no_label" if label is None else str(label)

    def apply(self, target):
        """Fixed dofs (0-based) of a node set, face set or element set, node by
        node in the target's node order (FixedFieldBoundaryConditions.jl:39-48)."""
        nodes = target.ids if isinstance(target, NodeSet) else target.conn.reshape(-1)
        comps = np.asarray(self.components, dtype=np.int64) - 1
        return (3 * nodes.astype(np.int64)[:, None] + comps[None, :]).reshape(-1)


class GlobalLoad:
    """Load in global coordinates; ``values(t)`` returns a 3-vector
    (GlobalLoadBoundaryConditions.jl:21-31)."""

    def __init__(self, field, values, label=None):
        _check_field(field)
        self.field = "u"
        self.values = values
        self.label = "no_label" if label is None else str(label)

    def __call__(self, t):
        return np.asarray(self.values(t), dtype=np.float64).reshape(3)

    def apply(self, target, t, coords):
        """(dofs, values) of the load at time t (GlobalLoadBoundaryConditions.jl:34-89):
        nodes get bc(t); faces bc(t) A / 3 per node; elements bc(t) V / nnodes per node."""
        f = self(t)
        if isinstance(target, NodeSet):
            nodes = target.ids
            vals = np.broadcast_to(f, (nodes.size, 3))
        elif isinstance(target, TriangularFaces):
            A = target.areas(coords)
            nodes = target.conn.reshape(-1)
            vals = np.repeat((A[:, None] * f[None, :]) / 3, 3, axis=0)
        elif isinstance(target, _Set):
            V = target.volumes(coords)
            k = target.nodes_per_entity
            nodes = target.conn.reshape(-1)
            vals = np.repeat((V[:, None] * f[None, :]) / k, k, axis=0)
        else:
            raise TypeError(f"GlobalLoad cannot be applied to {type(target)}")
        dofs = (3 * nodes.astype(np.int64)[:, None] + np.arange(3)[None, :]).reshape(-1)
        return dofs, np.asarray(vals, dtype=np.float64).reshape(-1)


class Pressure:
    """Pressure ``values(t)`` (a scalar) along minus the face normal, on the
    reference geometry (LocalLoadBoundaryConditions.jl:21-56).  Its faces are
    handed to the engine, which evaluates p(t) (-n) A / 3 per face node on the GPU."""

    def __init__(self, field, values, label=None):
        _check_field(field)
        v = values(np.random.rand())
        if not np.isscalar(v) and np.ndim(v) != 0:
            raise AssertionError("Pressure values must be a real function")
        self.field = "u"
        self.values = values
        self.label = "no_label" if label is None else str(label)

    def __call__(self, t):
        return float(self.values(t))

    def apply(self, target, t, coords):
        """Host form of the face load, for checks (LocalLoadBoundaryConditions.jl:36-56)."""
        if not isinstance(target, TriangularFaces):
            raise TypeError("Pressure applies to triangular faces")
        n = target.normal_directions(coords)
        A = target.areas(coords)
        p = self(t) * (-n) * A[:, None] / 3
        nodes = target.conn.reshape(-1)
        dofs = (3 * nodes.astype(np.int64)[:, None] + np.arange(3)[None, :]).reshape(-1)
        return dofs, np.repeat(p, 3, axis=0).reshape(-1)


class StructuralBoundaryCondition:
    """Boundary conditions mapped to their targets
    (src/StructuralModel/StructuralBoundaryConditions.jl).

    StructuralBoundaryCondition({bc: [target, ...], ...}) or from pairs."""

    def __init__(self, pairs=None, **kw):
        items = list(pairs.items()) if isinstance(pairs, dict) else list(pairs or [])
        self._pairs = [(bc, list(t) if isinstance(t, (list, tuple)) else [t]) for bc, t in items]

    def __iter__(self):
        return iter(self._pairs)

    def __getitem__(self, bc):
        for b, t in self._pairs:
            if b is bc:
                return t
        raise KeyError(bc)

    def fixed_dof_bcs(self):
        return [b for b, _ in self._pairs if isinstance(b, FixedField)]

    def load_bcs(self):
        return [b for b, _ in self._pairs if isinstance(b, (GlobalLoad, Pressure))]

    def fixed_dofs(self):
        """Union of fixed dofs, first-appearance order (StructuralBoundaryConditions.jl:170-185)."""
        out = [bc.apply(t) for bc, ts in self._pairs if isinstance(bc, FixedField) for t in ts]
        if not out:
            return np.zeros(0, dtype=np.int64)
        allf = np.concatenate(out)
        _, first = np.unique(allf, return_index=True)
        return allf[np.sort(first)]

    def has_nodal_loads(self):
        """True if any GlobalLoad is declared (else the nodal load vector is zero)."""
        return any(isinstance(bc, GlobalLoad) for bc, _ in self._pairs)

    def nodal_loads(self, t, coords, n):
        """Sum of every GlobalLoad at time t as a dense n-vector, duplicate dofs
        summed (StructuralBoundaryConditions.jl:195-220)."""
        f = np.zeros(n)
        for bc, ts in self._pairs:
            if isinstance(bc, GlobalLoad):
                for tg in ts:
                    d, v = bc.apply(tg, t, coords)
                    np.add.at(f, d, v)
        return f

    def pressure_groups(self):
        """[(Pressure bc, concatenated face connectivity)] in declaration order."""
        out = []
        for bc, ts in self._pairs:
            if isinstance(bc, Pressure):
                conn = np.concatenate([tg.conn for tg in ts], axis=0)
                out.append((bc, conn))
        return out
