"""Evaluation of nodal fields at arbitrary points (src/Meshes/Handlers.jl:24-129,
src/Meshes/Interpolators.jl:70-138), the point search on the GPU.

PointEvalHandler(state, points) locates every point in the reference
configuration of a tet group -- the first element in element order that
contains it, as the reference's Serial search (src/Meshes/Searches.jl:64-82)
-- with its barycentric weights (Tetrahedrons.jl:311-313); `interpolate`
sums weight x nodal value (Interpolators.jl:87-107).
"""
from __future__ import annotations

import numpy as np

from .entities import Tetrahedra


class PointEvalHandler:
    def __init__(self, state, points, element_set=None):
        """state: a GpuStaticState (or a Solution whose analysis still holds
        one); points: (n, 3).  element_set: the tet set to search (default:
        the first tet group)."""
        st = state.analysis.state if hasattr(state, "analysis") else state
        groups = [(es, gid) for _, es, gid in st.groups if isinstance(es, Tetrahedra)]
        if not groups:
            raise TypeError("point evaluation needs a tetrahedron element set")
        es, gid = groups[0] if element_set is None else next(
            (e, g) for e, g in groups if e is element_set or e.label == element_set)
        self.points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
        elem, w = st.ctx.locate_points(gid, self.points)
        inside = elem >= 0
        self.element_set = es
        self.in_mesh_points_idx = np.nonzero(inside)[0]
        self.not_in_mesh_points_idx = np.nonzero(~inside)[0]
        self.in_mesh_elements_idx = elem[inside]
        self.weights = w[inside]
        self.nodes = np.asarray(es.conn)[self.in_mesh_elements_idx]  # (n_in, 4)

    @property
    def in_mesh_points(self):
        return self.points[self.in_mesh_points_idx]

    @property
    def not_in_mesh_points(self):
        return self.points[self.not_in_mesh_points_idx]


def interpolate(handler: PointEvalHandler, nodal):
    """Interpolated nodal magnitude at the in-mesh points: nodal is (nn,) or
    (nn, k) (e.g. U.reshape(-1, 3)); returns (n_in,) or (n_in, k)."""
    nodal = np.asarray(nodal)
    vals = nodal[handler.nodes]  # (n_in, 4[, k])
    if vals.ndim == 2:
        return np.einsum("pa,pa->p", handler.weights, vals)
    return np.einsum("pa,pak->pk", handler.weights, vals)


def displacements(sol, handler: PointEvalHandler, step=-1):
    """Displacements of a solution at the handler's in-mesh points, load step `step`."""
    return interpolate(handler, np.asarray(sol.U[step]).reshape(-1, 3))
