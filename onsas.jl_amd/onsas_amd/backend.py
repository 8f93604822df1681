"""MI355XBackend: the linear-solve argument of solve! that moves the Newton hot
path onto the GPU, and the device-resident static state it drives.

In the reference, `solve!(problem, solver, linear_solve)` accepts any
`SciMLBase.AbstractLinearAlgorithm` (src/StructuralSolvers/StructuralSolvers.jl:
217-222, :31).  Passing an MI355XBackend instead selects the engine: the
element loop, assembly, residual, linear solve and update all run in
libonsas_hip; the host keeps the load-step / Newton control flow.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _capi
from .entities import Tetrahedra, Trusses
from .materials import IsotropicLinearElastic, NeoHookean, SVK

_SOLVERS = {"auto": _capi.LINSOLVE_AUTO, "pcg": _capi.LINSOLVE_PCG, "gmres": _capi.LINSOLVE_GMRES}
_PRECONDS = {"block-jacobi": _capi.PRECOND_BLOCK_JACOBI, "amg": _capi.PRECOND_AMG}
_AMG_REFRESH = {"step": _capi.AMG_REFRESH_STEP, "assembly": _capi.AMG_REFRESH_ASSEMBLY,
                "once": _capi.AMG_REFRESH_ONCE}


@dataclass
class MI355XBackend:
    """precision: "mixed" (FP32 tangent + PCG, FP64 f_int / residual / norms / U)
    or "fp64".  rtol/maxiter: inner-solve stopping rule (None: sqrt(eps) in fp64
    mode as the reference, 1e-6 in mixed mode; maxiter = free dofs).  solver:
    "auto" (PCG, GMRES if the tangent is indefinite), "pcg" or "gmres".
    preconditioner: "block-jacobi" (3x3 nodal blocks) or "amg" (smoothed-
    aggregation multigrid V-cycle with rigid-body modes and Chebyshev smoothing).
    amg_refresh: "step" (coarse levels rebuilt once per load step), "assembly"
    (after every assembly) or "once" (built from the first tangent only); the
    finest level is always the current tangent.  metrics_jsonl: a file that
    receives one JSON record per load step while the solve runs (the fields of
    Solution.metrics, with U_norm null)."""

    precision: str = "mixed"
    rtol: float = None
    maxiter: int = None
    solver: str = "auto"
    gmres_restart: int = 40
    check_every: int = None  # PCG iterations per replayed chunk (None: 16, multigrid 8)
    store_stresses: bool = True
    # inexact Newton: per-iteration inner tolerance (Eisenstat-Walker with a
    # floor at what the outer tolerance needs); rtol is then the tightest one
    forcing: bool = False
    eta_max: float = 0.1
    eta_first: float = 5e-3  # the fastest first inner tolerance on the 64^3 cube, both modes
    preconditioner: str = "block-jacobi"
    amg_refresh: str = "step"
    # path of a JSON-lines file that receives one record per load step as the
    # solve goes (Solution.metrics gives the same records afterwards)
    metrics_jsonl: str = None

    def precision_code(self):
        if self.precision not in ("mixed", "fp64"):
            raise ValueError("precision must be 'mixed' or 'fp64'")
        return _capi.PRECISION_MIXED if self.precision == "mixed" else _capi.PRECISION_FP64


class GpuStaticState:
    """Device-resident counterpart of FullStaticState
    (src/StructuralAnalyses/StaticStates.jl:33-102): U, f_ext, f_int, K live in
    HBM; accessors copy them back."""

    def __init__(self, structure, backend: MI355XBackend):
        self.s = structure
        self.backend = backend
        self.ctx = _capi.Context(backend.precision_code())
        ctx = self.ctx
        m = structure.mesh
        ctx.set_nodes(m.coords)
        self.groups = []  # (material, element set, group id)
        for mat, es in structure.materials:
            if isinstance(es, Tetrahedra):
                G = np.asarray(getattr(mat, "G", 0.0))
                per = None
                params = mat.params()
                if G.ndim:
                    if G.size != len(es):
                        raise ValueError("per-element shear modulus needs one value per element")
                    per = G
                    params = (params[0], float(np.median(G)))
                gid = ctx.add_tet_group(es.conn, mat.kind, params, per)
            elif isinstance(es, Trusses):
                if isinstance(mat, IsotropicLinearElastic):
                    raise NotImplementedError(
                        "internal_forces(::IsotropicLinearElastic, ::Truss) is not defined")
                if not isinstance(mat, (SVK, NeoHookean)):
                    raise TypeError(f"unsupported truss material {type(mat)}")
                gid = ctx.add_truss_group(es.conn, es.strain_model.code, mat.elasticity_modulus(),
                                          es.areas())
            else:
                raise TypeError(f"unsupported element set {type(es)}")
            self.groups.append((mat, es, gid))
        self.pressure_bcs = []
        for bc, conn in structure.bcs.pressure_groups():
            ctx.add_pressure_faces(conn)
            self.pressure_bcs.append(bc)
        ctx.set_fixed_dofs(structure.fixed_dofs)
        ctx.finalize()
        self.configure_linear_solver()
        self.n = m.num_dofs()
        self.free_dofs = structure.free_dofs

    def configure_linear_solver(self, newton_tol=0.0):
        b = self.backend
        chunk = b.check_every or (8 if b.preconditioner == "amg" else 16)
        self.ctx.set_linear_solver(_SOLVERS[b.solver], b.rtol or 0.0, b.maxiter or 0,
                                   b.gmres_restart, chunk, b.forcing, newton_tol,
                                   b.eta_max, b.eta_first, _PRECONDS[b.preconditioner],
                                   _AMG_REFRESH[b.amg_refresh])

    # -- loads
    def apply_loads(self, t):
        """external_forces .= 0; apply!(sa, load_bcs) at time t
        (NonLinearStaticAnalyses.jl:86-87, StructuralAnalyses.jl:228-241)."""
        # no GlobalLoad: the engine clears its nodal loads itself (no 3n upload)
        nodal = (self.s.bcs.nodal_loads(t, self.s.mesh.coords, self.n)
                 if self.s.bcs.has_nodal_loads() else None)
        pressures = [bc(t) for bc in self.pressure_bcs]
        self.ctx.set_loads(nodal, pressures)

    # -- accessors (StructuralAnalyses.jl:70-100)
    def displacements(self):
        return self.ctx.get_U()

    def external_forces(self):
        return self.ctx.get_fext()

    def internal_forces(self):
        return self.ctx.get_fint()

    def tangent_matrix(self):
        return self.ctx.tangent_bsr()

    def residual_forces(self):
        """f_ext[free] - f_int[free] (StaticStates.jl:113-116)."""
        return (self.external_forces() - self.internal_forces())[self.free_dofs]

    def stress_strain(self, at_last_assembly=True):
        """[(element set, stress (ne,3,3), strain (ne,3,3))] per material group."""
        out = []
        for mat, es, gid in self.groups:
            s, e = self.ctx.stress_strain(gid, at_last_assembly)
            out.append((es, s, e))
        return out

    def close(self):
        self.ctx.close()


def internal_forces(material, elements, U, backend: MI355XBackend = None, coords=None):
    """Batch internal_forces(mat, e, u_e) over an element set
    (src/Entities/Entities.jl:173-218): returns (f (ne, nd), K (ne, nd, nd),
    stress (ne, 3, 3), strain (ne, 3, 3)), computed on the GPU."""
    from .boundary_conditions import StructuralBoundaryCondition
    from .entities import Mesh
    from .structure import Structure, StructuralMaterial
    if coords is None:
        raise ValueError("coords (the mesh node coordinates) are required")
    backend = backend or MI355XBackend(precision="fp64")
    s = Structure(Mesh(coords, [elements]), StructuralMaterial({material: elements}),
                  StructuralBoundaryCondition({}))
    st = GpuStaticState(s, backend)
    try:
        st.ctx.set_U(U)
        return st.ctx.element_forces(st.groups[0][2])
    finally:
        st.close()
