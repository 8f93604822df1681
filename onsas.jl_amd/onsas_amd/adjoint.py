"""Reverse-mode sensitivity of a converged nonlinear static solve to the
per-element shear moduli of its NeoHookean / SVK tetrahedra.

The reference has no adjoint.  This is the rrule a ChainRulesCore user would
write around `solve!` (src/StructuralSolvers/StructuralSolvers.jl:217-222) for
the shear modulus G of `NeoHookean` (src/Materials/NeoHookeanMaterial.jl:94-108)
and `SVK` (src/Materials/SVKMaterial.jl:89-100): at the converged state
R(U, G) = f_int(U, G) - f_ext = 0, so for a scalar J(U)

    K lambda = dJ/dU   (K symmetric, free dofs),   dJ/dG_e = -lambda . d f_int / d G_e,

with d f_int / d G_e evaluated in closed form on the device
(NeoHookean: dP/dG = F - F^-T; SVK: dP/dG = 2 F E).
"""
from __future__ import annotations

import numpy as np

from .analysis import NonLinearStaticAnalysis, solve, solve_
from .entities import Tetrahedra
from .materials import NeoHookean, SVK


def _shear_groups(state):
    return [(mat, es, gid) for mat, es, gid in state.groups
            if isinstance(es, Tetrahedra) and isinstance(mat, (NeoHookean, SVK))]


def rrule(f, problem, solver=None, linear_solve=None):
    """rrule(solve, problem, solver, backend) -> (solution, pullback).

    pullback(Ubar) takes the cotangent of the displacements of the LAST load
    step (n values, full dof vector; fixed dofs are ignored) and returns
    {element-set label: dJ/dG (ne,)} for every NeoHookean / SVK tet group.
    The device model stays alive until `solution.analysis.state.close()`."""
    if f is not solve:
        raise TypeError("rrule is defined for onsas_amd.solve")
    sol = solve(problem, solver, linear_solve)
    state = sol.analysis.state

    def pullback(Ubar):
        Ubar = np.asarray(Ubar, dtype=np.float64)
        out = {}
        for mat, es, gid in _shear_groups(state):
            g, _ = state.ctx.adjoint_shear(gid, Ubar)
            out[es.label] = g
        return out

    return sol, pullback


def resolve(sol, G, alg, backend, element_set=None):
    """Re-solve the last load step of a finished solution with new shear moduli
    G (ne,) on the same device model, warm-started from its converged
    displacements (the inner loop of a gradient-based identification).
    Returns the new Solution (one step, load factor of the last step)."""
    sa_old = sol.analysis
    state = sa_old.state
    groups = _shear_groups(state)
    if element_set is None:
        if len(groups) != 1:
            raise ValueError("several shear groups: name the element set")
        gid = groups[0][2]
    else:
        gid = next(g for m, es, g in groups if es is element_set or es.label == element_set)
    state.ctx.set_shear_moduli(gid, G)
    U0 = state.displacements() if sol.U[-1] is None else sol.U[-1]
    sa = NonLinearStaticAnalysis(sa_old.s, load_factors=[sa_old.load_factors[-1]],
                                 initial_state=U0)
    sa.state = state
    return solve_(sa, alg, backend)


# ---- misfit functional of the identification workload -----------------------------

def misfit(U, obs_nodes, target):
    """J = 1/2 sum over observed nodes of |u_i - target_i|^2 and dJ/dU (n,)."""
    U = np.asarray(U)
    d = U.reshape(-1, 3)[obs_nodes] - target
    g = np.zeros_like(U).reshape(-1, 3)
    g[obs_nodes] = d
    return 0.5 * float(np.sum(d * d)), g.reshape(-1)
