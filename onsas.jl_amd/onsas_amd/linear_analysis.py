"""LinearStaticAnalysis on the engine (src/StructuralAnalyses/LinearStaticAnalyses.jl:
25-153): K assembled at the first load step, U = K^-1 f_ext per step (same
block-Jacobi PCG), a re-assembly per step for the stresses and strains.  In
mixed mode the FP32-tangent solve is followed by refinement steps against the
FP64 residual until ||f_ext - f_int(U)|| <= 1e-10 ||f_ext||.
"""
from __future__ import annotations

import time
import warnings

import numpy as np

from .analysis import ΔU_and_ResidualForce_Criteria, ResidualsIterationStep, Solution, store_
from .backend import GpuStaticState

REFINE_TOL = 1e-10  # relative residual the mixed mode refines to
REFINE_MAX = 4


class LinearStaticAnalysis:
    """LinearStaticAnalysis(s, final_time=1.0; NSTEPS=10, initial_step=1) or with
    load_factors (LinearStaticAnalyses.jl:48-68)."""

    def __init__(self, s, final_time=1.0, NSTEPS=10, initial_step=1, load_factors=None):
        if load_factors is None:
            load_factors = np.linspace(final_time / NSTEPS, final_time, NSTEPS)
        self.s = s
        self.load_factors = np.asarray(load_factors, dtype=np.float64)
        if not 1 <= initial_step <= len(self.load_factors):
            raise ValueError(f"initial_step must be in [1, {len(self.load_factors)}] but is: "
                             f"{initial_step}.")
        self.current_step = int(initial_step)
        self.state = None
        # LinearResidualsIterationStep (LinearStaticAnalyses.jl:26-28)
        self.iter_state = ResidualsIterationStep(None, None, None, None, 1,
                                                 ΔU_and_ResidualForce_Criteria())
        self.initial_state = None

    def is_done(self):
        if self.current_step > len(self.load_factors):
            self.current_step -= 1
            return True
        return False

    def current_time(self):
        return float(self.load_factors[self.current_step - 1])

    def next_(self):
        self.current_step += 1


def _solve_linear_(sa: LinearStaticAnalysis, backend):
    """_solve!(::LinearStaticAnalysis, nothing, linear_solver) (LinearStaticAnalyses.jl:78-114)."""
    if sa.state is None:
        sa.state = GpuStaticState(sa.s, backend)
    st = sa.state
    ctx = st.ctx
    sol = Solution(sa, None)
    sol.refine_r_rel = []  # mixed mode: ||f_ext - f_int(U)|| / ||f_ext|| after refinement
    t0 = time.perf_counter()
    while not sa.is_done():
        step = sa.current_step
        st.apply_loads(sa.current_time())
        if step == 1:
            ctx.assemble()
        # step!: res = f_ext[free]; solve K dU = res; U[free] .= dU (:117-153)
        fext = ctx.get_fext()
        b = np.zeros_like(fext)
        b[st.free_dofs] = fext[st.free_dofs]
        x, it, rr = ctx.linear_solve(b)
        U = np.zeros_like(fext)
        U[st.free_dofs] = x[st.free_dofs]
        ctx.set_U(U)
        sol.newton_log.append({"load_step": step, "lin_iters": it, "lin_rel_residual": rr})
        ctx.assemble()
        if backend.precision == "mixed":
            # the FP32 tangent solve is refined against the FP64 residual
            # f_ext - f_int(U) (f_int = K U in FP64 for a linear material) until
            # it meets the FP64 solve's accuracy; the residual of the last pass
            # is recomputed at the final U, and a step that still misses the
            # tolerance is reported (warning + refine_r_rel), not hidden
            r_rel = None
            for _ in range(REFINE_MAX):
                stats = ctx.newton_step()
                ctx.assemble()
                d = stats.as_dict()
                d["load_step"] = step
                sol.newton_log.append(d)
                r_rel = stats.r_rel
                if stats.r_rel <= REFINE_TOL or stats.fext_norm == 0:
                    break
            else:
                fext = ctx.get_fext()
                nf = np.linalg.norm(fext)
                r_rel = (np.linalg.norm((fext - ctx.get_fint())[st.free_dofs]) / nf
                         if nf > 0 else 0.0)
            sol.refine_r_rel.append(r_rel)
            if r_rel is not None and r_rel > REFINE_TOL:
                warnings.warn(f"LinearStaticAnalysis step {step}: mixed-precision refinement "
                              f"stopped after {REFINE_MAX} passes at ||r||/||f_ext|| = "
                              f"{r_rel:.3e} > {REFINE_TOL:g}", RuntimeWarning)
        store_(sol, sa, step)
        sa.next_()
    sol.wall_s = time.perf_counter() - t0
    return sol
