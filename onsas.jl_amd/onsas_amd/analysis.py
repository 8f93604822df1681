"""Nonlinear static analysis, Newton-Raphson and the solve entry points.

ConvergenceSettings, criteria, ResidualsIterationStep, isconverged! --
    src/StructuralSolvers/StructuralSolvers.jl:36-171
NewtonRaphson -- src/StructuralSolvers/Solvers.jl:14-26
NonLinearStaticAnalysis, _solve!, step! --
    src/StructuralAnalyses/NonLinearStaticAnalyses.jl:32-148
solve / solve! -- src/StructuralSolvers/StructuralSolvers.jl:201-222

Julia's `name!` functions are spelled `name_` here (solve_, step_, assemble_,
isconverged_, reset_, update_, next_).  The host runs exactly the reference's
load-step / Newton control flow; every array operation of the loop body runs in
libonsas_hip through MI355XBackend.
"""
from __future__ import annotations

import copy
import json
import time
import warnings
from dataclasses import dataclass, field

import numpy as np

from .backend import GpuStaticState, MI355XBackend

INITIAL_DELTA = 1e12  # StructuralSolvers.jl:27


# ---- convergence (StructuralSolvers.jl:36-171) ----------------------------------

@dataclass(frozen=True)
class ConvergenceSettings:
    """ConvergenceSettings(rel_U_tol, rel_res_force_tol, max_iter) (StructuralSolvers.jl:36-43)."""
    rel_U_tol: float = 1e-6
    rel_res_force_tol: float = 1e-6
    max_iter: int = 20


def residual_forces_tol(x):
    if isinstance(x, ConvergenceSettings):
        return x.rel_res_force_tol
    return (x.dr_rel, x.dr_norm)


def displacement_tol(x):
    if isinstance(x, ConvergenceSettings):
        return x.rel_U_tol
    return (x.dU_rel, x.dU_norm)


def max_iter_tol(cs):
    return cs.max_iter


class AbstractConvergenceCriterion:
    def __repr__(self):
        return type(self).__name__ + "()"

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self))


class ResidualForceCriterion(AbstractConvergenceCriterion):
    pass


class ΔUCriterion(AbstractConvergenceCriterion):
    pass


DeltaUCriterion = ΔUCriterion


class ΔU_and_ResidualForce_Criteria(AbstractConvergenceCriterion):
    pass


class MaxIterCriterion(AbstractConvergenceCriterion):
    pass


class NotConvergedYet(AbstractConvergenceCriterion):
    pass


@dataclass
class ResidualsIterationStep:
    """Convergence record of the current iteration (StructuralSolvers.jl:85-97)."""
    dU_norm: float = INITIAL_DELTA
    dr_norm: float = INITIAL_DELTA
    dU_rel: float = INITIAL_DELTA
    dr_rel: float = INITIAL_DELTA
    iter: int = 0
    criterion: AbstractConvergenceCriterion = field(default_factory=NotConvergedYet)


def iterations(x):
    return x.iter if isinstance(x, ResidualsIterationStep) else x.iterations()


def criterion(x):
    return x.criterion if isinstance(x, ResidualsIterationStep) else x.criteria()


def reset_(ri: ResidualsIterationStep):
    """reset! (StructuralSolvers.jl:115-121)."""
    ri.dU_norm = ri.dr_norm = ri.dU_rel = ri.dr_rel = INITIAL_DELTA
    ri.iter = 0
    ri.criterion = NotConvergedYet()
    return ri


def update_(ri: ResidualsIterationStep, *args):
    """update!(ri, criterion) or update!(ri, dU_norm, dU_rel, dr_norm, dr_rel)
    (StructuralSolvers.jl:131-145)."""
    if len(args) == 1:
        ri.criterion = args[0]
        return ri
    ri.dU_norm, ri.dU_rel, ri.dr_norm, ri.dr_rel = (float(a) for a in args)
    ri.iter += 1
    return ri


def isconverged_(ri: ResidualsIterationStep, cs: ConvergenceSettings):
    """isconverged! (StructuralSolvers.jl:148-171): residual criterion first, then
    dU, then the iteration cap (warns), else not converged."""
    dU_rel, _ = displacement_tol(ri)
    dr_rel, _ = residual_forces_tol(ri)
    if not dU_rel > 0:
        raise AssertionError("Residual displacements norm must be greater than 0.")
    if not dr_rel > 0:
        raise AssertionError("Residual forces norm must be greater than 0.")
    if dr_rel <= cs.rel_res_force_tol:
        c = ResidualForceCriterion()
    elif dU_rel <= cs.rel_U_tol:
        c = ΔUCriterion()
    elif ri.iter > cs.max_iter:
        warnings.warn("Maximum number of iterations was reached.")
        c = MaxIterCriterion()
    else:
        c = NotConvergedYet()
    update_(ri, c)
    return c


# ---- solvers (Solvers.jl:14-26) -------------------------------------------------

class NewtonRaphson:
    def __init__(self, tol: ConvergenceSettings = None):
        self.tol = tol if tol is not None else ConvergenceSettings()


def tolerances(alg):
    return alg.tol


# ---- analysis (NonLinearStaticAnalyses.jl:32-60) ---------------------------------

class NonLinearStaticAnalysis:
    """NonLinearStaticAnalysis(s, t1=1.0; NSTEPS=10, initial_step=1) with load
    factors LinRange(t1/NSTEPS, t1, NSTEPS), or NonLinearStaticAnalysis(s,
    load_factors=...) (NonLinearStaticAnalyses.jl:45-60)."""

    def __init__(self, s, t1=1.0, NSTEPS=10, initial_step=1, load_factors=None,
                 initial_state=None):
        if load_factors is None:
            t0 = t1 / NSTEPS
            load_factors = np.linspace(t0, t1, NSTEPS)
        self.s = s
        self.load_factors = np.asarray(load_factors, dtype=np.float64)
        if not 1 <= initial_step <= len(self.load_factors):
            raise ValueError(f"initial_step must be in [1, {len(self.load_factors)}] but is: "
                             f"{initial_step}.")
        self.current_step = int(initial_step)
        self.initial_state = initial_state  # optional U to resume from
        self.state = None                   # GpuStaticState, created by solve_
        self.iter_state = ResidualsIterationStep()

    def is_done(self):
        """is_done (StaticAnalyses.jl:70-77)."""
        if self.current_step > len(self.load_factors):
            self.current_step -= 1
            return True
        return False

    def current_time(self):
        return float(self.load_factors[self.current_step - 1])

    def next_(self):
        self.current_step += 1


def structure(sa):
    return sa.s


def load_factors(sa):
    return sa.load_factors


def current_iteration(sa):
    return sa.iter_state


def current_state(sa):
    return sa.state


# ---- solution (src/StructuralSolvers/Solutions.jl) --------------------------------

class Solution:
    """Per-load-step states: U, per-group stress/strain (at the last assembly
    of the step, like store!, StaticAnalyses.jl:157-174) and the convergence
    record, plus the engine's linear-solver statistics."""

    def __init__(self, sa, alg):
        self.analysis = sa
        self.algorithm = alg
        self.U = [None] * len(sa.load_factors)
        self.stress = [None] * len(sa.load_factors)
        self.strain = [None] * len(sa.load_factors)
        self.records = [None] * len(sa.load_factors)
        self.newton_log = []  # one dict per Newton iteration
        self.step_wall_s = []  # host wall time of each load step
        self.wall_s = 0.0

    def displacements(self, step=None, node=None, component=None):
        U = self.U if step is None else [self.U[step]]
        out = np.array([u for u in U])
        if node is not None:
            out = out[:, 3 * node: 3 * node + 3]
            if component is not None:
                out = out[:, component - 1]
        return out if step is None else out[0]

    def criteria(self):
        return [r["criterion"] if r else None for r in self.records]

    def iterations(self):
        return [r["iterations"] if r else None for r in self.records]

    def linear_iterations(self):
        return sum(d["lin_iters"] for d in self.newton_log)

    def metrics(self):
        """One record per stored load step: the reference's per-step table
        (Solutions.jl:106-134: ||U||, ||dU||/||U||, ||dR||, ||dR||/||Fext||,
        criterion, Newton iterations) plus the engine's inner solves (iterations,
        GMRES fallbacks, tolerances) and phase times (device ms, host wall s)."""
        out = []
        for k, r in enumerate(self.records):
            if r is None:
                continue
            its = [d for d in self.newton_log if d.get("load_step") == k + 1]
            out.append(step_metrics(k + 1, self.analysis.load_factors[k], self.U[k], r, its,
                                    self.step_wall_s[k] if k < len(self.step_wall_s) else None))
        return out

    def write_jsonl(self, path):
        """The per-step records as JSON lines."""
        with open(path, "w") as f:
            for m in self.metrics():
                f.write(json.dumps(m) + "\n")

    def table(self):
        """Rows of _print_table (Solutions.jl:106-134)."""
        rows = []
        for k, r in enumerate(self.records):
            if r is None:
                continue
            rows.append((k + 1, np.linalg.norm(self.U[k]), r["dU_rel"], r["dr_norm"], r["dr_rel"],
                         r["criterion"], r["iterations"]))
        return rows


class _UStore:
    """store!'s per-step copy of U, overlapped with the next load step: the
    device copy is enqueued at once into one of two pinned snapshot buffers,
    and a worker thread waits for it and fills the step's array (a fresh
    3n-value array: page faults and a host copy of ~1 ms at 64^3) while this
    thread drives the next step.  finish() joins the worker."""

    def __init__(self, state):
        from concurrent.futures import ThreadPoolExecutor
        self.ctx, self.n = state.ctx, state.n
        self.pool = ThreadPoolExecutor(max_workers=1)
        self.pending = [None, None]
        self.k = 0

    def take(self):
        slot = self.k % 2
        self.k += 1
        if self.pending[slot] is not None:
            self.pending[slot].result()  # the slot is read before it is reused
        self.ctx.snapshot_U(slot)
        out = np.empty(self.n)
        self.pending[slot] = self.pool.submit(self.ctx.read_snapshot, slot, out)
        return out

    def finish(self):
        try:
            for f in self.pending:
                if f is not None:
                    f.result()
        finally:
            self.pool.shutdown()


def store_(sol: Solution, sa: NonLinearStaticAnalysis, step: int, ustore: _UStore = None):
    st = sa.state
    sol.U[step - 1] = ustore.take() if ustore is not None else st.displacements()
    if st.backend.store_stresses:
        ss = st.stress_strain(at_last_assembly=True)
        sol.stress[step - 1] = [s for _, s, _ in ss]
        sol.strain[step - 1] = [e for _, _, e in ss]
    ri = sa.iter_state
    sol.records[step - 1] = dict(dU_norm=ri.dU_norm, dU_rel=ri.dU_rel, dr_norm=ri.dr_norm,
                                 dr_rel=ri.dr_rel, iterations=ri.iter, criterion=ri.criterion)


_LIN_SOLVER = {1: "pcg", 2: "gmres", 3: "pcg+gmres"}


def step_metrics(step, t, U, rec, its, wall_s):
    """The JSON record of one load step (see Solution.metrics); its: the step's
    Newton iterations as newton_step returned them."""
    crit = rec["criterion"]
    return dict(
        step=int(step), time=float(t), U_norm=float(np.linalg.norm(U)) if U is not None else None,
        dU_rel=float(rec["dU_rel"]), dr_norm=float(rec["dr_norm"]), dr_rel=float(rec["dr_rel"]),
        criterion=type(crit).__name__ if not isinstance(crit, str) else crit,
        newton_iterations=int(rec["iterations"]),
        lin_iters=int(sum(d["lin_iters"] for d in its)),
        lin_iters_per_iteration=[int(d["lin_iters"]) for d in its],
        lin_solver=[_LIN_SOLVER.get(int(d["lin_solver"]), "none") for d in its],
        gmres_fallbacks=int(sum(1 for d in its if int(d["lin_solver"]) == 3)),
        lin_rtol=[float("%.3e" % d["lin_rtol"]) for d in its],
        lin_converged=all(bool(d["lin_converged"]) for d in its),
        assemble_ms=round(float(sum(d["assemble_ms"] for d in its)), 4),
        solve_ms=round(float(sum(d["solve_ms"] for d in its)), 4),
        wall_s=None if wall_s is None else round(float(wall_s), 6))


# ---- the hot path ---------------------------------------------------------------

def assemble_(s, sa):
    """assemble!(s, sa) (StaticAnalyses.jl:99-122): f_int and K at the current U, on the GPU."""
    sa.state.ctx.assemble()


def step_(sa, alg, linear_solver=None, sol=None):
    """step!(sa, ::NewtonRaphson, linear_solver) (NonLinearStaticAnalyses.jl:107-148):
    r = f_ext - f_int, solve K dU = r, the four norms, U[free] += dU, update!."""
    stats = sa.state.ctx.newton_step()
    update_(sa.iter_state, stats.dU_norm, stats.dU_rel, stats.r_norm, stats.r_rel)
    if sol is not None:
        d = stats.as_dict()
        d["load_step"] = sa.current_step
        d["iteration"] = sa.iter_state.iter
        sol.newton_log.append(d)
    return stats


def _solve_nonlinear_(sa, alg, backend):
    """_solve!(::NonLinearStaticAnalysis, alg, linear_solver)
    (NonLinearStaticAnalyses.jl:70-104) with the MI355X engine."""
    s = sa.s
    if sa.state is None:
        sa.state = GpuStaticState(s, backend)
    if backend.forcing:
        sa.state.configure_linear_solver(newton_tol=tolerances(alg).rel_res_force_tol)
    if sa.initial_state is not None:
        sa.state.ctx.set_U(sa.initial_state)
    sol = Solution(sa, alg)
    t0 = time.perf_counter()
    ustore = _UStore(sa.state)
    jsonl = open(backend.metrics_jsonl, "w") if backend.metrics_jsonl else None
    try:
        while not sa.is_done():
            step = sa.current_step
            ts = time.perf_counter()
            reset_(sa.iter_state)
            sa.state.apply_loads(sa.current_time())
            while isinstance(isconverged_(sa.iter_state, tolerances(alg)), NotConvergedYet):
                assemble_(s, sa)
                step_(sa, alg, backend, sol)
            store_(sol, sa, step, ustore)
            sol.step_wall_s.append(time.perf_counter() - ts)
            if jsonl is not None:  # the step's record as it completes (U's norm comes later)
                its = [d for d in sol.newton_log if d.get("load_step") == step]
                jsonl.write(json.dumps(step_metrics(step, sa.current_time(), None,
                                                    sol.records[step - 1], its,
                                                    sol.step_wall_s[-1])) + "\n")
                jsonl.flush()
            sa.next_()
    finally:
        ustore.finish()  # every stored U is complete before the solution is returned
        if jsonl is not None:
            jsonl.close()
    sol.wall_s = time.perf_counter() - t0
    return sol


def solve_(problem, solver=None, linear_solve=None):
    """solve!(problem, solver, linear_solve) (StructuralSolvers.jl:217-222)."""
    if linear_solve is None:
        linear_solve = MI355XBackend()
    if not isinstance(linear_solve, MI355XBackend):
        raise TypeError("the engine path needs linear_solve = MI355XBackend(...)")
    if isinstance(problem, NonLinearStaticAnalysis):
        return _solve_nonlinear_(problem, solver or NewtonRaphson(), linear_solve)
    from .linear_analysis import LinearStaticAnalysis, _solve_linear_
    if isinstance(problem, LinearStaticAnalysis):
        return _solve_linear_(problem, linear_solve)
    raise TypeError(f"unsupported problem {type(problem)}")


def solve(problem, solver=None, linear_solve=None):
    """solve(problem, ...) = solve!(deepcopy(problem), ...) (StructuralSolvers.jl:201-206)."""
    p = copy.copy(problem)
    p.iter_state = ResidualsIterationStep()
    p.state = None
    return solve_(p, solver, linear_solve)
