"""Per-load-step metric records (Solution.metrics / write_jsonl): the
reference's per-step table (src/StructuralSolvers/Solutions.jl:106-134) plus
the engine's inner-solver data and phase times, one JSON object per line."""
import json

import numpy as np

from onsas_amd.analysis import ResidualForceCriterion, Solution, step_metrics


class _FakeAnalysis:
    load_factors = [0.5, 1.0]


def _solution():
    sol = Solution(_FakeAnalysis(), None)
    sol.U = [np.array([3.0, 4.0]), np.array([6.0, 8.0])]
    crit = ResidualForceCriterion()
    sol.records = [dict(dU_norm=1e-3, dU_rel=1e-4, dr_norm=2e-9, dr_rel=1e-9, iterations=3,
                        criterion=crit),
                   dict(dU_norm=1e-3, dU_rel=1e-4, dr_norm=3e-9, dr_rel=2e-9, iterations=2,
                        criterion=crit)]
    it = dict(lin_iters=5, lin_solver=1, lin_converged=1, assemble_ms=0.4, solve_ms=1.0,
              lin_rtol=5e-3)
    sol.newton_log = [dict(it, load_step=1, iteration=1), dict(it, load_step=1, iteration=2),
                      dict(it, load_step=1, iteration=3, lin_solver=3, lin_iters=9),
                      dict(it, load_step=2, iteration=1), dict(it, load_step=2, iteration=2)]
    sol.step_wall_s = [0.01, 0.02]
    return sol


def test_metrics_fields_and_sums():
    m = _solution().metrics()
    assert [r["step"] for r in m] == [1, 2] and [r["time"] for r in m] == [0.5, 1.0]
    r = m[0]
    assert r["U_norm"] == 5.0 and r["criterion"] == "ResidualForceCriterion"
    assert r["newton_iterations"] == 3 and r["lin_iters"] == 19
    assert r["lin_iters_per_iteration"] == [5, 5, 9]
    assert r["lin_solver"] == ["pcg", "pcg", "pcg+gmres"] and r["gmres_fallbacks"] == 1
    assert r["assemble_ms"] == 1.2 and r["solve_ms"] == 3.0 and r["wall_s"] == 0.01
    assert m[1]["lin_iters"] == 10 and m[1]["gmres_fallbacks"] == 0


def test_write_jsonl_one_object_per_line(tmp_path):
    sol = _solution()
    path = tmp_path / "steps.jsonl"
    sol.write_jsonl(path)
    lines = path.read_text().splitlines()
    assert len(lines) == 2
    assert [json.loads(x) for x in lines] == sol.metrics()


def test_step_metrics_without_U():
    rec = dict(dU_norm=0.0, dU_rel=0.0, dr_norm=0.0, dr_rel=0.0, iterations=1, criterion="x")
    r = step_metrics(1, 1.0, None, rec, [], None)
    assert r["U_norm"] is None and r["lin_iters"] == 0 and r["wall_s"] is None
    json.dumps(r)
