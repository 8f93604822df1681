"""GPU parity: the HIP engine (called through the C ABI) against the FP64 oracle
and the reference examples' closed forms.

Tolerances: FP64 quantities (f_int, residual, U, element outputs) to ~1e-12
relative; the FP32 tangent of the mixed mode to 2e-6 of the largest entry of
its row block; converged displacements to 1e-8 relative of max|U| when the
Newton residual is driven below 1e-10 (both precisions iterate to the same FP64
residual, so the converged states must agree).
"""
import contextlib
import os

import numpy as np
import pytest

from _helpers import fext_steps
from onsas_amd import _capi, meshes, workloads
from onsas_amd.analysis import ConvergenceSettings, NewtonRaphson, NonLinearStaticAnalysis, solve
from onsas_amd.backend import MI355XBackend

pytestmark = pytest.mark.gpu

PREC = {"mixed": _capi.PRECISION_MIXED, "fp64": _capi.PRECISION_FP64}
MATS = {"svk": (_capi.MAT_SVK, (0.6e6, 0.4e6)), "nh": (_capi.MAT_NEOHOOKEAN, (1.3e6, 0.4e6)),
        "linear": (_capi.MAT_LINEAR, (1.0e6, 0.3))}


@contextlib.contextmanager
def _assembly_mode(mode):
    """ONSAS_ASSEMBLY and ONSAS_GATHER are read when a context is finalized;
    mode "staged" is the gather assembly with its states staged in LDS."""
    old = {k: os.environ.get(k) for k in ("ONSAS_ASSEMBLY", "ONSAS_GATHER")}
    if mode == "staged":
        os.environ.pop("ONSAS_ASSEMBLY", None)
        os.environ["ONSAS_GATHER"] = "staged"
    elif mode is not None:
        os.environ["ONSAS_ASSEMBLY"] = mode
    try:
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _ctx(prec, coords, tets=None, mat="svk", trusses=None, strain=0, E=1.0, A=None, fixed=(),
         faces=None, G=None, assembly=None):
    with _assembly_mode(assembly):
        return _ctx_build(prec, coords, tets, mat, trusses, strain, E, A, fixed, faces, G)


def _ctx_build(prec, coords, tets, mat, trusses, strain, E, A, fixed, faces, G):
    c = _capi.Context(PREC[prec])
    c.set_nodes(coords)
    if tets is not None:
        code, p = MATS[mat]
        c.add_tet_group(tets, code, p, G)
    if trusses is not None:
        c.add_truss_group(trusses, strain, E, A)
    if faces is not None:
        c.add_pressure_faces(faces)
    c.set_fixed_dofs(np.asarray(fixed, dtype=np.int64))
    c.finalize()
    return c


def _oracle_model(oracle, coords, tets, mat, fixed=(), G=None):
    code = {"svk": oracle.MAT_SVK, "nh": oracle.MAT_NEOHOOKEAN, "linear": oracle.MAT_LINEAR}[mat]
    return oracle.Model(coords, tets=tets, tet_mat=code, tet_p=MATS[mat][1], tet_G=G,
                        fixed_dofs=fixed)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("mat", ["svk", "nh", "linear"])
def test_tet_element_outputs(oracle, prec, mat):
    """Batch internal_forces on the GPU vs the oracle, element by element."""
    print(f"[gpu] tet element outputs {mat} {prec}", flush=True)
    sm = meshes.cube(2, jitter=0.1, seed=7)
    rng = np.random.default_rng(11)
    U = rng.normal(scale=0.04, size=3 * sm.nn)
    c = _ctx(prec, sm.coords, sm.tets, mat)
    c.set_U(U)
    f, K, s, e = c.element_forces(0)
    code, p = MATS[mat]
    for k, tet in enumerate(sm.tets):
        X = sm.coords[tet].reshape(-1)
        u = U.reshape(-1, 3)[tet].reshape(-1)
        if mat == "linear":
            fo, Ko, so, eo = oracle.tet_linear(p[0], p[1], X, u)
        else:
            fo, Ko, so, eo = oracle.tet_hyperelastic({"svk": oracle.MAT_SVK, "nh": oracle.MAT_NEOHOOKEAN}[mat],
                                                     p[0], p[1], X, u)
        np.testing.assert_allclose(f[k], fo, rtol=1e-11, atol=1e-11 * np.abs(fo).max())
        np.testing.assert_allclose(K[k], Ko, rtol=1e-11, atol=1e-11 * np.abs(Ko).max())
        np.testing.assert_allclose(s[k], so, rtol=1e-11, atol=1e-11 * np.abs(so).max())
        np.testing.assert_allclose(e[k], eo, rtol=1e-12, atol=1e-14)
    c.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("model", [0, 1])
def test_truss_element_outputs(oracle, model, prec):
    """Batch internal_forces of trusses vs the oracle; the element outputs are
    evaluated in FP64 in both precision modes."""
    print(f"[gpu] truss element outputs model={model} {prec}", flush=True)
    tl = meshes.truss_lattice(2, h=0.5, jitter=0.05)
    rng = np.random.default_rng(3)
    U = rng.normal(scale=0.01, size=3 * tl.nn)
    A = rng.uniform(0.5, 1.5, size=len(tl.trusses))
    c = _ctx(prec, tl.coords, trusses=tl.trusses, strain=model, E=2.0, A=A)
    c.set_U(U)
    f, K, s, e = c.element_forces(0)
    for k, (a, b) in enumerate(tl.trusses):
        d = np.r_[3 * a:3 * a + 3, 3 * b:3 * b + 3]
        fo, Ko, so, eo = oracle.truss(model, 2.0, A[k], tl.coords[[a, b]].reshape(-1), U[d])
        np.testing.assert_allclose(f[k], fo, rtol=1e-11, atol=1e-14)
        np.testing.assert_allclose(K[k], Ko, rtol=1e-10, atol=1e-12 * np.abs(Ko).max())
        assert s[k][0, 0] == pytest.approx(so, rel=1e-11, abs=1e-15)
        assert e[k][0, 0] == pytest.approx(eo, rel=1e-11, abs=1e-15)
    c.close()


def _bsr_vs_oracle(Kg, Ko, fixed, n, tol):
    """Compare the engine's tangent (identity on fixed rows/cols) with the
    oracle's full tangent restricted to free dofs."""
    Kg = Kg.toarray()
    Ko = Ko.toarray()
    free = np.setdiff1d(np.arange(n), fixed)
    scale = np.abs(Ko).max()
    np.testing.assert_allclose(Kg[np.ix_(free, free)], Ko[np.ix_(free, free)], rtol=0,
                               atol=tol * scale)
    np.testing.assert_array_equal(Kg[np.ix_(fixed, fixed)], np.eye(len(fixed)))
    assert np.abs(Kg[np.ix_(fixed, free)]).max(initial=0) == 0
    assert np.abs(Kg[np.ix_(free, fixed)]).max(initial=0) == 0


@pytest.mark.parametrize("assembly", ["gather", "staged", "scatter"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("mat", ["svk", "nh", "linear"])
def test_assembly_matches_oracle(oracle, prec, mat, assembly):
    """Gather (per-block contribution lists), staged gather (element states in
    LDS per workgroup of rows) and coloured-scatter assembly of f_int and K vs
    the oracle's COO->CSR assembly."""
    print(f"[gpu] assembly {mat} {prec} {assembly}", flush=True)
    s, sm = workloads.cube_extension(n=4, jitter=0.1, seed=3)
    fixed = s.fixed_dofs
    rng = np.random.default_rng(5)
    U = rng.normal(scale=0.02, size=3 * sm.nn)
    U[fixed] = 0
    c = _ctx(prec, sm.coords, sm.tets, mat, fixed=fixed, assembly=assembly)
    assert c.timings()["assembly_mode"] == 0  # no assembly yet
    c.set_U(U)
    fint = c.assemble(want_fint=True)
    Kg = c.tangent_bsr()
    out = _oracle_model(oracle, sm.coords, sm.tets, mat, fixed).assemble(U)
    np.testing.assert_allclose(fint, out["fint"], rtol=0, atol=1e-12 * np.abs(out["fint"]).max())
    _bsr_vs_oracle(Kg, out["K"], fixed, 3 * sm.nn, 1e-12 if prec == "fp64" else 2e-6)
    # y = K x through the C ABI
    x = rng.normal(size=3 * sm.nn)
    y = c.matvec(x)
    xf = x.copy()
    xf[fixed] = 0
    yo = out["K"] @ xf
    yo[fixed] = x[fixed]
    np.testing.assert_allclose(y, yo, rtol=0, atol=(1e-11 if prec == "fp64" else 3e-6) * np.abs(yo).max())
    assert c.timings()["assembly_mode"] == {"gather": 1, "staged": 2, "scatter": 0}[assembly]
    c.close()


@pytest.mark.parametrize("assembly", ["gather", "scatter"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_assembly_mixed_groups_vs_oracle(oracle, prec, assembly):
    """Tets and trusses in one model (two element groups sharing nodes), plus a
    fully fixed node that no element touches (identity block-Jacobi row)."""
    print(f"[gpu] assembly tets+trusses {prec} {assembly}", flush=True)
    sm = meshes.cube(3, jitter=0.1, seed=4)
    X = np.vstack([sm.coords, [[2.0, 2.0, 2.0]]])  # the isolated node
    nn = len(X)
    bars = np.array([[0, 5], [3, 17], [10, 40], [21, 63], [1, 2]], dtype=np.int32)
    A = np.array([1e-3, 2e-3, 1.5e-3, 1e-3, 3e-3])
    fixed = np.concatenate([np.arange(3 * 0, 3 * 0 + 3), [3 * 4 + 1], np.arange(3 * (nn - 1), 3 * nn)])
    rng = np.random.default_rng(8)
    U = rng.normal(scale=0.02, size=3 * nn)
    U[fixed] = 0
    with _assembly_mode(assembly):
        c = _capi.Context(PREC[prec])
        c.set_nodes(X)
        c.add_tet_group(sm.tets, _capi.MAT_SVK, MATS["svk"][1], None)
        c.add_truss_group(bars, 1, 2e7, A)
        c.set_fixed_dofs(fixed)
        c.finalize()
    c.set_U(U)
    fint = c.assemble(want_fint=True)
    Kg = c.tangent_bsr()
    m = oracle.Model(X, tets=sm.tets, tet_mat=oracle.MAT_SVK, tet_p=MATS["svk"][1], trusses=bars,
                     truss_strain=oracle.STRAIN_GREEN, truss_E=2e7, truss_A=A, fixed_dofs=fixed)
    out = m.assemble(U)
    np.testing.assert_allclose(fint, out["fint"], rtol=0, atol=1e-12 * np.abs(out["fint"]).max())
    # the isolated node has no block in the pattern: compare the rest
    m3 = 3 * (nn - 1)
    assert np.abs(Kg.toarray()[m3:]).max() == 0
    _bsr_vs_oracle(Kg.tocsr()[:m3, :m3], out["K"].tocsr()[:m3, :m3], fixed[fixed < m3], m3,
                   1e-12 if prec == "fp64" else 2e-6)
    # a solve through the block-Jacobi PCG touches the identity row of the isolated node
    b = rng.normal(size=3 * nn)
    x, it, rr = c.linear_solve(b)
    free = np.setdiff1d(np.arange(3 * nn), fixed)
    Kff = Kg.tocsr()[free][:, free]  # the engine's own (FP32 in mixed mode) tangent
    res = Kff @ x[free] - b[free]
    # FP32 PCG: the true residual can trail the recurrence by ~ eps32 cond(K) (cond ~ 7e3 here)
    assert np.linalg.norm(res) <= (1e-7 if prec == "fp64" else 5e-4) * np.linalg.norm(b[free])
    c.close()


@pytest.mark.parametrize("assembly", ["gather", "staged", "scatter"])
def test_assembly_is_deterministic(assembly):
    """Two assemblies of the same state give the same bits (no float atomics)."""
    print(f"[gpu] assembly determinism {assembly}", flush=True)
    sm = meshes.cube(8, jitter=0.1, seed=9)
    U = np.random.default_rng(2).normal(scale=0.01, size=3 * sm.nn)
    for prec in ("mixed", "fp64"):
        c = _ctx(prec, sm.coords, sm.tets, "nh", assembly=assembly)
        c.set_U(U)
        f1 = c.assemble(True)
        K1 = c.tangent_bsr().data.copy()
        f2 = c.assemble(True)
        K2 = c.tangent_bsr().data.copy()
        assert np.array_equal(f1, f2) and np.array_equal(K1, K2)
        c.close()


def test_pressure_loads_match_oracle(oracle):
    print("[gpu] pressure loads", flush=True)
    sm = meshes.cube(3, jitter=0.1, seed=1)
    faces = np.concatenate([sm.faces["x1"], sm.faces["y1"], sm.faces["z0"]])
    c = _ctx("fp64", sm.coords, sm.tets, "svk", faces=faces)
    nodal = np.random.default_rng(0).normal(size=3 * sm.nn)
    c.set_loads(nodal, [2.5])
    f = c.get_fext()
    ref = nodal.copy()
    for tri in faces:
        fe = oracle.face_pressure(sm.coords[tri].reshape(-1), 2.5)
        for a in range(3):
            ref[3 * tri[a]:3 * tri[a] + 3] += fe[a]
    np.testing.assert_allclose(f, ref, rtol=1e-14, atol=1e-15)
    # pressure-only loads stay queued (staged through pinned memory): two calls
    # back to back, the second one's values win
    c.set_loads(None, [7.0])
    c.set_loads(None, [2.5])
    np.testing.assert_allclose(c.get_fext(), ref - nodal, rtol=1e-13, atol=1e-14)
    c.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_pcg_solve_matches_direct(oracle, prec):
    """Block-Jacobi PCG on the assembled tangent vs a sparse direct solve."""
    import scipy.sparse.linalg as spla
    print(f"[gpu] pcg vs direct {prec}", flush=True)
    s, sm = workloads.cube_extension(n=6, jitter=0.1, seed=4)
    fixed = s.fixed_dofs
    c = _ctx(prec, sm.coords, sm.tets, "svk", fixed=fixed)
    c.set_linear_solver(_capi.LINSOLVE_PCG, rtol=1e-12 if prec == "fp64" else 1e-6, maxiter=5000)
    c.assemble()
    b = np.random.default_rng(1).normal(size=3 * sm.nn)
    x, it, rr = c.linear_solve(b)
    out = _oracle_model(oracle, sm.coords, sm.tets, "svk", fixed).assemble(np.zeros(3 * sm.nn),
                                                                         want_full=False,
                                                                         want_ff=True)
    free = np.setdiff1d(np.arange(3 * sm.nn), fixed)
    xo = spla.spsolve(out["Kff"].tocsc(), b[free])
    err = np.abs(x[free] - xo).max() / np.abs(xo).max()
    assert err < (1e-9 if prec == "fp64" else 2e-4), err
    assert np.all(x[fixed] == 0)
    assert 0 < it < 5000
    c.close()


def _newton(structure, nsteps, prec, tol_f=1e-10, **kw):
    sa = NonLinearStaticAnalysis(structure, NSTEPS=nsteps)
    alg = NewtonRaphson(ConvergenceSettings(rel_U_tol=1e-14, rel_res_force_tol=tol_f, max_iter=30))
    be = MI355XBackend(precision=prec, maxiter=20000, **kw)
    return solve(sa, alg, be)


def test_metrics_jsonl_streamed(tmp_path):
    """MI355XBackend(metrics_jsonl=...) writes one record per load step while
    the solve runs; the records equal Solution.metrics() (U's norm aside, which
    the streamed records leave null) and their inner iterations add up."""
    import json
    print("[gpu] metrics jsonl", flush=True)
    s, sm = workloads.cube_extension(n=4)
    path = tmp_path / "steps.jsonl"
    sol = _newton(s, 4, "mixed", preconditioner="amg", forcing=True, metrics_jsonl=str(path))
    lines = [json.loads(x) for x in path.read_text().splitlines()]
    m = sol.metrics()
    assert len(lines) == len(m) == 4
    for a, b in zip(lines, m):
        assert a["U_norm"] is None and b["U_norm"] > 0
        assert {k: v for k, v in a.items() if k != "U_norm"} == \
            {k: v for k, v in b.items() if k != "U_norm"}
        assert b["newton_iterations"] == len(b["lin_iters_per_iteration"]) or \
            b["newton_iterations"] == len(b["lin_iters_per_iteration"]) + 1
        assert b["dr_rel"] <= 1e-10 and b["wall_s"] > 0
    assert sum(b["lin_iters"] for b in m) == sol.linear_iterations()
    assert [b["step"] for b in m] == [1, 2, 3, 4]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_uniaxial_extension_cube(prec):
    """SVK cube 4^3 x 6 tets, tension 3e5 Pa, 10 steps, tol 1e-10: every load
    step is the homogeneous closed-form state (uniaxial_extension.jl:151-205)."""
    print(f"[gpu] uniaxial extension {prec}", flush=True)
    s, sm = workloads.cube_extension(n=4)
    sol = _newton(s, 10, prec)
    for k, t in enumerate(np.linspace(0.1, 1.0, 10)):
        a = workloads.svk_uniaxial_alpha(1e6, 3e5 * t)
        b = workloads.svk_uniaxial_beta(a, 0.3)
        ex = workloads.homogeneous_displacement(sm.coords, a, b)
        np.testing.assert_allclose(sol.U[k], ex, rtol=0, atol=1e-9 * np.abs(ex).max())
        assert sol.records[k]["dr_rel"] <= 1e-10
        assert type(sol.records[k]["criterion"]).__name__ == "ResidualForceCriterion"
    # stress at the converged state: P11 = p, C = F'F
    P = sol.stress[-1][0]
    a = workloads.svk_uniaxial_alpha(1e6, 3e5)
    np.testing.assert_allclose(P[:, 0, 0], 3e5, rtol=1e-6)
    np.testing.assert_allclose(sol.strain[-1][0][:, 0, 0], a * a, rtol=1e-6)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_nh_compression_vs_oracle(oracle, prec):
    """NeoHookean jittered cube in compression: per-step displacements equal
    the oracle's Newton (reference algorithm) and the closed form."""
    print(f"[gpu] NH compression vs oracle {prec}", flush=True)
    s, sm, p = workloads.cube_compression(n=4, jitter=0.05, seed=2)
    nsteps = 5
    sol = _newton(s, nsteps, prec)
    lfs = np.linspace(1 / nsteps, 1.0, nsteps)
    m = oracle.Model(sm.coords, tets=sm.tets, tet_mat=oracle.MAT_NEOHOOKEAN, tet_p=(8.3e5, 3.8e5),
                     fixed_dofs=s.fixed_dofs)
    _, Us, _ = m.newton(fext_steps(oracle, s, lfs), rel_U_tol=1e-14, rel_res_tol=1e-10, max_iter=30)
    for k in range(nsteps):
        scale = np.abs(Us[k]).max()
        np.testing.assert_allclose(sol.U[k], Us[k], rtol=0, atol=1e-8 * scale)
        a = workloads.nh_uniaxial_alpha(-p * lfs[k], 8.3e5, 3.8e5)
        b = workloads.nh_uniaxial_beta(a, 8.3e5, 3.8e5)
        ex = workloads.homogeneous_displacement(sm.coords, a, b)
        np.testing.assert_allclose(sol.U[k], ex, rtol=0, atol=1e-8 * scale)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_forcing_converges_to_same_state(prec):
    """Inexact-Newton forcing changes only the inner tolerances: same residual
    criterion, same converged displacements, fewer inner iterations."""
    print(f"[gpu] forcing vs fixed {prec}", flush=True)
    s, sm, p = workloads.cube_compression(n=8, jitter=0.05, seed=2)
    fixed = _newton(s, 5, prec)
    forced = _newton(s, 5, prec, forcing=True)
    for k in range(5):
        scale = np.abs(fixed.U[k]).max()
        np.testing.assert_allclose(forced.U[k], fixed.U[k], rtol=0, atol=1e-8 * scale)
        assert forced.records[k]["dr_rel"] <= 1e-10
    assert forced.linear_iterations() < fixed.linear_iterations()
    floor = 1e-6 if prec == "mixed" else 1.4e-8
    for d in forced.newton_log:
        # eta_max = 0.1 caps the tolerance; only the solve whose residual already
        # meets the outer tolerance (r_rel <= 1e-10, kept by isconverged's
        # residual-first order) may be loose up to 0.5
        cap = 0.5 if d["r_rel"] <= 1e-10 else 0.1
        assert floor * (1 - 1e-12) <= d["lin_rtol"] <= cap, d


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_gmres_matches_direct(oracle, prec):
    """Restarted GMRES (block-Jacobi right preconditioning) on the assembled
    tangent vs a sparse direct solve, with restarts exercised (m = 8)."""
    import scipy.sparse.linalg as spla
    print(f"[gpu] gmres vs direct {prec}", flush=True)
    s, sm = workloads.cube_extension(n=4, jitter=0.1, seed=4)
    fixed = s.fixed_dofs
    c = _ctx(prec, sm.coords, sm.tets, "svk", fixed=fixed)
    c.set_linear_solver(_capi.LINSOLVE_GMRES, rtol=1e-10 if prec == "fp64" else 1e-6,
                        maxiter=3000, gmres_restart=8)
    c.assemble()
    b = np.random.default_rng(7).normal(size=3 * sm.nn)
    x, it, rr = c.linear_solve(b)
    Kff = _oracle_model(oracle, sm.coords, sm.tets, "svk", fixed).assemble(
        np.zeros(3 * sm.nn), want_full=False, want_ff=True)["Kff"]
    free = np.setdiff1d(np.arange(3 * sm.nn), fixed)
    xo = spla.spsolve(Kff.tocsc(), b[free])
    err = np.abs(x[free] - xo).max() / np.abs(xo).max()
    assert err < (1e-8 if prec == "fp64" else 1e-3), err
    assert it > 8 and np.all(x[fixed] == 0)
    c.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_gmres_fallback_on_indefinite_tangent(prec):
    """AUTO: PCG breaks down on an indefinite tangent (a compressed bar chain
    has negative transverse stiffness) and GMRES solves the system."""
    print(f"[gpu] gmres fallback {prec}", flush=True)
    X = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0.0], [3, 0.1, 0.2]])
    c = _ctx(prec, X, trusses=[[0, 1], [1, 2], [2, 3]], E=1.0, A=[1.0, 1.0, 1.0],
             fixed=[0, 1, 2, 9, 10, 11])
    c.set_linear_solver(_capi.LINSOLVE_AUTO, rtol=1e-10 if prec == "fp64" else 1e-6, maxiter=200)
    U = np.zeros(12)
    U[6] = -0.2
    U[9] = -0.3
    c.set_U(U)
    c.assemble()
    b = np.array([0, 0, 0, 0.1, 1.0, -0.3, 0.2, 0.5, 0.4, 0, 0, 0.0])
    x, it, rr = c.linear_solve(b)
    K = c.tangent_bsr().toarray()
    free = [3, 4, 5, 6, 7, 8]
    assert np.linalg.eigvalsh(K[np.ix_(free, free)]).min() < 0  # genuinely indefinite
    res = K[np.ix_(free, free)] @ x[free] - b[free]
    assert np.linalg.norm(res) <= (1e-9 if prec == "fp64" else 1e-5) * np.linalg.norm(b[free])
    c.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_newton_with_gmres_only(prec):
    print(f"[gpu] newton with gmres {prec}", flush=True)
    s, sm = workloads.cube_extension(n=3)
    sol = _newton(s, 3, prec, solver="gmres", gmres_restart=20)
    for k, t in enumerate(np.linspace(1 / 3, 1.0, 3)):
        a = workloads.svk_uniaxial_alpha(1e6, 3e5 * t)
        b = workloads.svk_uniaxial_beta(a, 0.3)
        ex = workloads.homogeneous_displacement(sm.coords, a, b)
        np.testing.assert_allclose(sol.U[k], ex, rtol=0, atol=1e-8 * np.abs(ex).max())
    assert all(d["lin_solver"] == 2 for d in sol.newton_log)


def test_trusses_green_clamped(golden):
    """Clamped Green-strain bar (examples/clamped_truss) through the API on the GPU."""
    print("[gpu] clamped truss", flush=True)
    from onsas_amd import (FixedField, GlobalLoad, Mesh, NodeSet, SVK, Square, StructuralBoundaryCondition,
                           StructuralMaterial, Structure, Trusses, GreenStrain)
    g = golden["clamped_truss"]
    N, E, L, A, F = g["N"], g["E"], g["L"], g["A"], g["F"]
    x = np.linspace(0, L, N + 1)
    coords = np.stack([x, 0 * x, 0 * x], 1)
    bars = Trusses(np.stack([np.arange(N), np.arange(1, N + 1)], 1), Square(np.sqrt(A)), GreenStrain)
    s = Structure(Mesh(coords, [bars]), StructuralMaterial({SVK(E=E, nu=g["nu"]): bars}),
                  StructuralBoundaryCondition({
                      FixedField("u", [1]): [NodeSet([0])],
                      FixedField("u", [2, 3]): [NodeSet(np.arange(N + 1))],
                      GlobalLoad("u", lambda t: [F * t, 0, 0]): [NodeSet([N])]}))
    sa = NonLinearStaticAnalysis(s, NSTEPS=g["NSTEPS"])
    for prec in ("fp64", "mixed"):
        sol = solve(sa, NewtonRaphson(), MI355XBackend(precision=prec))
        u = np.array([U[3 * N] for U in sol.U])
        eg = 0.5 * ((L + u) ** 2 - L ** 2) / L ** 2
        P = (L + u) / L * E * eg * A
        np.testing.assert_allclose(P, F * sa.load_factors, rtol=g["rtol"])


def test_error_paths():
    """Inverted tet -> ERR_NEGATIVE_VOLUME; call order -> ERR_STATE; injected
    error -> its code; an indefinite tangent breaks PCG down (ERR_LINEAR_SOLVE)."""
    print("[gpu] error paths", flush=True)
    X = np.array([[0, 0, 0], [0, 1, 0], [0, 0, 1], [2, 0, 1.0]])
    c = _capi.Context(_capi.PRECISION_FP64)
    c.set_nodes(X)
    c.add_tet_group([[1, 0, 2, 3]], _capi.MAT_SVK, (1.0, 1.0))
    with pytest.raises(_capi.OnsasError) as e:
        c.finalize()
    assert e.value.code == _capi.ERR_NEGATIVE_VOLUME
    c.close()
    c = _ctx("fp64", X, [[0, 1, 2, 3]], "svk", fixed=[0, 1, 2, 3, 4, 5, 6, 7, 8])
    with pytest.raises(_capi.OnsasError) as e:
        c.newton_step()
    assert e.value.code == _capi.ERR_STATE
    c.inject_error(_capi.ERR_INJECTED)
    with pytest.raises(_capi.OnsasError) as e:
        c.assemble()
    assert e.value.code == _capi.ERR_INJECTED
    c.assemble()
    c.close()
    # a straight two-bar chain compressed axially: negative transverse stiffness
    X = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0.0]])
    c = _ctx("fp64", X, trusses=[[0, 1], [1, 2]], E=1.0, A=[1.0, 1.0], fixed=[0, 1, 2, 6, 7, 8])
    c.set_linear_solver(_capi.LINSOLVE_PCG, rtol=1e-10, maxiter=100)
    c.set_U(np.array([0, 0, 0, 0.0, 0, 0, -0.2, 0, 0]))
    c.assemble()
    with pytest.raises(_capi.OnsasError) as e:
        c.linear_solve(np.array([0, 0, 0, 0, 1.0, 0, 0, 0, 0]))
    assert e.value.code == _capi.ERR_LINEAR_SOLVE
    c.close()


def test_full_size_compression_cube_mixed():
    """The 64^3 NeoHookean compression cube (1,572,864 tets) in mixed precision:
    the converged state is homogeneous, so every node must match the closed
    form to 1e-6 relative at full load (size-independent property)."""
    print("[gpu] 64^3 compression cube, mixed, 20 steps", flush=True)
    s, sm, p = workloads.cube_compression(n=64)
    sol = _newton(s, 20, "mixed", tol_f=1e-8, store_stresses=False)
    a = workloads.nh_uniaxial_alpha(-p, 8.3e5, 3.8e5)
    assert a == pytest.approx(0.85, abs=1e-9)
    b = workloads.nh_uniaxial_beta(a, 8.3e5, 3.8e5)
    ex = workloads.homogeneous_displacement(sm.coords, a, b)
    err = np.abs(sol.U[-1] - ex).max() / np.abs(ex).max()
    print(f"[gpu] 64^3 max rel error {err:.3e}, wall {sol.wall_s:.2f}s, "
          f"lin iters {sol.linear_iterations()}", flush=True)
    assert err <= 1e-6
    assert all(r["dr_rel"] <= 1e-8 for r in sol.records)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_amg_pcg_matches_direct(oracle, prec):
    """PCG with the smoothed-aggregation V-cycle vs a sparse direct solve, and
    far fewer iterations than block-Jacobi on the same system."""
    import scipy.sparse.linalg as spla
    print(f"[gpu] AMG PCG vs direct {prec}", flush=True)
    s, sm, p = workloads.cube_compression(n=8, jitter=0.05, seed=2)
    fixed = s.fixed_dofs
    a = workloads.nh_uniaxial_alpha(-p * 0.5, 8.3e5, 3.8e5)
    b = workloads.nh_uniaxial_beta(a, 8.3e5, 3.8e5)
    U = workloads.homogeneous_displacement(sm.coords, a, b)
    rtol = 1e-10 if prec == "fp64" else 1e-6
    rhs = np.random.default_rng(3).normal(size=3 * sm.nn)
    its = {}
    for pc in (_capi.PRECOND_BLOCK_JACOBI, _capi.PRECOND_AMG):
        c = _ctx(prec, sm.coords, sm.tets, "nh", fixed=fixed)
        c.set_linear_solver(_capi.LINSOLVE_PCG, rtol=rtol, maxiter=5000, precond=pc)
        c.set_U(U)
        c.assemble()
        x, it, rr = c.linear_solve(rhs)
        its[pc] = it
        if pc == _capi.PRECOND_AMG:
            Kff = _oracle_model(oracle, sm.coords, sm.tets, "nh", fixed).assemble(
                U, want_full=False, want_ff=True)["Kff"]
            free = np.setdiff1d(np.arange(3 * sm.nn), fixed)
            xo = spla.spsolve(Kff.tocsc(), rhs[free])
            err = np.abs(x[free] - xo).max() / np.abs(xo).max()
            assert err < (1e-8 if prec == "fp64" else 2e-4), err
            assert np.all(x[fixed] == 0)
            # a second solve reuses the hierarchy; a new assembly refreshes it
            x2, it2, _ = c.linear_solve(rhs)
            assert np.array_equal(x, x2) and it2 == it
        c.close()
    print(f"  iterations block-Jacobi {its[0]} AMG {its[1]}", flush=True)
    assert its[_capi.PRECOND_AMG] < 0.2 * its[_capi.PRECOND_BLOCK_JACOBI]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_amg_large_hierarchy(prec):
    """40^3 cube (68,921 nodes): four levels, the sliced-ELL coarse operator, the
    four-wave restriction and a dense coarsest level of a few hundred dofs.  The
    solution's true residual against the engine's own tangent meets the
    tolerance, in few iterations."""
    print(f"[gpu] AMG 40^3 hierarchy {prec}", flush=True)
    s, sm, p = workloads.cube_compression(n=40, jitter=0.02, seed=2)
    fixed = np.asarray(s.fixed_dofs)
    a = workloads.nh_uniaxial_alpha(-p * 0.5, 8.3e5, 3.8e5)
    b = workloads.nh_uniaxial_beta(a, 8.3e5, 3.8e5)
    U = workloads.homogeneous_displacement(sm.coords, a, b)
    rtol = 1e-9 if prec == "fp64" else 1e-6
    c = _capi.Context(PREC[prec])
    c.set_nodes(sm.coords)
    c.add_tet_group(sm.tets, _capi.MAT_NEOHOOKEAN, (8.3e5, 3.8e5), None)
    c.set_fixed_dofs(fixed.astype(np.int64))
    c.finalize()
    c.set_linear_solver(_capi.LINSOLVE_PCG, rtol=rtol, maxiter=500, precond=_capi.PRECOND_AMG)
    c.set_U(U)
    c.assemble()
    rhs = np.random.default_rng(4).normal(size=3 * sm.nn)
    rhs[fixed] = 0
    x, it, _ = c.linear_solve(rhs)
    K = c.tangent_bsr()
    res = np.linalg.norm(rhs - K @ x) / np.linalg.norm(rhs)
    print(f"  iterations {it}, true relative residual {res:.2e}", flush=True)
    # mixed: the FP32 recurrence residual meets rtol; the true residual, evaluated
    # in FP64 arithmetic, adds the FP32 rounding of the products (a few eps32
    # times the conditioning), measured 3.7e-6 here
    assert res <= (3 * rtol if prec == "fp64" else 2e-5)
    assert it <= 60
    c.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_amg_hierarchy_galerkin(prec):
    """24^3 compression cube: the exported three-level hierarchy satisfies the
    Galerkin identity A_{l+1} = P_l^T A_l P_l on both coarse levels (level 0 is
    the tangent, P_l the stored prolongator values), and PCG reaches 1e-6 on
    the load vector in few iterations.  The iteration bound guards the coarse
    near-nullspace: rigid-body rotations taken about a different point in each
    aggregate span the same first coarse space but describe no rigid motion on
    the second level, which costs ~35% more iterations here (22 instead of 16)."""
    print(f"[gpu] AMG 24^3 Galerkin hierarchy {prec}", flush=True)
    s, sm, p = workloads.cube_compression(n=24)
    fixed = np.asarray(s.fixed_dofs)
    a = workloads.nh_uniaxial_alpha(-p * 0.7, 8.3e5, 3.8e5)
    b = workloads.nh_uniaxial_beta(a, 8.3e5, 3.8e5)
    U = workloads.homogeneous_displacement(sm.coords, a, b)
    c = _capi.Context(PREC[prec])
    c.set_nodes(sm.coords)
    c.add_tet_group(sm.tets, _capi.MAT_NEOHOOKEAN, (8.3e5, 3.8e5), None)
    c.add_pressure_faces(sm.faces["x1"])
    c.set_fixed_dofs(fixed.astype(np.int64))
    c.finalize()
    c.set_linear_solver(_capi.LINSOLVE_PCG, rtol=1e-6, maxiter=500, precond=_capi.PRECOND_AMG)
    assert c.amg_levels() == 0
    with pytest.raises(_capi.OnsasError):  # no hierarchy to time yet
        c.time_cycle_products(10)
    c.set_U(U)
    c.set_loads(None, [0.05 * p])
    c.assemble()
    f = c.get_fext()
    f[fixed] = 0
    x, it, _ = c.linear_solve(f)
    print(f"  iterations {it}", flush=True)
    assert it <= 19
    assert c.amg_levels() == 3
    # the timing batch uses only the cycle's work vectors: the next solve is unchanged
    ms = c.time_cycle_products(20)
    print(f"  fine-level cycle product {ms * 1e3:.1f} us", flush=True)
    assert 0 < ms < 100
    with pytest.raises(_capi.OnsasError):
        c.time_cycle_products(1)
    x2, it2, _ = c.linear_solve(f)
    assert it2 == it and np.array_equal(x2, x)
    info = [c.amg_level_info(l) for l in range(3)]
    assert [i["bs"] for i in info] == [3, 6, 6]
    assert info[0]["n"] == sm.nn and info[2]["nnz_p"] == 0
    assert all(i["lam"] > 0 for i in info[:2])
    # the stored operators: FP64 products over the stored P; W = A P and A_c are
    # kept in FP64 (fp64 mode) or FP32 (mixed), whose rounding bounds the gap
    tol = 1e-12 if prec == "fp64" else 1e-6
    A = c.tangent_bsr().tocsr()
    for l in range(2):
        P = c.amg_prolongator(l).tocsr()
        Ac = c.amg_operator(l + 1).tocsr()
        assert P.shape == (A.shape[0], Ac.shape[0])
        G = (P.T @ A @ P).tocsr()
        gap = abs(G - Ac).max() / abs(G).max()
        print(f"  level {l + 1}: {Ac.shape[0]} dofs, Galerkin gap {gap:.2e}", flush=True)
        assert gap <= tol
        A = Ac
    c.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_amg_edge_cases(prec):
    """Multigrid on models the cube tests do not reach: tets + trusses with a
    fully fixed node that no element touches (no block of its own in the
    pattern), and a one-cell cube (8 nodes: the first coarse level is already
    the dense one).  The solves meet the same residual bounds as block-Jacobi."""
    print(f"[gpu] AMG edge cases {prec}", flush=True)
    rng = np.random.default_rng(8)
    sm = meshes.cube(3, jitter=0.1, seed=4)
    X = np.vstack([sm.coords, [[2.0, 2.0, 2.0]]])  # the isolated node
    nn = len(X)
    bars = np.array([[0, 5], [3, 17], [10, 40], [21, 63], [1, 2]], dtype=np.int32)
    A = np.array([1e-3, 2e-3, 1.5e-3, 1e-3, 3e-3])
    fixed = np.concatenate([np.arange(3), [3 * 4 + 1], np.arange(3 * (nn - 1), 3 * nn)])
    c = _capi.Context(PREC[prec])
    c.set_nodes(X)
    c.add_tet_group(sm.tets, _capi.MAT_SVK, MATS["svk"][1], None)
    c.add_truss_group(bars, 1, 2e7, A)
    c.set_fixed_dofs(fixed)
    c.finalize()
    c.set_linear_solver(_capi.LINSOLVE_PCG, rtol=1e-10 if prec == "fp64" else 1e-6, maxiter=500,
                        precond=_capi.PRECOND_AMG)
    U = rng.normal(scale=0.02, size=3 * nn)
    U[fixed] = 0
    c.set_U(U)
    c.assemble()
    b = rng.normal(size=3 * nn)
    x, it, rr = c.linear_solve(b)
    free = np.setdiff1d(np.arange(3 * nn), fixed)
    Kff = c.tangent_bsr().tocsr()[free][:, free]
    res = np.linalg.norm(Kff @ x[free] - b[free]) / np.linalg.norm(b[free])
    print(f"  tets + trusses + isolated node: {it} iterations, residual {res:.1e}", flush=True)
    assert res <= (1e-7 if prec == "fp64" else 5e-4)
    assert np.all(x[fixed] == 0)
    c.close()
    s, sm1 = workloads.cube_extension(n=1)
    c = _ctx(prec, sm1.coords, sm1.tets, "svk", fixed=s.fixed_dofs)
    c.set_linear_solver(_capi.LINSOLVE_PCG, rtol=1e-10 if prec == "fp64" else 1e-6, maxiter=100,
                        precond=_capi.PRECOND_AMG)
    c.set_U(np.zeros(3 * sm1.nn))
    c.assemble()
    b = rng.normal(size=3 * sm1.nn)
    x, it, rr = c.linear_solve(b)
    free = np.setdiff1d(np.arange(3 * sm1.nn), s.fixed_dofs)
    Kff = c.tangent_bsr().tocsr()[free][:, free]
    res = np.linalg.norm(Kff @ x[free] - b[free]) / np.linalg.norm(b[free])
    print(f"  one-cell cube: {it} iterations, residual {res:.1e}", flush=True)
    assert res <= (1e-9 if prec == "fp64" else 1e-5) and it <= 20  # two levels, 6 coarse modes
    c.close()


@pytest.mark.parametrize("refresh", ["step", "assembly", "once"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_newton_with_amg(prec, refresh):
    """Compression cube with the AMG preconditioner, its coarse levels rebuilt
    once per load step, after every assembly or only once: the closed-form
    state."""
    print(f"[gpu] Newton with AMG {prec} refresh {refresh}", flush=True)
    s, sm, p = workloads.cube_compression(n=6, jitter=0.05, seed=2)
    nsteps = 4
    sol = _newton(s, nsteps, prec, preconditioner="amg", amg_refresh=refresh)
    for k, t in enumerate(np.linspace(1 / nsteps, 1.0, nsteps)):
        a = workloads.nh_uniaxial_alpha(-p * t, 8.3e5, 3.8e5)
        b = workloads.nh_uniaxial_beta(a, 8.3e5, 3.8e5)
        ex = workloads.homogeneous_displacement(sm.coords, a, b)
        np.testing.assert_allclose(sol.U[k], ex, rtol=0, atol=1e-8 * np.abs(ex).max())
        assert sol.records[k]["dr_rel"] <= 1e-10



@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_amg_pcg_maxiter_exact(prec):
    """The multigrid PCG stops exactly where the reference's cg! would: with
    maxiter = N (N the iterations it needs) it converges in N iterations to the
    same x, and with N - 1 it reports no convergence.  The check runs right
    after each update and again at the next product, and N falls at different
    places of the 2 + 8 + 8 ... launch chunks for the three tolerances."""
    print(f"[gpu] AMG PCG maxiter {prec}", flush=True)
    s, sm, p = workloads.cube_compression(n=10, jitter=0.05, seed=2)
    fixed = np.asarray(s.fixed_dofs)
    a = workloads.nh_uniaxial_alpha(-p * 0.5, 8.3e5, 3.8e5)
    b = workloads.nh_uniaxial_beta(a, 8.3e5, 3.8e5)
    U = workloads.homogeneous_displacement(sm.coords, a, b)
    c = _ctx(prec, sm.coords, sm.tets, "nh", fixed=fixed)
    c.set_U(U)
    c.assemble()
    rhs = np.random.default_rng(7).normal(size=3 * sm.nn)
    seen = set()
    for rtol in (1e-2, 1e-4, 1e-6 if prec == "mixed" else 1e-9):
        c.set_linear_solver(_capi.LINSOLVE_PCG, rtol=rtol, maxiter=500, precond=_capi.PRECOND_AMG)
        x, n, _ = c.linear_solve(rhs)
        seen.add(n)
        c.set_linear_solver(_capi.LINSOLVE_PCG, rtol=rtol, maxiter=n, precond=_capi.PRECOND_AMG)
        x2, n2, _ = c.linear_solve(rhs)
        assert n2 == n and np.array_equal(x, x2)
        if n > 1:
            c.set_linear_solver(_capi.LINSOLVE_PCG, rtol=rtol, maxiter=n - 1,
                                precond=_capi.PRECOND_AMG)
            with pytest.raises(_capi.OnsasError):
                c.linear_solve(rhs)
    print(f"  iteration counts {sorted(seen)}", flush=True)
    assert len(seen) >= 2
    c.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_amg_fine_degree_two(oracle, prec, monkeypatch):
    """The degree-2 Chebyshev fine smoother (ONSAS_AMG_FINE_DEGREE=2, read when
    the hierarchy is built) solves the same system as the default degree-1
    step, to the same tolerance, in no more iterations."""
    import scipy.sparse.linalg as spla
    print(f"[gpu] AMG fine degree 2 {prec}", flush=True)
    s, sm, p = workloads.cube_compression(n=8, jitter=0.05, seed=2)
    fixed = s.fixed_dofs
    a = workloads.nh_uniaxial_alpha(-p * 0.5, 8.3e5, 3.8e5)
    b = workloads.nh_uniaxial_beta(a, 8.3e5, 3.8e5)
    U = workloads.homogeneous_displacement(sm.coords, a, b)
    rtol = 1e-10 if prec == "fp64" else 1e-6
    rhs = np.random.default_rng(3).normal(size=3 * sm.nn)
    Kff = _oracle_model(oracle, sm.coords, sm.tets, "nh", fixed).assemble(
        U, want_full=False, want_ff=True)["Kff"]
    free = np.setdiff1d(np.arange(3 * sm.nn), fixed)
    xo = spla.spsolve(Kff.tocsc(), rhs[free])
    its = {}
    for deg in ("1", "2"):
        monkeypatch.setenv("ONSAS_AMG_FINE_DEGREE", deg)
        c = _ctx(prec, sm.coords, sm.tets, "nh", fixed=fixed)
        c.set_linear_solver(_capi.LINSOLVE_PCG, rtol=rtol, maxiter=500, precond=_capi.PRECOND_AMG)
        c.set_U(U)
        c.assemble()
        x, its[deg], _ = c.linear_solve(rhs)
        err = np.abs(x[free] - xo).max() / np.abs(xo).max()
        assert err < (1e-8 if prec == "fp64" else 2e-4), (deg, err)
        c.close()
    print(f"  iterations degree 1 {its['1']}, degree 2 {its['2']}", flush=True)
    assert its["2"] <= its["1"]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_amg_coarse_degree_one(prec, monkeypatch):
    """The damped-Jacobi smoothing on the coarser levels (ONSAS_AMG_COARSE_DEGREE
    =1, read when the hierarchy is built) on the 40^3 cube, whose level 1 has
    the sliced-ELL operator: the degree-1 step reads its input at neighbouring
    rows and writes out of place.  Both degrees meet the tolerance in the true
    residual against the engine's tangent; degree 2 needs no more iterations."""
    print(f"[gpu] AMG coarse degree 1 {prec}", flush=True)
    s, sm, p = workloads.cube_compression(n=40, jitter=0.02, seed=2)
    fixed = np.asarray(s.fixed_dofs)
    a = workloads.nh_uniaxial_alpha(-p * 0.5, 8.3e5, 3.8e5)
    b = workloads.nh_uniaxial_beta(a, 8.3e5, 3.8e5)
    U = workloads.homogeneous_displacement(sm.coords, a, b)
    rtol = 1e-9 if prec == "fp64" else 1e-6
    rhs = np.random.default_rng(4).normal(size=3 * sm.nn)
    rhs[fixed] = 0
    its = {}
    for deg in ("1", "2"):
        monkeypatch.setenv("ONSAS_AMG_COARSE_DEGREE", deg)
        c = _ctx(prec, sm.coords, sm.tets, "nh", fixed=fixed)
        c.set_linear_solver(_capi.LINSOLVE_PCG, rtol=rtol, maxiter=500, precond=_capi.PRECOND_AMG)
        c.set_U(U)
        c.assemble()
        x, its[deg], _ = c.linear_solve(rhs)
        K = c.tangent_bsr()
        res = np.linalg.norm(rhs - K @ x) / np.linalg.norm(rhs)
        assert res <= (3 * rtol if prec == "fp64" else 2e-5), (deg, res)
        assert c.amg_levels() >= 3
        c.close()
    print(f"  iterations degree 1 {its['1']}, degree 2 {its['2']}", flush=True)
    assert its["2"] <= its["1"] <= 80


@pytest.mark.parametrize("pc", ["block-jacobi", "amg"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_warm_start_matches_cold(prec, pc, monkeypatch):
    """Warm-started Newton solves (each PCG starts from the previous load
    step's update of the same Newton iteration, scaled to minimise the initial
    residual) take the same Newton iterations to the same states as zero
    starts (ONSAS_WARM_START=0), with fewer PCG iterations."""
    print(f"[gpu] warm vs cold start {prec} {pc}", flush=True)
    s, sm, p = workloads.cube_compression(n=6, jitter=0.05, seed=2)
    runs = {}
    for w in ("0", "1"):
        monkeypatch.setenv("ONSAS_WARM_START", w)
        runs[w] = _newton(s, 6, prec, tol_f=1e-9, preconditioner=pc, forcing=True)
    cold, warm = runs["0"], runs["1"]
    assert warm.iterations() == cold.iterations()
    for k in range(6):
        np.testing.assert_allclose(warm.U[k], cold.U[k], rtol=0,
                                   atol=1e-8 * np.abs(cold.U[k]).max())
        assert warm.records[k]["dr_rel"] <= 1e-9
    print(f"  PCG iterations cold {cold.linear_iterations()} warm {warm.linear_iterations()}",
          flush=True)
    assert warm.linear_iterations() < cold.linear_iterations()


@pytest.mark.parametrize("pc,chunk", [("block-jacobi", None), ("amg", None), ("amg", 6)])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_pcg_predicted_launches_bit_identical(prec, pc, chunk, monkeypatch):
    """Launching the previous load step's PCG iteration count before the first
    host check (graphs of chunk, 4 and 2 iterations) changes only how many
    iterations are queued past convergence, which exit on the device: every
    state and every PCG count equals that of ONSAS_PCG_PREDICT=0 bit for bit."""
    print(f"[gpu] predicted PCG launches {prec} {pc} chunk {chunk}", flush=True)
    s, sm, p = workloads.cube_compression(n=6, jitter=0.05, seed=2)
    runs = {}
    for w in ("0", "1"):
        monkeypatch.setenv("ONSAS_PCG_PREDICT", w)
        runs[w] = _newton(s, 6, prec, tol_f=1e-9, preconditioner=pc, forcing=True,
                          check_every=chunk)
    off, on = runs["0"], runs["1"]
    assert [d["lin_iters"] for d in on.newton_log] == [d["lin_iters"] for d in off.newton_log]
    for k in range(6):
        assert np.array_equal(on.U[k], off.U[k])
        assert on.records[k]["dr_rel"] == off.records[k]["dr_rel"]


def test_u_snapshots():
    """onsas_snapshot_U / onsas_read_snapshot: two slots, read on another
    thread after later work was enqueued, equal to the U of the moment each
    snapshot was taken (the per-step copy the Newton driver overlaps)."""
    from concurrent.futures import ThreadPoolExecutor
    print("[gpu] U snapshots", flush=True)
    sm = meshes.cube(3, jitter=0.1, seed=4)
    c = _ctx("mixed", sm.coords, sm.tets, "nh")
    rng = np.random.default_rng(12)
    U0, U1 = rng.normal(size=3 * sm.nn), rng.normal(size=3 * sm.nn)
    c.set_U(U0)
    c.snapshot_U(0)
    c.set_U(U1)
    c.snapshot_U(1)
    c.assemble()  # more work on the stream after both snapshots
    with ThreadPoolExecutor(max_workers=1) as pool:
        got0 = pool.submit(c.read_snapshot, 0, np.empty(3 * sm.nn)).result()
        got1 = pool.submit(c.read_snapshot, 1, np.empty(3 * sm.nn)).result()
    assert np.array_equal(got0, U0) and np.array_equal(got1, U1)
    assert np.array_equal(c.get_U(), U1)
    # a slot taken again after it was read holds the U of the new moment, not
    # the U written after it
    c.set_U(U0 + 1.0)
    c.snapshot_U(0)
    c.set_U(U1 - 1.0)
    c.assemble()
    got0 = np.empty(3 * sm.nn)
    c.read_snapshot(0, got0)
    assert np.array_equal(got0, U0 + 1.0)
    with pytest.raises(_capi.OnsasError):
        c.snapshot_U(2)
    c.close()
