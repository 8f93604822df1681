"""GPU parity on the truss-lattice and thick-cylinder workloads, and the linear
static analysis on the engine.

  * truss lattice (SVK bars, both strain models, jittered nodes, z = 0 pinned,
    top load scaled so that the linear response is 1 % of the height): converged
    Newton states vs the oracle's FP64 Newton at 1e-8 of max |U|, both driven to
    a relative residual of 1e-11 (the FP64 floor of the first, lightest load
    step is about 1.3e-12; the equilibrium path is unique, so the converged
    states must agree whatever the inner solver did).
  * thick cylinder (linear tets, inner pressure): LinearStaticAnalysis on the
    GPU vs the oracle's solve at 1e-8 of max |U|, and the plane-strain Lame
    solution of examples/cylinder_internal_pressure within the refinement error.
"""
import numpy as np
import pytest

from _helpers import fext_steps, oracle_model
from onsas_amd import workloads
from onsas_amd.analysis import ConvergenceSettings, NewtonRaphson, NonLinearStaticAnalysis, solve
from onsas_amd.backend import MI355XBackend
from onsas_amd.entities import GreenStrain, RotatedEngineeringStrain
from onsas_amd.linear_analysis import LinearStaticAnalysis

pytestmark = pytest.mark.gpu


def _scaled_lattice(oracle, n, strain):
    """Lattice with the top load scaled so that the linear (U = 0 tangent)
    response has a mean top displacement of 1 % of the height."""
    import scipy.sparse.linalg as spla
    s, sm = workloads.truss_lattice(n=n, strain_model=strain)
    m = oracle_model(oracle, s)
    f1 = fext_steps(oracle, s, [1.0])[0]
    Kff = m.assemble(np.zeros(m.n), want_full=False, want_ff=True)["Kff"]
    u = np.zeros(m.n)
    u[m.free] = spla.spsolve(Kff.tocsc(), f1[m.free])
    mean_top = u.reshape(-1, 3)[sm.node_sets["z1"], 2].mean()
    F = 0.01 * n * 0.1 / abs(mean_top) * np.sign(mean_top)
    return workloads.truss_lattice(n=n, strain_model=strain, F=F)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("strain", ["rotated", "green"])
def test_truss_lattice_vs_oracle(oracle, prec, strain):
    print(f"[gpu] truss lattice {strain} {prec}", flush=True)
    model = RotatedEngineeringStrain if strain == "rotated" else GreenStrain
    s, sm = _scaled_lattice(oracle, 5, model)
    nsteps = 5
    alg = NewtonRaphson(ConvergenceSettings(rel_U_tol=1e-16, rel_res_force_tol=1e-11, max_iter=30))
    sol = solve(NonLinearStaticAnalysis(s, NSTEPS=nsteps), alg,
                MI355XBackend(precision=prec, maxiter=20000))
    m = oracle_model(oracle, s)
    lfs = np.linspace(1 / nsteps, 1.0, nsteps)
    _, Us, rec = m.newton(fext_steps(oracle, s, lfs), rel_U_tol=1e-16, rel_res_tol=1e-11,
                          max_iter=30, cg_reltol=1e-13)
    assert all(r["criterion"] == 1 for r in rec)
    top = sm.node_sets["z1"]
    assert abs(Us[-1].reshape(-1, 3)[top, 2].mean()) > 0.002 * 0.5  # a visible deflection
    for k in range(nsteps):
        scale = np.abs(Us[k]).max()
        err = np.abs(sol.U[k] - Us[k]).max() / scale
        assert err < 1e-8, (k, err)
        assert sol.records[k]["dr_rel"] <= 1e-11


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_cylinder_linear_analysis_vs_oracle(oracle, prec):
    print(f"[gpu] cylinder linear analysis {prec}", flush=True)
    s, sm = workloads.thick_cylinder(nr=3, ntheta=24, nz=3, Lz=0.1)
    nsteps = 2
    sol = solve(LinearStaticAnalysis(s, NSTEPS=nsteps), None,
                MI355XBackend(precision=prec, maxiter=20000))
    m = oracle_model(oracle, s)
    lfs = np.linspace(1 / nsteps, 1.0, nsteps)
    _, Us, _ = m.newton(fext_steps(oracle, s, lfs), rel_U_tol=1e-16, rel_res_tol=1e-13,
                        max_iter=30, cg_reltol=1e-14)
    for k in range(nsteps):
        err = np.abs(sol.U[k] - Us[k]).max() / np.abs(Us[k]).max()
        assert err < 1e-8, (k, err)
    # linear: the second step is twice the first
    np.testing.assert_allclose(sol.U[1], 2 * sol.U[0], rtol=0, atol=1e-8 * np.abs(sol.U[1]).max())
    # stresses of the last step: sigma_zz = nu (sigma_rr + sigma_tt) holds only in
    # the mean for a discretised cylinder; check the element stress of the GPU
    # against the oracle's element routine instead
    sig = sol.stress[-1][0]
    X = sm.coords
    U = sol.U[-1]
    for e in range(0, len(sm.tets), 97):
        tet = sm.tets[e]
        dofs = (3 * tet[:, None] + np.arange(3)).reshape(-1)
        _, _, so, _ = oracle.tet_linear(210e9, 0.3, X[tet].reshape(-1), U[dofs])
        np.testing.assert_allclose(sig[e], so, rtol=0, atol=1e-9 * np.abs(so).max())


def test_cylinder_lame_on_gpu():
    """Refinement toward the plane-strain Lame solution (the reference's own
    example tolerance is 1e-2 (Re - Ri)); same meshes as the oracle's test."""
    print("[gpu] cylinder Lame", flush=True)
    ex_i = workloads.lame_ur(0.1, 0.1, 0.2, 210e9, 0.3, 1e7)
    ex_o = workloads.lame_ur(0.2, 0.1, 0.2, 210e9, 0.3, 1e7)
    errs = []
    for nr, nt in ((4, 32), (8, 64), (16, 128)):
        s, sm = workloads.thick_cylinder(nr=nr, ntheta=nt, nz=2, Lz=0.05)
        sol = solve(LinearStaticAnalysis(s, NSTEPS=1), None, MI355XBackend(precision="mixed"))
        U = sol.U[-1].reshape(-1, 3)
        r = np.linalg.norm(sm.coords[:, :2], axis=1)
        ur = (U[:, 0] * sm.coords[:, 0] + U[:, 1] * sm.coords[:, 1]) / r
        errs.append(max(np.abs(ur[sm.node_sets["inner"]] / ex_i - 1).max(),
                        np.abs(ur[sm.node_sets["outer"]] / ex_o - 1).max()))
    print("  Lame errors", errs, flush=True)
    assert errs[1] < 0.6 * errs[0] and errs[2] < 0.6 * errs[1] and errs[2] < 0.02


def test_gmsh_mesh_solves_like_the_direct_mesh(tmp_path):
    """A cube written as an MSH 2.2 file with the reference's physical-name
    labels, read back, solved on the GPU: the same displacements as the cube
    built directly (same arrays, so the same bits), written to ParaView files."""
    from onsas_amd import meshes
    from onsas_amd.boundary_conditions import FixedField, Pressure, StructuralBoundaryCondition
    from onsas_amd.entities import Tetrahedra, TriangularFaces
    from onsas_amd.gmsh import MshFile, mesh_from_msh, write_msh
    from onsas_amd.materials import SVK
    from onsas_amd.structure import StructuralMaterial, Structure
    from onsas_amd.vtk import read_vtu, write_vtk
    print("[gpu] gmsh cube", flush=True)
    sm = meshes.cube(3, jitter=0.1, seed=3)
    fn = write_msh(str(tmp_path / "cube.msh"), sm.coords,
                   [("svk_tetrahedron_", 4, sm.tets), ("_triangle_tension", 2, sm.faces["x1"]),
                    ("_triangle_fixed-ux", 2, sm.faces["x0"]),
                    ("_triangle_fixed-uy", 2, sm.faces["y0"]),
                    ("_triangle_fixed-uz", 2, sm.faces["z0"])])
    gm = mesh_from_msh(MshFile(fn), {"tetrahedron": Tetrahedra, "triangle": TriangularFaces})
    bcs = {FixedField("u", [1], "ux"): [gm.face_set("fixed-ux")],
           FixedField("u", [2], "uy"): [gm.face_set("fixed-uy")],
           FixedField("u", [3], "uz"): [gm.face_set("fixed-uz")],
           Pressure("u", lambda t: -3e5 * t, "tension"): [gm.face_set("tension")]}
    s = Structure(gm, StructuralMaterial({SVK(E=1e6, nu=0.3, label="svk"): gm.element_set("tetrahedron")}),
                  StructuralBoundaryCondition(bcs))
    alg = NewtonRaphson(ConvergenceSettings(rel_U_tol=1e-14, rel_res_force_tol=1e-10, max_iter=30))
    sol = solve(NonLinearStaticAnalysis(s, NSTEPS=3), alg, MI355XBackend(precision="mixed"))
    s_ref, _ = workloads.cube_extension(n=3, jitter=0.1, seed=3)
    ref = solve(NonLinearStaticAnalysis(s_ref, NSTEPS=3), alg, MI355XBackend(precision="mixed"))
    for k in range(3):
        np.testing.assert_array_equal(sol.U[k], ref.U[k])
    pvd = write_vtk(sol, str(tmp_path / "cube"))
    d = read_vtu(str(tmp_path / "cube_timestep_3.vtu"))
    np.testing.assert_array_equal(d["point"]["Displacement"], sol.U[2].reshape(-1, 3))
    np.testing.assert_array_equal(d["cell"]["σxx"], sol.stress[2][0][:, 0, 0])
    assert pvd.endswith("cube.pvd")


def _brute_locate(coords, tets, P, tol=1e-12):
    """First element (in element order) whose barycentric coordinates
    lambda = A^-T [1; p], A rows [1 x y z] (Tetrahedrons.jl:287-313), are all
    >= -tol; -1 outside."""
    out = np.full(len(P), -1)
    W = np.zeros((len(P), 4))
    for e, t in enumerate(tets):
        A = np.hstack([np.ones((4, 1)), coords[t]])
        lam = np.linalg.solve(A.T, np.vstack([np.ones(len(P)), P.T])).T  # (np, 4)
        hit = (out < 0) & np.all(lam >= -tol, axis=1)
        out[hit] = e
        W[hit] = lam[hit]
    return out, W


def test_point_location_matches_brute_force():
    """GPU bin-grid search vs a brute-force search over every element, with
    points inside, outside, on nodes and on shared faces (the first element in
    element order wins, as the reference's Serial search)."""
    from onsas_amd.backend import GpuStaticState
    from onsas_amd.points import PointEvalHandler, interpolate
    print("[gpu] point location", flush=True)
    s, sm = workloads.cube_extension(n=4, jitter=0.15, seed=11)
    st = GpuStaticState(s, MI355XBackend(precision="fp64"))
    rng = np.random.default_rng(3)
    P = np.vstack([rng.uniform(-0.1, 1.1, size=(400, 3)),          # some outside
                   sm.coords[rng.choice(sm.nn, 40, replace=False)],  # on nodes
                   sm.coords[sm.tets[:30]].mean(axis=1),               # centroids
                   sm.coords[sm.tets[:30, :3]].mean(axis=1)])          # on faces
    h = PointEvalHandler(st, P)
    eb, wb = _brute_locate(sm.coords, sm.tets, P)
    elem = np.full(len(P), -1)
    elem[h.in_mesh_points_idx] = h.in_mesh_elements_idx
    np.testing.assert_array_equal(elem, eb)
    np.testing.assert_allclose(h.weights, wb[h.in_mesh_points_idx], rtol=0, atol=1e-12)
    assert len(h.not_in_mesh_points_idx) > 0 and len(h.in_mesh_points_idx) > 300
    # linear fields are interpolated exactly
    f = sm.coords @ np.array([1.0, -2.0, 0.5]) + 3.0
    np.testing.assert_allclose(interpolate(h, f), h.in_mesh_points @ np.array([1.0, -2.0, 0.5]) + 3.0,
                               rtol=0, atol=1e-12)
    st.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_cylinder_mid_size_vs_oracle(oracle, prec):
    """A mid-size cylinder (6 x 48 x 12 cells, 20,736 tets, 13,104 dofs): the
    GPU LinearStaticAnalysis vs the oracle's FP64 solve at 1e-8 of max |U|."""
    print(f"[gpu] cylinder 6x48x12 linear analysis {prec}", flush=True)
    s, sm = workloads.thick_cylinder(nr=6, ntheta=48, nz=12, Lz=0.1)
    sol = solve(LinearStaticAnalysis(s, NSTEPS=1), None,
                MI355XBackend(precision=prec, maxiter=50000, preconditioner="amg"))
    m = oracle_model(oracle, s)
    _, Us, _ = m.newton(fext_steps(oracle, s, [1.0]), rel_U_tol=1e-16, rel_res_tol=1e-13,
                        max_iter=30, cg_reltol=1e-14)
    err = np.abs(sol.U[0] - Us[0]).max() / np.abs(Us[0]).max()
    print(f"  max rel difference {err:.2e}", flush=True)
    assert err < 1e-8


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_stored_stress_is_at_the_last_assembly(oracle, monkeypatch, prec):
    """store! keeps the sigma/eps of the step's last assemble!, which precedes the
    final U += dU (NonLinearStaticAnalyses.jl:90-98, StaticAnalyses.jl:157-174).
    With a loose Newton tolerance the last update is visible: the stored P and C
    must equal the oracle's element outputs at the U of that last assembly
    (U_k), and differ from those at the stored U (U_k + dU)."""
    from onsas_amd import analysis
    print(f"[gpu] stored stress at the last assembly {prec}", flush=True)
    s, sm, p = workloads.cube_compression(n=4)
    u_before = []
    orig = analysis.step_

    def recording_step(sa, alg, linear_solver=None, sol=None):
        u_before.append((sa.current_step, sa.state.ctx.get_U()))
        return orig(sa, alg, linear_solver, sol)

    monkeypatch.setattr(analysis, "step_", recording_step)
    alg = NewtonRaphson(ConvergenceSettings(rel_U_tol=1e-16, rel_res_force_tol=1e-3, max_iter=30))
    sol = solve(NonLinearStaticAnalysis(s, NSTEPS=3), alg,
                MI355XBackend(precision=prec, store_stresses=True))
    (mat, es), = list(s.materials)
    K, G = mat.params()
    seen_gap = 0.0
    for k in range(3):
        Uk = [u for step, u in u_before if step == k + 1][-1]  # U at the step's last assembly
        Uk1 = sol.U[k]
        P, Cm = sol.stress[k][0], sol.strain[k][0]
        for e in range(0, len(sm.tets), 37):
            tet = sm.tets[e]
            X = sm.coords[tet].reshape(-1)
            dofs = (3 * tet[:, None] + np.arange(3)).reshape(-1)
            _, _, Po, Co = oracle.tet_hyperelastic(oracle.MAT_NEOHOOKEAN, K, G, X, Uk[dofs])
            _, _, P1, _ = oracle.tet_hyperelastic(oracle.MAT_NEOHOOKEAN, K, G, X, Uk1[dofs])
            sc = np.abs(Po).max()
            np.testing.assert_allclose(P[e], Po, rtol=0, atol=1e-10 * sc)
            np.testing.assert_allclose(Cm[e], Co, rtol=0, atol=1e-12)
            seen_gap = max(seen_gap, np.abs(P1 - Po).max() / sc)
    print(f"  largest |P(U_k + dU) - P(U_k)| / |P| over the checked elements {seen_gap:.2e}",
          flush=True)
    assert seen_gap > 1e-7  # the test can tell U_k from U_k + dU
    sol.analysis.state.close()
