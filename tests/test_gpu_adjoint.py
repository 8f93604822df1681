"""Adjoint sensitivity to per-element shear moduli on the GPU (no counterpart in
the reference; see onsas_amd/adjoint.py).

Pins, from the most to the least local:
  * the oracle: lambda from a sparse direct solve with the oracle's FP64 tangent
    at the GPU's converged U, and d f_int,e / d G_e as the oracle's element force
    with the other modulus set to 0 and G = 1 (S is linear in G for both
    NeoHookean and SVK, so that force IS the derivative).  Tolerance 1e-9 of
    max |dJ/dG| in both precisions (the mixed mode refines its FP32 solves with
    an FP64 residual to 1e-11).
  * finite differences of J(G) = 1/2 sum |u_obs - target|^2 through complete
    Newton solves (warm-started, FP64 residual below 1e-13): a random direction
    over all elements and the single elements of largest sensitivity, central
    differences, relative agreement 1e-5.
"""
import numpy as np
import pytest

from onsas_amd import _capi, meshes, workloads
from onsas_amd.adjoint import misfit, resolve, rrule
from onsas_amd.analysis import ConvergenceSettings, NewtonRaphson, NonLinearStaticAnalysis, solve
from onsas_amd.backend import MI355XBackend

pytestmark = pytest.mark.gpu

TIGHT = NewtonRaphson(ConvergenceSettings(rel_U_tol=1e-16, rel_res_force_tol=1e-13, max_iter=30))


def _backend(prec, pc="block-jacobi"):
    return MI355XBackend(precision=prec, maxiter=20000, store_stresses=False, preconditioner=pc)


@pytest.mark.parametrize("pc", ["block-jacobi", "amg"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("mat", ["nh", "svk"])
def test_adjoint_matches_oracle(oracle, prec, mat, pc):
    """Forward Newton solve and refined adjoint solve with either PCG
    preconditioner: dJ/dG against the oracle's direct adjoint."""
    import scipy.sparse.linalg as spla
    print(f"[gpu] adjoint vs oracle {mat} {prec} {pc}", flush=True)
    if mat == "nh":
        s, sm, p, G, _ = workloads.identification_cube(n=4, n_obs=50)
        code, p0 = oracle.MAT_NEOHOOKEAN, 8.3e5
        nsteps = 4
    else:
        s, sm = workloads.cube_extension(n=4, jitter=0.05, seed=3)
        cent = sm.coords[sm.tets].mean(axis=1)
        G = workloads.shear_field(cent, G0=384615.38, seed=9)
        mat_obj = next(iter(s.materials))[0]
        mat_obj.G = G
        code, p0 = oracle.MAT_SVK, mat_obj.lam
        nsteps = 3
    sa = NonLinearStaticAnalysis(s, NSTEPS=nsteps)
    sol, pullback = rrule(solve, sa, TIGHT, _backend(prec, pc))
    U = sol.U[-1]
    Ubar = np.random.default_rng(5).normal(size=U.size)
    g = pullback(Ubar)["tets"]
    fixed = s.fixed_dofs
    m = oracle.Model(sm.coords, tets=sm.tets, tet_mat=code, tet_p=(p0, 1.0), tet_G=G,
                     fixed_dofs=fixed)
    Kff = m.assemble(U, want_full=False, want_ff=True)["Kff"]
    lam = np.zeros_like(U)
    lam[m.free] = spla.spsolve(Kff.tocsc(), Ubar[m.free])
    go = np.empty(len(sm.tets))
    for e, tet in enumerate(sm.tets):
        dofs = (3 * tet[:, None] + np.arange(3)).reshape(-1)
        f, _, _, _ = oracle.tet_hyperelastic(code, 0.0, 1.0, sm.coords[tet].reshape(-1), U[dofs])
        go[e] = -lam[dofs] @ f
    err = np.abs(g - go).max() / np.abs(go).max()
    print(f"  max |dJ/dG - oracle| / max |oracle| = {err:.2e}", flush=True)
    assert err < 1e-9, err
    sol.analysis.state.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_set_shear_moduli_matches_construction(oracle, prec):
    print(f"[gpu] set_shear_moduli {prec}", flush=True)
    sm = meshes.cube(3, jitter=0.1, seed=8)
    G = workloads.shear_field(sm.coords[sm.tets].mean(axis=1), seed=12)
    U = np.random.default_rng(3).normal(scale=0.03, size=3 * sm.nn)
    ref = _capi.Context(_capi.PRECISION_FP64 if prec == "fp64" else _capi.PRECISION_MIXED)
    ref.set_nodes(sm.coords)
    ref.add_tet_group(sm.tets, _capi.MAT_NEOHOOKEAN, (8.3e5, 3.8e5), G)
    ref.set_fixed_dofs(np.zeros(0, dtype=np.int64))
    ref.finalize()
    c = _capi.Context(_capi.PRECISION_FP64 if prec == "fp64" else _capi.PRECISION_MIXED)
    c.set_nodes(sm.coords)
    c.add_tet_group(sm.tets, _capi.MAT_NEOHOOKEAN, (8.3e5, 3.8e5), None)  # scalar G first
    c.set_fixed_dofs(np.zeros(0, dtype=np.int64))
    c.finalize()
    c.set_shear_moduli(0, G)
    for ctx in (ref, c):
        ctx.set_U(U)
    f0, _, s0, _ = ref.element_forces(0)
    f1, _, s1, _ = c.element_forces(0)
    assert np.array_equal(f0, f1) and np.array_equal(s0, s1)
    ref.assemble()
    c.assemble()
    assert np.array_equal(ref.get_fint(), c.get_fint())
    ref.close()
    c.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_adjoint_vs_finite_differences(prec):
    print(f"[gpu] adjoint vs FD {prec}", flush=True)
    s_true, sm, p, G_true, obs = workloads.identification_cube(n=6)
    be = _backend(prec)
    sol_true = solve(NonLinearStaticAnalysis(s_true, NSTEPS=4), TIGHT, be)
    rng = np.random.default_rng(5)
    target = sol_true.U[-1].reshape(-1, 3)[obs] + rng.normal(scale=1e-4, size=(len(obs), 3))
    sol_true.analysis.state.close()
    # evaluate at another field (the start of an identification)
    G_eval = workloads.shear_field(sm.coords[sm.tets].mean(axis=1), seed=21)
    s, sm, p, _, _ = workloads.identification_cube(n=6)
    next(iter(s.materials))[0].G = G_eval
    sol, pullback = rrule(solve, NonLinearStaticAnalysis(s, NSTEPS=4), TIGHT, be)
    J0, dJdU = misfit(sol.U[-1], obs, target)
    g = pullback(dJdU)["tets"]

    def J(Gp):
        sp = resolve(sol, Gp, TIGHT, be)
        assert sp.records[-1]["dr_rel"] <= 1e-13
        return misfit(sp.U[-1], obs, target)[0]

    # random direction over all elements
    d = G_eval * np.random.default_rng(6).uniform(-1, 1, size=G_eval.size)
    eps = 1e-4
    fd = (J(G_eval + eps * d) - J(G_eval - eps * d)) / (2 * eps)
    ad = g @ d
    print(f"  directional: fd {fd:.10e} adjoint {ad:.10e} rel {abs(fd - ad) / abs(ad):.2e}",
          flush=True)
    assert abs(fd - ad) <= 1e-5 * abs(ad), (fd, ad)
    # single elements with the largest sensitivities
    for e in np.argsort(-np.abs(g))[:4]:
        h = 1e-3 * G_eval[e]
        Gp, Gm = G_eval.copy(), G_eval.copy()
        Gp[e] += h
        Gm[e] -= h
        fd = (J(Gp) - J(Gm)) / (2 * h)
        print(f"  element {e}: rel {abs(fd - g[e]) / abs(g[e]):.2e}", flush=True)
        assert abs(fd - g[e]) <= 1e-5 * abs(g[e]), (e, fd, g[e])
    sol.analysis.state.close()
