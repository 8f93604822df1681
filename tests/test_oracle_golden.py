"""The oracle against every golden vector the reference's own tests hold for
this path (tests/golden/reference_vectors.json, transcribed with file:line)."""
import numpy as np
import pytest


def test_tet_svk_golden(golden, oracle):
    """test/entities/tetrahedrons.jl:48-99: f, K, C of an SVK tet, rtol 1e-3."""
    g = golden["tet_svk"]
    X = np.array(g["X"], float).reshape(-1)
    u = np.array(g["u"])
    V, st = oracle.tet_volume(X)
    assert st == 0 and V == g["volume"]
    f, K, P, C = oracle.tet_hyperelastic(oracle.MAT_SVK, g["lambda"], g["G"], X, u)
    np.testing.assert_allclose(f, g["fint"], rtol=g["rtol"])
    np.testing.assert_allclose(K, np.array(g["K"]), rtol=g["rtol"], atol=1e-4)
    np.testing.assert_allclose(C, np.array(g["C"]), rtol=g["rtol"])
    # a plausible mistake must fail: transposing H (u ⊗ dN vs dN ⊗ u) changes f
    f2, *_ = oracle.tet_hyperelastic(oracle.MAT_SVK, g["lambda"], g["G"], X,
                                     u.reshape(4, 3)[[1, 0, 2, 3]].reshape(-1))
    assert not np.allclose(f2, g["fint"], rtol=g["rtol"])


def test_tet_linear_equals_svk_at_zero(golden, oracle):
    """test/entities/tetrahedrons.jl:122-138: linear K == SVK K at u = 0 and
    f_linear = K_svk(0) u."""
    g = golden["tet_svk"]
    X = np.array(g["X"], float).reshape(-1)
    u = np.array(g["u"])
    lam, G = g["lambda"], g["G"]
    E = G * (3 * lam + 2 * G) / (lam + G)
    nu = lam / (2 * (lam + G))
    f_lin, K_lin, s, e = oracle.tet_linear(E, nu, X, u)
    _, K0, _, _ = oracle.tet_hyperelastic(oracle.MAT_SVK, lam, G, X, np.zeros(12))
    np.testing.assert_allclose(K_lin, K0, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(f_lin, K0 @ u, rtol=1e-12, atol=1e-14)


def test_svk_material_golden(golden, oracle):
    """test/materials/materials.jl:112-139: S and dS/dE of SVK at a given E."""
    g = golden["svk_material"]
    S, D = oracle.svk_stress(g["lambda"], g["G"], np.array(g["E"]))
    np.testing.assert_allclose(S, np.array(g["S"]), rtol=g["rtol"])
    np.testing.assert_allclose(D, np.array(g["dSdE"]), rtol=g["rtol"])


def test_linear_material_golden(golden, oracle):
    """test/materials/materials.jl:41-75: sigma = C voigt(eps, 2) for steel."""
    g = golden["linear_material"]
    E, nu = g["E"], g["nu"]
    c = g["eps_voigt_components"]
    eps = np.array([[c["e11"], c["e12"], c["e13"]], [c["e12"], c["e22"], c["e23"]],
                    [c["e13"], c["e23"], c["e33"]]])
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    G = E / (2 * (1 + nu))
    Cm = np.zeros((6, 6))
    Cm[:3, :3] = lam
    Cm[:3, :3] += 2 * G * np.eye(3)
    Cm[3:, 3:] = G * np.eye(3)
    ev = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[1, 2], 2 * eps[0, 2], 2 * eps[0, 1]])
    sv = Cm @ ev
    expect = np.array([[sv[0], sv[5], sv[4]], [sv[5], sv[1], sv[3]], [sv[4], sv[3], sv[2]]])
    S, D = oracle.linear_stress(E, nu, eps)
    np.testing.assert_allclose(S, expect, rtol=g["rtol"])
    np.testing.assert_allclose(D, Cm, rtol=1e-12)


def test_truss_1d_golden(golden, oracle):
    """test/entities/trusses.jl:16-62 (1D truss, embedded in 3D along x)."""
    g = golden["truss_1d"]
    E, A = g["E"], g["A"]
    X = np.array([g["x1"], 0, 0, g["x2"], 0, 0])
    u = np.array([g["u1"], 0, 0, g["u2"], 0, 0])
    l_ref = abs(g["x2"] - g["x1"])
    l_def = abs(g["x2"] + g["u2"] - g["x1"] - g["u1"])
    eps_rot = (l_def ** 2 - l_ref ** 2) / (l_ref * (l_ref + l_def))
    f, K, s11, e11 = oracle.truss(oracle.STRAIN_ROTATED, E, A, X, u)
    assert e11 == pytest.approx(eps_rot, rel=g["rtol"])
    assert s11 == pytest.approx(E * eps_rot * l_def / l_ref, rel=g["rtol"])
    assert f[0] == pytest.approx(-E * eps_rot * A, rel=g["rtol"])
    assert f[3] == pytest.approx(E * eps_rot * A, rel=g["rtol"])
    kx = K[np.ix_([0, 3], [0, 3])]
    np.testing.assert_allclose(kx, kx[0, 0] * np.array([[1, -1], [-1, 1]]), rtol=g["rtol"])


def test_truss_3d_zero_golden(golden, oracle):
    """test/entities/trusses.jl:64-118: zero displacement -> f = 0, K[1] = E A / l."""
    g = golden["truss_3d_zero"]
    X = np.array(g["x1"] + g["x2"], float)
    for model in (oracle.STRAIN_ROTATED, oracle.STRAIN_GREEN):
        f, K, s11, e11 = oracle.truss(model, g["E"], g["A"], X, np.zeros(6))
        assert np.linalg.norm(f) == 0 and s11 == 0 and e11 == 0
        assert K[0, 0] == pytest.approx(g["E"] * g["A"] / 2.0, rel=1e-15)


def test_face_and_pressure_golden(golden, oracle):
    """test/entities/triangular_faces.jl:32-33 and
    test/boundary_conditions/boundary_conditions.jl:90-105."""
    g = golden["triangular_face"]
    X = np.array(g["X"], float).reshape(-1)
    f = oracle.face_pressure(X, 1.0)  # p = 1: each node gets -n A / 3
    n = -3 * f[0] / g["area"]
    np.testing.assert_allclose(n, g["normal"], atol=1e-15)
    # the boundary-condition test face: (0,0,0), (0,1,0), (0,0,1), p(t) = t^2 at t = 2
    X2 = np.array([0, 0, 0, 0, 1, 0, 0, 0, 1], float)
    t = golden["pressure"]["t"]
    f2 = oracle.face_pressure(X2, t ** 2)
    expect = -np.tile([t ** 2 * 0.5 / 3, 0, 0], 3).reshape(3, 3)
    np.testing.assert_allclose(f2, expect, rtol=1e-15, atol=1e-15)


def test_assembler_golden(golden, oracle):
    """test/structural_solvers/structural_solvers.jl:77-126: COO order and K_glob."""
    g = golden["assembler"]
    Ke = np.array(g["Ke"])
    I, J, V = [], [], []
    for dofs in g["dofs"]:
        d = np.array(dofs) - 1
        for c in range(2):  # column-major append (Assemblers.jl:59-66)
            for r in range(2):
                I.append(d[r])
                J.append(d[c])
                V.append(Ke[r, c])
    assert [i + 1 for i in I] == g["I"] and [j + 1 for j in J] == g["J"]
    K = oracle.coo_sum(I, J, V, 3).toarray()
    np.testing.assert_array_equal(K, np.array(g["K_glob"]))


def test_two_truss_assembly_manufactured(oracle):
    """test/structural_analyses/static_analyses.jl:150-188: assembly equals the
    dense sum of the element contributions."""
    X = np.array([0, 0, 0, 1, 0.2, 0, 2, 0.1, 0.3])
    rng = np.random.default_rng(0)
    U = rng.normal(scale=1e-2, size=9)
    m = oracle.Model(X, trusses=[[0, 1], [1, 2]], truss_E=3.0, truss_A=[0.5, 0.7],
                     truss_strain=oracle.STRAIN_GREEN)
    out = m.assemble(U)
    F = np.zeros(9)
    K = np.zeros((9, 9))
    for e, (a, b) in enumerate([(0, 1), (1, 2)]):
        d = np.r_[3 * a:3 * a + 3, 3 * b:3 * b + 3]
        fe, ke, _, _ = oracle.truss(oracle.STRAIN_GREEN, 3.0, [0.5, 0.7][e], X[d], U[d])
        F[d] += fe
        K[np.ix_(d, d)] += ke
    np.testing.assert_allclose(out["fint"], F, rtol=1e-14)
    np.testing.assert_allclose(out["K"].toarray(), K, rtol=1e-14)
