"""The synthetic workloads of the benchmark suite, built with the reference's
API, plus the closed-form solutions their examples check against.

cube_extension     -- SVK cube in uniaxial tension (configuration of
                      examples/uniaxial_extension/uniaxial_extension.jl, structured mesh)
truss_lattice      -- SVK space-truss lattice, z=0 pinned, vertical top load
cube_compression   -- NeoHookean cube in uniaxial compression
                      (examples/uniaxial_compression/uniaxial_compression.jl)
thick_cylinder     -- linear elastic thick cylinder under internal pressure
                      (examples/cylinder_internal_pressure/cylinder_internal_pressure.jl)
"""
from __future__ import annotations

import numpy as np

from . import meshes
from .boundary_conditions import FixedField, GlobalLoad, Pressure, StructuralBoundaryCondition
from .entities import (Circle, GreenStrain, Mesh, NodeSet, RotatedEngineeringStrain, Tetrahedra,
                       TriangularFaces, Trusses)
from .materials import IsotropicLinearElastic, NeoHookean, SVK
from .structure import StructuralMaterial, Structure


def _rollers(sm):
    """FixedField u1 on x=0, u2 on y=0, u3 on z=0 (symmetry rollers)."""
    return {FixedField("u", [1], "fixed-ux"): [TriangularFaces(sm.faces["x0"], "x0")],
            FixedField("u", [2], "fixed-uy"): [TriangularFaces(sm.faces["y0"], "y0")],
            FixedField("u", [3], "fixed-uz"): [TriangularFaces(sm.faces["z0"], "z0")]}


def cube_extension(n=4, E=1e6, nu=0.3, p=3e5, L=1.0, jitter=0.0, seed=0):
    """Unit cube, n^3 hexes x 6 tets, SVK(E, nu), rollers, Pressure(-p t) on x=1
    (tension: the load p(t) (-n) A / 3 points along +x)."""
    sm = meshes.cube(n, L=L, jitter=jitter, seed=seed)
    tets = Tetrahedra(sm.tets, "tets")
    mesh = Mesh(sm.coords, [tets], [TriangularFaces(f, k) for k, f in sm.faces.items()])
    mat = SVK(E=E, nu=nu, label="svk")
    bcs = _rollers(sm)
    bcs[Pressure("u", lambda t: -p * t, "tension")] = [TriangularFaces(sm.faces["x1"], "x1")]
    s = Structure(mesh, StructuralMaterial({mat: tets}), StructuralBoundaryCondition(bcs))
    return s, sm


def cube_compression(n=64, K=8.3e5, G=3.8e5, p=None, alpha_final=0.85, L=1.0, jitter=0.02,
                     seed=2, G_field=None):
    """Unit cube, n^3 hexes x 6 tets, NeoHookean(K, G), rollers, compressive
    Pressure(p t) on x=1 with p the nominal stress that gives stretch
    alpha_final at t = 1; interior-node jitter keeps the solution homogeneous."""
    if p is None:
        p = -nh_uniaxial_P11(alpha_final, K, G)
    sm = meshes.cube(n, L=L, jitter=jitter, seed=seed)
    tets = Tetrahedra(sm.tets, "tets")
    mesh = Mesh(sm.coords, [tets], [TriangularFaces(f, k) for k, f in sm.faces.items()])
    mat = NeoHookean(K, G if G_field is None else G_field, label="neo")
    bcs = _rollers(sm)
    bcs[Pressure("u", lambda t: p * t, "compression")] = [TriangularFaces(sm.faces["x1"], "x1")]
    s = Structure(mesh, StructuralMaterial({mat: tets}), StructuralBoundaryCondition(bcs))
    return s, sm, p


def thick_cylinder(nr=24, ntheta=192, nz=96, Ri=0.1, Re=0.2, Lz=0.5, E=210e9, nu=0.3, p=1e7):
    """Linear elastic thick cylinder, inner Pressure(p t), u_z fixed on both ends
    (plane strain), u_y fixed on the theta = 0 line and u_x on theta = 90 deg
    (rigid-body modes)."""
    sm = meshes.cylinder(nr, ntheta, nz, Ri, Re, Lz)
    tets = Tetrahedra(sm.tets, "tets")
    mesh = Mesh(sm.coords, [tets], [TriangularFaces(f, k) for k, f in sm.faces.items()])
    mat = IsotropicLinearElastic(E, nu, "mat")
    bcs = {FixedField("u", [3], "fixed-uz"): [NodeSet(np.concatenate([sm.node_sets["z0"],
                                                                      sm.node_sets["z1"]]))],
           FixedField("u", [2], "fixed-uy"): [NodeSet(sm.node_sets["theta0"])],
           FixedField("u", [1], "fixed-ux"): [NodeSet(sm.node_sets["theta90"])],
           Pressure("u", lambda t: p * t, "pressure"): [TriangularFaces(sm.faces["inner"], "in")]}
    s = Structure(mesh, StructuralMaterial({mat: tets}), StructuralBoundaryCondition(bcs))
    return s, sm


def truss_lattice(n=40, h=0.1, E=210e9, nu=0.3, A=1e-4, jitter=0.01, seed=1, F=None,
                  strain_model=RotatedEngineeringStrain):
    """Space-truss lattice on (n+1)^3 nodes, z=0 pinned, vertical nodal load
    F t (per top node) on z = n h.  F defaults to 1 (callers scale it)."""
    sm = meshes.truss_lattice(n, h=h, jitter=jitter, seed=seed)
    d = np.sqrt(4 * A / np.pi)
    bars = Trusses(sm.trusses, Circle(d), strain_model, "bars")
    mesh = Mesh(sm.coords, [bars])
    mat = SVK(E=E, nu=nu, label="steel")
    Fz = 1.0 if F is None else F
    bcs = {FixedField("u", [1, 2, 3], "pinned"): [NodeSet(sm.node_sets["z0"])],
           GlobalLoad("u", lambda t: [0.0, 0.0, Fz * t], "top"): [NodeSet(sm.node_sets["z1"])]}
    s = Structure(mesh, StructuralMaterial({mat: bars}), StructuralBoundaryCondition(bcs))
    return s, sm


def shear_field(points, G0=3.8e5, amp=0.3, ell=0.2, modes=256, seed=3):
    """G(x) = G0 exp(amp g(x)) at the given points, with g a unit-variance
    Gaussian field of squared-exponential correlation length ell, sampled by
    `modes` random Fourier features: g = sqrt(2/M) sum cos(w_m . x + phi_m),
    w_m ~ N(0, I / ell^2), phi_m ~ U(0, 2 pi)."""
    rng = np.random.default_rng(seed)
    w = rng.normal(size=(modes, 3)) / ell
    phi = rng.uniform(0.0, 2 * np.pi, modes)
    pts = np.asarray(points, dtype=np.float64)
    g = np.empty(len(pts))
    for i in range(0, len(pts), 65536):
        g[i:i + 65536] = np.cos(pts[i:i + 65536] @ w.T + phi).sum(axis=1)
    return G0 * np.exp(amp * np.sqrt(2.0 / modes) * g)


def identification_cube(n=48, K=8.3e5, G0=3.8e5, alpha_final=0.85, jitter=0.02, seed=3,
                        n_obs=2000, obs_seed=4):
    """NeoHookean cube in compression with a heterogeneous per-element shear
    modulus G_e = shear_field(centroid_e) (the parameters of an identification
    problem), plus `n_obs` observed nodes drawn from the free surfaces
    (x = 1, y = 1, z = 1).  The pressure is the homogeneous-G0 value of
    cube_compression.  Returns (structure, mesh, p, G_true, obs_nodes)."""
    sm = meshes.cube(n, jitter=jitter, seed=2)
    cent = sm.coords[sm.tets].mean(axis=1)
    G_true = shear_field(cent, G0=G0, seed=seed)
    s, sm, p = cube_compression(n=n, K=K, G=G0, alpha_final=alpha_final, jitter=jitter, seed=2,
                                G_field=G_true)
    surf = np.unique(np.concatenate([sm.node_sets[k] for k in ("x1", "y1", "z1")]))
    rng = np.random.default_rng(obs_seed)
    obs = np.sort(rng.choice(surf, size=min(n_obs, len(surf)), replace=False))
    return s, sm, p, G_true, obs


# ---- closed forms ----------------------------------------------------------------

def svk_uniaxial_alpha(E, p):
    """alpha with E/2 alpha (alpha^2 - 1) = p (uniaxial_extension.jl:188); p > 0 tension."""
    a = 1.0
    for _ in range(100):
        f = E / 2 * a * (a * a - 1) - p
        a -= f / (E / 2 * (3 * a * a - 1))
    return a


def svk_uniaxial_beta(alpha, nu):
    """beta^2 = 1 - nu (alpha^2 - 1) (uniaxial_extension.jl:189)."""
    return np.sqrt(1 - nu * (alpha ** 2 - 1))


def nh_uniaxial_beta(alpha, K, G):
    """Lateral stretch with P22 = G (b - 1/b) + K b (a^2 b^2 - a) = 0
    (uniaxial_compression.jl:252-255)."""
    b = 1.0
    for _ in range(200):
        f = G * (b - 1 / b) + K * b * (alpha ** 2 * b ** 2 - alpha)
        df = G * (1 + 1 / b ** 2) + K * (3 * alpha ** 2 * b ** 2 - alpha)
        b -= f / df
    return b


def nh_uniaxial_P11(alpha, K, G):
    """P11 = G (a - 1/a) + K b^2 (a b^2 - 1) (uniaxial_compression.jl:247-250)."""
    b = nh_uniaxial_beta(alpha, K, G)
    return G * (alpha - 1 / alpha) + K * b ** 2 * (alpha * b ** 2 - 1)


def nh_uniaxial_alpha(P11, K, G):
    """alpha(P11) by Newton on P11(alpha) (monotone)."""
    a = 1.0
    for _ in range(200):
        f = nh_uniaxial_P11(a, K, G) - P11
        h = 1e-7
        df = (nh_uniaxial_P11(a + h, K, G) - nh_uniaxial_P11(a - h, K, G)) / (2 * h)
        a -= f / df
    return a


def homogeneous_displacement(coords, alpha, beta):
    """u = ((alpha - 1) x, (beta - 1) y, (beta - 1) z)."""
    u = np.empty_like(coords)
    u[:, 0] = (alpha - 1) * coords[:, 0]
    u[:, 1] = (beta - 1) * coords[:, 1]
    u[:, 2] = (beta - 1) * coords[:, 2]
    return u.reshape(-1)


def lame_ur(r, Ri, Re, E, nu, p):
    """Plane-strain Lame radial displacement (cylinder_internal_pressure.jl:203-214)."""
    A = (1 + nu) * (1 - 2 * nu) * Ri ** 2 * p / (E * (Re ** 2 - Ri ** 2))
    B = (1 + nu) * Ri ** 2 * Re ** 2 * p / (E * (Re ** 2 - Ri ** 2))
    return A * r + B / r
