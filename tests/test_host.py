"""Host-side logic (no GPU): mesh generators, workloads, boundary-condition
rules and the convergence state machine against the reference's own test
fixtures, and the C-ABI library's exports."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from onsas_amd import _capi, meshes, workloads
from onsas_amd.analysis import (ConvergenceSettings, MaxIterCriterion, NotConvergedYet,
                                ResidualForceCriterion, ResidualsIterationStep, ΔUCriterion,
                                displacement_tol, isconverged_, iterations, reset_,
                                residual_forces_tol, update_)
from onsas_amd.boundary_conditions import FixedField, GlobalLoad, Pressure
from onsas_amd.entities import NodeSet, Tetrahedra, TriangularFaces

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cube_sizes_and_orientation():
    sm = meshes.cube(4)
    assert sm.nn == 125 and len(sm.tets) == 384
    v = meshes.tet_volumes(sm.coords, sm.tets)
    assert v.min() > 0 and v.sum() == pytest.approx(1.0, rel=1e-14)
    for name, normal in (("x1", (1, 0, 0)), ("x0", (-1, 0, 0)), ("z1", (0, 0, 1))):
        f = TriangularFaces(sm.faces[name])
        np.testing.assert_allclose(f.normal_directions(sm.coords), np.tile(normal, (32, 1)),
                                   atol=1e-15)
        assert f.areas(sm.coords).sum() == pytest.approx(1.0, rel=1e-14)


def test_cube_matches_reference_hex_split():
    """One hexahedron gives exactly the six tets of uniaxial_extension.jl:66-71."""
    sm = meshes.cube(1, L=1.0)
    # reference corner -> our node id for the unit hex (x fastest numbering)
    ref = {1: (0, 0, 0), 2: (0, 0, 1), 3: (0, 1, 1), 4: (0, 1, 0), 5: (1, 0, 0), 6: (1, 0, 1),
           7: (1, 1, 1), 8: (1, 1, 0)}
    nid = {k: i + 2 * (j + 2 * kk) for k, (i, j, kk) in ref.items()}
    expect = [(1, 4, 2, 6), (6, 2, 3, 4), (4, 3, 6, 7), (4, 1, 5, 6), (4, 6, 5, 8), (4, 7, 6, 8)]
    got = [tuple(t) for t in sm.tets.tolist()]
    assert got == [tuple(nid[a] for a in t) for t in expect]


def test_large_mesh_sizes():
    """Sizes of the benchmark meshes (the 64^3 compression cube and the 41^3 lattice)."""
    sm = meshes.cube(64, jitter=0.02, seed=2)
    assert sm.nn == 274625 and len(sm.tets) == 1572864
    assert meshes.tet_volumes(sm.coords, sm.tets).min() > 0
    tl = meshes.truss_lattice(40, jitter=0.01)
    assert tl.nn == 68921 and len(tl.trusses) == 462520
    assert len(np.unique(np.sort(tl.trusses, 1), axis=0)) == 462520
    cy = meshes.cylinder(24, 192, 96, 0.1, 0.2, 0.5)
    assert cy.nn == 465600 and len(cy.tets) == 2654208


def test_jitter_deterministic_and_keeps_boundary():
    a = meshes.cube(6, jitter=0.02, seed=2)
    b = meshes.cube(6, jitter=0.02, seed=2)
    np.testing.assert_array_equal(a.coords, b.coords)
    c = meshes.cube(6, jitter=0.02, seed=3)
    assert not np.array_equal(a.coords, c.coords)
    ref = meshes.cube(6)
    bnd = np.concatenate([a.node_sets[k] for k in ("x0", "x1", "y0", "y1", "z0", "z1")])
    np.testing.assert_array_equal(a.coords[bnd], ref.coords[bnd])
    assert np.abs(a.coords - ref.coords).max() <= 0.02 / 6 + 1e-15


# ---- boundary conditions (test/boundary_conditions/boundary_conditions.jl) ----------
# reference fixture: n1 (0,0,0), n2 (0,1,0), n3 (0,0,1), n4 (2,0,1); 0-based here
X4 = np.array([[0, 0, 0], [0, 1, 0], [0, 0, 1], [2, 0, 1]], float)
FACE = TriangularFaces([[0, 1, 2]])
TET = Tetrahedra([[0, 1, 2, 3]])


def test_fixed_field_dofs():
    """boundary_conditions.jl:29-47 (dofs 1-based there)."""
    bc = FixedField("u", [1, 3], "fixed_bc_generic")
    assert (bc.apply(NodeSet([2])) + 1).tolist() == [7, 9]
    assert (bc.apply(FACE) + 1).tolist() == [1, 3, 4, 6, 7, 9]
    assert (bc.apply(TET) + 1).tolist() == [1, 3, 4, 6, 7, 9, 10, 12]


def test_global_load_lumping():
    """boundary_conditions.jl:62-88: node f = bc(t); face A/3 per node; tet V/4 per node."""
    t = 2.0
    bc = GlobalLoad("u", lambda t: [np.sin(t), t, t ** 2], "bc_generic")
    f = bc(t)
    d, v = bc.apply(NodeSet([0]), t, X4)
    assert d.tolist() == [0, 1, 2] and np.array_equal(v, f)
    d, v = bc.apply(FACE, t, X4)
    assert d.tolist() == list(range(9))
    np.testing.assert_allclose(v, np.tile(f * 0.5 / 3, 3), rtol=1e-15)
    d, v = bc.apply(TET, t, X4)
    assert d.tolist() == list(range(12))
    np.testing.assert_allclose(v, np.tile(f * (1 / 3) / 4, 4), rtol=1e-15)


def test_pressure_apply():
    """boundary_conditions.jl:90-105: p_vec = -repeat([p(t) A / 3, 0, 0], 3)."""
    t = 2.0
    bc = Pressure("u", lambda t: t ** 2, "normal_pressure")
    d, v = bc.apply(FACE, t, X4)
    assert d.tolist() == list(range(9))
    np.testing.assert_allclose(v, -np.tile([t ** 2 * 0.5 / 3, 0, 0], 3), atol=1e-15)


def test_structure_free_dofs():
    s, sm = workloads.cube_extension(n=4)
    assert s.num_dofs() == 375 and s.num_free_dofs() == 300
    assert np.all(np.diff(s.free_dofs) > 0)
    s3, sm3, p = workloads.cube_compression(n=8)
    assert p == pytest.approx(174320.69, abs=0.01)


# ---- convergence state machine (test/structural_solvers/structural_solvers.jl:21-70) ----

def test_convergence_state_machine(golden):
    g = golden["convergence"]
    tols = ConvergenceSettings(g["tol_u"], g["tol_f"], g["max_iter"])
    assert residual_forces_tol(tols) == g["tol_f"] and displacement_tol(tols) == g["tol_u"]
    ri = ResidualsIterationStep()
    assert iterations(ri) == 0 and isinstance(ri.criterion, NotConvergedYet)
    reset_(ri)
    assert all(x >= 1e3 for x in displacement_tol(ri)) and all(x >= 1e3 for x in residual_forces_tol(ri))
    update_(ri, ΔUCriterion())
    assert isinstance(isconverged_(ri, tols), NotConvergedYet)
    dU, U, dr, fe = (np.array(g[k]) for k in ("dU", "U", "dr", "fext"))
    update_(ri, np.linalg.norm(dU), np.linalg.norm(dU) / np.linalg.norm(U), np.linalg.norm(dr),
            np.linalg.norm(dr) / np.linalg.norm(fe))
    assert iterations(ri) == 1
    assert displacement_tol(ri) == (np.linalg.norm(dU) / np.linalg.norm(U), np.linalg.norm(dU))
    assert not isinstance(isconverged_(ri, tols), NotConvergedYet)
    reset_(ri)
    assert iterations(ri) == 0 and isinstance(isconverged_(ri, tols), NotConvergedYet)


def test_convergence_order_and_assertions():
    tols = ConvergenceSettings(1e-6, 1e-6, 2)
    ri = ResidualsIterationStep()
    update_(ri, 1.0, 1e-7, 1.0, 1e-7)
    assert isinstance(isconverged_(ri, tols), ResidualForceCriterion)  # residual checked first
    ri = ResidualsIterationStep(1.0, 1.0, 1e-7, 1.0, 0)
    assert isinstance(isconverged_(ri, tols), ΔUCriterion)
    ri = ResidualsIterationStep(1.0, 1.0, 1.0, 1.0, 3)
    with pytest.warns(UserWarning):
        assert isinstance(isconverged_(ri, tols), MaxIterCriterion)
    ri = ResidualsIterationStep(0.0, 1.0, 0.0, 1.0, 1)
    with pytest.raises(AssertionError):
        isconverged_(ri, tols)


# ---- the C ABI library ------------------------------------------------------------

def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "onsas_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(onsas_[A-Za-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    syms = _header_symbols()
    assert len(syms) >= 25
    assert sorted(_capi.EXPORTED) == syms
    L = ctypes.CDLL(_capi.LIB_PATH)
    for s in syms:
        assert hasattr(L, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r" T (onsas_\w+)", out))
    assert set(syms) <= exported


def test_library_version_and_no_device_error():
    L = _capi.load()
    assert b"gfx950" in L.onsas_version()
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(_capi.OnsasError) as e:
        _capi.Context()
    assert e.value.code == _capi.ERR_HIP


def test_integration_c_example_compiles(tmp_path):
    """The C usage example of INTEGRATION.md compiles against include/onsas_hip.h
    (warnings are errors), so its calls match the declared signatures."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    doc = open(os.path.join(root, "INTEGRATION.md")).read()
    blocks = re.findall(r"```c\n(.*?)```", doc, flags=re.S)
    assert len(blocks) == 1
    src = tmp_path / "example.c"
    src.write_text(blocks[0])
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(root, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_host_setup_under_sanitizers(tmp_path):
    """The host algorithms of the setup (block pattern, slot maps, gather order,
    the multigrid's aggregation, tentative prolongator, sum lists and
    sliced-ELL layout: onsas.jl_amd/csrc/host_setup.hpp) built with g++ under
    AddressSanitizer and UndefinedBehaviorSanitizer and run on structured and
    scrambled meshes with invariant checks (tests/host/host_setup_check.cpp)."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "host_setup_check")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=all", "-I", os.path.join(root, "onsas.jl_amd", "csrc"),
                        os.path.join(root, "tests", "host", "host_setup_check.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok ") == 4, r.stdout
