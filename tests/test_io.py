"""Host-side I/O: Gmsh .msh ingest (src/Interfaces/Gmsh.jl,
src/StructuralModel/StructuralEntities.jl:76-156) and the ParaView writer
(src/Interfaces/VTK.jl).  No GPU needed."""
import numpy as np
import pytest

from onsas_amd import meshes, workloads
from onsas_amd.analysis import NonLinearStaticAnalysis, Solution
from onsas_amd.boundary_conditions import FixedField, Pressure, StructuralBoundaryCondition
from onsas_amd.entities import Tetrahedra, TriangularFaces
from onsas_amd.gmsh import MshFile, PHYSICAL_NODE_LABEL, _getlabels, mesh_from_msh, read_msh, write_msh
from onsas_amd.materials import SVK
from onsas_amd.structure import StructuralMaterial, Structure
from onsas_amd.vtk import FIELD_NAMES, read_vtu, write_vtk


def _cube_msh(tmp_path, n=3):
    sm = meshes.cube(n, jitter=0.1, seed=3)
    corner = [int(np.argmin(np.linalg.norm(sm.coords, axis=1)))]
    blocks = [("svk_tetrahedron_", 4, sm.tets),
              ("_triangle_tension", 2, sm.faces["x1"]),
              ("_triangle_fixed-ux", 2, sm.faces["x0"]),
              ("_triangle_fixed-uy", 2, sm.faces["y0"]),
              ("_triangle_fixed-uz", 2, sm.faces["z0"]),
              ("_node_corner", 15, np.array(corner)[:, None])]
    return sm, write_msh(str(tmp_path / "cube.msh"), sm.coords, blocks)


def test_physical_name_labels():
    """Gmsh.jl:62-77: names with three `_` parts give (material, entity, bc)."""
    m, e, b = _getlabels(["svk_tetrahedron_", "_triangle_load", "bad", "_node_fixed-ui"])
    assert m == ["svk", "", ""] and e == ["tetrahedron", "triangle", "node"]
    assert b == ["", "load", "fixed-ui"]


def test_read_msh_round_trip(tmp_path):
    sm, fn = _cube_msh(tmp_path)
    coords, conn, names, phys = read_msh(fn)
    np.testing.assert_array_equal(coords, sm.coords)  # repr() round trip is exact
    ntet = len(sm.tets)
    np.testing.assert_array_equal(np.array(conn[:ntet]), sm.tets)
    assert names[0] == "svk_tetrahedron_" and phys[0] == 1 and phys[-1] == 6
    assert len(conn) == ntet + sum(len(sm.faces[k]) for k in ("x1", "x0", "y0", "z0")) + 1
    with pytest.raises(ValueError):
        read_msh("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n")


def test_mesh_from_msh_builds_the_same_structure(tmp_path):
    """The .msh path and the direct path give the same element set, faces and
    fixed dofs (the reference examples' workflow: labels -> sets -> BCs)."""
    sm, fn = _cube_msh(tmp_path)
    msh = MshFile(fn)
    assert msh.entity_label(0) == "tetrahedron" and msh.material_label(0) == "svk"
    gm = mesh_from_msh(msh, {"tetrahedron": Tetrahedra, "triangle": TriangularFaces})
    tets = gm.element_set("tetrahedron")
    np.testing.assert_array_equal(tets.conn, sm.tets)
    np.testing.assert_array_equal(gm.element_set("svk").conn, sm.tets)
    np.testing.assert_array_equal(gm.face_set("tension").conn, sm.faces["x1"])
    assert gm.node_set("corner").ids.tolist() == [0] and PHYSICAL_NODE_LABEL in gm.node_sets
    mat = SVK(E=1e6, nu=0.3, label="svk")
    bcs = {FixedField("u", [1], "ux"): [gm.face_set("fixed-ux")],
           FixedField("u", [2], "uy"): [gm.face_set("fixed-uy")],
           FixedField("u", [3], "uz"): [gm.face_set("fixed-uz")],
           Pressure("u", lambda t: -3e5 * t, "tension"): [gm.face_set("tension")]}
    s = Structure(gm, StructuralMaterial({mat: tets}), StructuralBoundaryCondition(bcs))
    s_ref, _ = workloads.cube_extension(n=3, jitter=0.1, seed=3)
    np.testing.assert_array_equal(np.sort(s.fixed_dofs), np.sort(s_ref.fixed_dofs))
    (_, pc), = s.bcs.pressure_groups()
    (_, pr), = s_ref.bcs.pressure_groups()
    np.testing.assert_array_equal(pc, pr)


def _fake_solution():
    s, sm = workloads.cube_extension(n=2)
    sa = NonLinearStaticAnalysis(s, NSTEPS=2)
    sol = Solution(sa, None)
    rng = np.random.default_rng(0)
    ne = len(sm.tets)
    for k in range(2):
        sol.U[k] = rng.normal(size=3 * sm.nn)
        sol.stress[k] = [rng.normal(size=(ne, 3, 3))]
        sol.strain[k] = [rng.normal(size=(ne, 3, 3))]
    return sol, sm


@pytest.mark.parametrize("ascii", [False, True])
def test_vtk_step_file(tmp_path, ascii):
    sol, sm = _fake_solution()
    fn = write_vtk(sol, str(tmp_path / "step.vtu"), 2, ascii=ascii)
    d = read_vtu(fn)
    np.testing.assert_array_equal(d["points"], sm.coords)
    np.testing.assert_array_equal(d["connectivity"], sm.tets.reshape(-1))
    np.testing.assert_array_equal(d["offsets"], 4 * np.arange(1, len(sm.tets) + 1))
    assert set(d["types"].tolist()) == {10}
    np.testing.assert_array_equal(d["point"]["Displacement"], sol.U[1].reshape(-1, 3))
    S = sol.stress[1][0]
    # component names and their tensor entries (VTK.jl:113-116, 139-146)
    for name, (i, j) in (("σxx", (0, 0)), ("τyz", (1, 2)), ("τxz", (0, 2)), ("τxy", (0, 1)),
                         ("τzy", (2, 1)), ("τzx", (2, 0)), ("τyx", (1, 0))):
        np.testing.assert_array_equal(d["cell"][name], S[:, i, j])
    assert set(d["cell"]) == set(FIELD_NAMES["σ"][1]) | set(FIELD_NAMES["ϵ"][1])


def test_vtk_collection(tmp_path):
    import xml.etree.ElementTree as ET
    sol, sm = _fake_solution()
    pvd = write_vtk(sol, str(tmp_path / "run"))
    root = ET.parse(pvd).getroot()
    ds = root.findall("Collection/DataSet")
    assert [float(e.get("timestep")) for e in ds] == [0.5, 1.0]
    assert [e.get("file") for e in ds] == ["run_timestep_1.vtu", "run_timestep_2.vtu"]
    d = read_vtu(str(tmp_path / "run_timestep_1.vtu"))
    np.testing.assert_array_equal(d["point"]["Displacement"], sol.U[0].reshape(-1, 3))


FIXTURE41 = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                       "cube_manufactured_v41.msh")
# The same mesh as the MSH 4.1 fixture, written by hand in MSH 2.2 (elements
# carry their physical tag first, their elementary entity second).
CUBE22 = """$MeshFormat
2.2 0 8
$EndMeshFormat
$PhysicalNames
6
0 6 "_node_origin"
2 1 "_triangle_load"
2 2 "_triangle_fixed-ux"
2 3 "_triangle_fixed-uy"
2 4 "_triangle_fixed-uz"
3 5 "svk_tetrahedron_"
$EndPhysicalNames
$Nodes
8
1 0 0 0
2 0 0 1
3 0 1 1
4 0 1 0
5 1 0 0
6 1 0 1
7 1 1 1
8 1 1 0
$EndNodes
$Elements
15
1 15 2 6 1 1
2 2 2 1 1 5 8 6
3 2 2 1 1 6 8 7
4 2 2 2 2 4 1 2
5 2 2 2 2 4 2 3
6 2 2 3 3 6 2 1
7 2 2 3 3 6 1 5
8 2 2 4 4 1 4 5
9 2 2 4 4 4 8 5
10 4 2 5 1 1 4 2 6
11 4 2 5 1 6 2 3 4
12 4 2 5 1 4 3 6 7
13 4 2 5 1 4 1 5 6
14 4 2 5 1 4 6 5 8
15 4 2 5 1 4 7 6 8
$EndElements
"""


def test_read_msh41_fixture():
    """A hand-written MSH 4.1 file (the manufactured cube of
    examples/uniaxial_extension/uniaxial_extension.jl:45-73 with the physical
    names of its Gmsh case): nodes by entity block, physical tags through
    $Entities."""
    coords, conn, names, phys = read_msh(FIXTURE41)
    np.testing.assert_array_equal(coords, [[0, 0, 0], [0, 0, 1], [0, 1, 1], [0, 1, 0],
                                           [1, 0, 0], [1, 0, 1], [1, 1, 1], [1, 1, 0]])
    assert names == ["_triangle_load", "_triangle_fixed-ux", "_triangle_fixed-uy",
                     "_triangle_fixed-uz", "svk_tetrahedron_", "_node_origin"]
    np.testing.assert_array_equal(phys, [6, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 5, 5, 5, 5])
    assert [c.tolist() for c in conn[:3]] == [[0], [4, 7, 5], [5, 7, 6]]
    np.testing.assert_array_equal(np.array(conn[9:]), [[0, 3, 1, 5], [5, 1, 2, 3], [3, 2, 5, 6],
                                                       [3, 0, 4, 5], [3, 5, 4, 7], [3, 6, 5, 7]])


def test_msh41_and_msh22_of_one_mesh_agree():
    a, b = read_msh(FIXTURE41), read_msh(CUBE22)
    np.testing.assert_array_equal(a[0], b[0])
    assert [c.tolist() for c in a[1]] == [c.tolist() for c in b[1]]
    assert a[2] == b[2]
    np.testing.assert_array_equal(a[3], b[3])


def test_msh41_entity_blocks_and_tags():
    """Node tags out of file order (numbered by ascending tag), a parametric node
    block (x y z u v), an entity in two physical groups (its elements listed once
    per group, as MSH 2.2 lists them) and one in none (tag 0)."""
    src = """$MeshFormat
4.1 0 8
$EndMeshFormat
$PhysicalNames
2
2 1 "_triangle_a"
2 2 "_triangle_b"
$EndPhysicalNames
$Entities
0 0 2 0
1 0 0 0 1 1 0 2 1 2 0
2 0 0 0 1 1 0 0 0
$EndEntities
$Nodes
2 4 1 4
2 1 0 2
4
3
1 1 0
0 1 0
2 1 1 2
1
2
0 0 0 0.0 0.0
1 0 0 1.0 0.0
$EndNodes
$Elements
2 2 1 2
2 1 2 1
1 1 2 3
2 2 2 1
2 2 3 4
$EndElements
"""
    coords, conn, names, phys = read_msh(src)
    np.testing.assert_array_equal(coords, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]])
    assert [c.tolist() for c in conn] == [[0, 1, 2], [0, 1, 2], [1, 2, 3]]
    np.testing.assert_array_equal(phys, [1, 2, 0])
    assert names == ["_triangle_a", "_triangle_b"]


def test_msh41_fixture_builds_the_reference_example_structure():
    msh = MshFile(FIXTURE41)
    assert msh.entity_label(9) == "tetrahedron" and msh.material_label(9) == "svk"
    gm = mesh_from_msh(msh, {"tetrahedron": Tetrahedra, "triangle": TriangularFaces})
    np.testing.assert_array_equal(gm.face_set("load").conn, [[4, 7, 5], [5, 7, 6]])
    assert gm.node_set("origin").ids.tolist() == [0]
    assert len(gm.element_set("svk").conn) == 6


def test_read_msh_rejects_other_versions():
    with pytest.raises(ValueError):
        read_msh("$MeshFormat\n4.0 0 8\n$EndMeshFormat\n")
    with pytest.raises(ValueError):
        read_msh("$MeshFormat\n4.1 1 8\n$EndMeshFormat\n")
