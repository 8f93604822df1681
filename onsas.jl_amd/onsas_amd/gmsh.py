"""Gmsh .msh ingest: meshes of the reference's examples on the engine.

Mirrors src/Interfaces/Gmsh.jl (MshFile, the "material_entity_bc" physical
names, material_label / entity_label / bc_label, PHYSICAL_NODE_LABEL) and
Mesh(::MshFile, ::StructuralEntity) of
src/StructuralModel/StructuralEntities.jl:76-156.  The reference reads the file
with MshReader.jl (not vendored; its MshFileReader returns node coordinates,
element connectivities, the physical names indexed by physical tag and each
element's physical tag).  `read_msh` restates that reader from the published
Gmsh file format for ASCII files of version 2.2 and 4.1 (4.1 is what
`gmsh.write` produces by default, e.g. examples/uniaxial_extension/
uniaxial_mesh.jl:89-90): $PhysicalNames, $Nodes, $Elements, and in 4.1 the
$Entities section that carries the physical tags.  Element types 15 point,
1 line, 2 triangle, 4 tetrahedron.  Nodes are numbered 0-based in ascending
node-tag order (the order gmsh writes them in).
"""
from __future__ import annotations

import numpy as np

from .entities import Mesh, NodeSet, Tetrahedra, TriangularFaces, Trusses

PHYSICAL_NODE_LABEL = "node"  # Gmsh.jl:16
_NODES_PER_TYPE = {15: 1, 1: 2, 2: 3, 4: 4}


def read_msh(source):
    """ASCII MSH 2.2 or 4.1 reader -> (coords (nn, 3), connectivity [int array per
    element, 0-based], physical_names [name of tag 1, 2, ...], physical_index
    [1-based tag per element, 0 when the element has none]).

    MSH 2.2 gives an element's physical tag as its first tag.  MSH 4.1 gives it
    through the element's entity ($Entities): an element of an entity with k
    physical tags is listed k times, once per tag, as the same mesh saved in
    MSH 2.2 lists it; an entity without one gives tag 0."""
    text = source if "\n" in source else open(source, encoding="utf-8").read()
    lines = iter(text.splitlines())
    version = None
    names = {}
    for line in lines:
        tag = line.strip()
        if tag == "$MeshFormat":
            version, ftype, _ = next(lines).split()
            if ftype != "0":
                raise ValueError(f"only ASCII MSH files are read (file type {ftype})")
            if version.startswith("2"):
                return _read_msh2(lines, names)
            if version.startswith("4.1"):
                return _read_msh41(lines, names)
            raise ValueError(f"MSH version {version} is not read (2.x and 4.1 are)")
        if tag == "$PhysicalNames":
            _read_names(lines, names)
    return _read_msh2(iter(text.splitlines()), names)  # no $MeshFormat: 2.x layout


def _read_names(lines, names):
    for _ in range(int(next(lines))):
        dim, ptag, name = next(lines).split(maxsplit=2)
        names[int(ptag)] = name.strip().strip('"')


def _finish(tags, coords, conn_tags, phys, names):
    if coords is None:
        raise ValueError("no $Nodes section")
    order = np.argsort(tags, kind="stable")
    tags, coords = tags[order], coords[order]
    pos = {int(t): k for k, t in enumerate(tags)}
    conn = [np.array([pos[v] for v in c], dtype=np.int64) for c in conn_tags]
    ntag = max(names) if names else 0
    physical_names = [names.get(t, "") for t in range(1, ntag + 1)]
    return coords, conn, physical_names, np.array(phys, dtype=np.int64)


def _element_nodes(etype, nodes, etag):
    if etype not in _NODES_PER_TYPE:
        raise ValueError(f"element type {etype} is not supported")
    if len(nodes) != _NODES_PER_TYPE[etype]:
        raise ValueError(f"element {etag}: {len(nodes)} nodes for type {etype}")
    return nodes


def _read_msh2(lines, names):
    tags = coords = None
    conn, phys = [], []
    for line in lines:
        tag = line.strip()
        if tag == "$PhysicalNames":
            _read_names(lines, names)
        elif tag == "$Nodes":
            n = int(next(lines))
            tags = np.empty(n, dtype=np.int64)
            coords = np.empty((n, 3))
            for k in range(n):
                f = next(lines).split()
                tags[k] = int(f[0])
                coords[k] = [float(f[1]), float(f[2]), float(f[3])]
        elif tag == "$Elements":
            for _ in range(int(next(lines))):
                f = [int(v) for v in next(lines).split()]
                etype, ntags = f[1], f[2]
                conn.append(_element_nodes(etype, f[3 + ntags:], f[0]))
                phys.append(f[3] if ntags > 0 else 0)
    return _finish(tags, coords, conn, phys, names)


def _read_msh41(lines, names):
    ent_phys = {}  # (dim, entity tag) -> physical tags
    tags = coords = None
    conn, phys = [], []
    for line in lines:
        tag = line.strip()
        if tag == "$PhysicalNames":
            _read_names(lines, names)
        elif tag == "$Entities":
            counts = [int(v) for v in next(lines).split()]
            for dim, cnt in enumerate(counts):
                for _ in range(cnt):
                    f = next(lines).split()
                    # points: tag x y z nphys ...; others: tag 6 bounds nphys ...
                    k = 4 if dim == 0 else 7
                    nphys = int(f[k])
                    ent_phys[(dim, int(f[0]))] = [int(v) for v in f[k + 1:k + 1 + nphys]]
        elif tag == "$Nodes":
            nblocks, nnodes, _, _ = (int(v) for v in next(lines).split())
            tags = np.empty(nnodes, dtype=np.int64)
            coords = np.empty((nnodes, 3))
            k = 0
            for _ in range(nblocks):
                dim, etag, param, nb = (int(v) for v in next(lines).split())
                tags[k:k + nb] = [int(next(lines)) for _ in range(nb)]
                for q in range(nb):  # x y z (and u [v [w]] when parametric)
                    f = next(lines).split()
                    coords[k + q] = [float(f[0]), float(f[1]), float(f[2])]
                k += nb
            if k != nnodes:
                raise ValueError(f"$Nodes lists {k} nodes, its header {nnodes}")
        elif tag == "$Elements":
            nblocks, _, _, _ = (int(v) for v in next(lines).split())
            for _ in range(nblocks):
                dim, etag, etype, nb = (int(v) for v in next(lines).split())
                ptags = ent_phys.get((dim, etag)) or [0]
                for _ in range(nb):
                    f = [int(v) for v in next(lines).split()]
                    nodes = _element_nodes(etype, f[1:], f[0])
                    for pt in ptags:
                        conn.append(nodes)
                        phys.append(pt)
    return _finish(tags, coords, conn, phys, names)


def _getlabels(physical_names):
    """material, entity and boundary-condition labels of the names with three
    `_`-separated parts (Gmsh.jl:62-77)."""
    parts = [n.split("_") for n in physical_names]
    keep = [p for p in parts if len(p) == 3]
    return [p[0] for p in keep], [p[1] for p in keep], [p[2] for p in keep]


class MshFile:
    """Gmsh.jl:23-89: nodes, connectivity, physical index and the three label
    vectors.  Physical tags index the label vectors (1-based)."""

    def __init__(self, filename):
        self.filename = filename
        self.coords, self.connectivity, names, self.physical_index = read_msh(filename)
        if len(self.physical_index) != len(self.connectivity):
            raise AssertionError("The number of physical indexes must be equal to the number of "
                                 "elements.")
        self.material_labels, self.entities_labels, self.bc_labels = _getlabels(names)

    def _label(self, labels, k):
        return labels[self.physical_index[k] - 1]

    def material_label(self, k=None):
        return self.material_labels if k is None else self._label(self.material_labels, k)

    def entity_label(self, k=None):
        return self.entities_labels if k is None else self._label(self.entities_labels, k)

    def bc_label(self, k=None):
        return self.bc_labels if k is None else self._label(self.bc_labels, k)

    def dimension(self):
        return 3


class GmshMesh(Mesh):
    """Mesh built from an MshFile: the element, face and node sets of the
    reference's EntitySets, keyed by label (entity type, boundary condition and,
    for elements, material label)."""

    def __init__(self, coords, elements, faces, element_sets, face_sets, node_sets, msh):
        super().__init__(coords, elements, faces)
        self.element_sets, self.face_sets, self.node_sets = element_sets, face_sets, node_sets
        self.msh = msh

    def element_set(self, label):
        """The elements labelled `label` as one element set of their type."""
        kind, conn, proto = self.element_sets[label]
        if kind == "tet":
            return Tetrahedra(conn, label)
        return Trusses(conn, proto.cross_section, proto.strain_model, label)

    def face_set(self, label):
        return TriangularFaces(self.face_sets[label], label)

    def node_set(self, label):
        return NodeSet(self.node_sets[label], label)


def mesh_from_msh(msh: MshFile, entities):
    """Mesh(msh_file, entities) (StructuralEntities.jl:98-156).  `entities` maps
    an entity label of the file to a prototype: "tetrahedron"-like labels to
    Tetrahedra, truss labels to a Trusses prototype (its cross section and
    strain model are used), face labels to TriangularFaces.  Nodes labelled
    PHYSICAL_NODE_LABEL go to node sets under their boundary-condition label."""
    elems, faces = {}, {}
    el_sets, fa_sets, no_sets = {}, {}, {}

    def add(d, label, kind, proto, item):
        if label == "":
            return
        d.setdefault(label, (kind, proto, []))[2].append(item)

    for k, nodes in enumerate(msh.connectivity):
        elabel = msh.entity_label(k)
        bc = msh.bc_label(k)
        if elabel == PHYSICAL_NODE_LABEL:
            for lab in (elabel, bc):
                if lab:
                    no_sets.setdefault(lab, []).append(int(nodes[0]))
            continue
        if elabel not in entities:
            raise KeyError(f"entity label {elabel!r} has no entity type")
        proto = entities[elabel]
        if isinstance(proto, TriangularFaces) or proto is TriangularFaces:
            for lab in (elabel, bc):
                if lab:
                    fa_sets.setdefault(lab, []).append(nodes)
            faces.setdefault(elabel, []).append(nodes)
        else:
            kind = "tet" if (isinstance(proto, Tetrahedra) or proto is Tetrahedra) else "truss"
            for lab in (elabel, bc, msh.material_label(k)):
                add(el_sets, lab, kind, proto, nodes)
            elems.setdefault(elabel, (kind, proto, []))[2].append(nodes)
    elements = []
    for label, (kind, proto, conn) in elems.items():
        conn = np.array(conn, dtype=np.int64)
        if kind == "tet":
            elements.append(Tetrahedra(conn, label))
        else:
            elements.append(Trusses(conn, proto.cross_section, proto.strain_model, label))
    face_list = [TriangularFaces(np.array(c, dtype=np.int64), label) for label, c in faces.items()]
    element_sets = {label: (kind, np.array(conn, dtype=np.int64), proto)
                    for label, (kind, proto, conn) in el_sets.items()}
    face_sets = {label: np.array(c, dtype=np.int64) for label, c in fa_sets.items()}
    node_sets = {label: np.unique(np.array(c, dtype=np.int64)) for label, c in no_sets.items()}
    return GmshMesh(msh.coords, elements, face_list, element_sets, face_sets, node_sets, msh)


def write_msh(filename, coords, blocks):
    """Write an ASCII MSH 2.2 file.  blocks: [(physical name, element type
    (15/1/2/4), connectivity (k, n) 0-based)].  Used to hand meshes to tools
    that read Gmsh files and by the tests of read_msh."""
    names = [b[0] for b in blocks]
    out = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", str(len(names))]
    for t, (name, etype, _) in enumerate(blocks, start=1):
        dim = {15: 0, 1: 1, 2: 2, 4: 3}[etype]
        out.append(f'{dim} {t} "{name}"')
    out += ["$EndPhysicalNames", "$Nodes", str(len(coords))]
    out += [f"{i + 1} {float(x)!r} {float(y)!r} {float(z)!r}"
            for i, (x, y, z) in enumerate(np.asarray(coords, float))]
    out += ["$EndNodes", "$Elements"]
    rows = []
    for t, (name, etype, conn) in enumerate(blocks, start=1):
        for c in np.asarray(conn).reshape(len(conn), -1):
            rows.append(f"{len(rows) + 1} {etype} 2 {t} {t} " + " ".join(str(int(v) + 1) for v in c))
    out += [str(len(rows))] + rows + ["$EndElements"]
    with open(filename, "w", encoding="utf-8") as f:
        f.write("\n".join(out) + "\n")
    return filename
