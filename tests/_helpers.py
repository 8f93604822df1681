"""Oracle-side helpers shared by the tests: reference load vectors and the
oracle model of a Structure built with the onsas_amd API."""
import numpy as np


def fext_steps(oracle, structure, lfs):
    """Full external-force vector at each load factor, as the reference
    assembles it: GlobalLoad contributions plus p(t) (-n) A / 3 per pressure
    face node from the oracle's face routine, duplicate dofs summed."""
    n = structure.num_dofs()
    X = structure.mesh.coords
    out = []
    for t in lfs:
        f = structure.bcs.nodal_loads(t, X, n)
        for bc, conn in structure.bcs.pressure_groups():
            for tri in conn:
                fe = oracle.face_pressure(X[tri].reshape(-1), bc(t))
                for a in range(3):
                    f[3 * tri[a]:3 * tri[a] + 3] += fe[a]
        out.append(f)
    return np.array(out)


def oracle_model(oracle, structure):
    """oracle.Model of a single-material Structure."""
    from onsas_amd.entities import Tetrahedra
    from onsas_amd.materials import IsotropicLinearElastic, NeoHookean, SVK
    (mat, es), = list(structure.materials)
    if isinstance(es, Tetrahedra):
        code = {SVK: oracle.MAT_SVK, NeoHookean: oracle.MAT_NEOHOOKEAN,
                IsotropicLinearElastic: oracle.MAT_LINEAR}[type(mat)]
        G = getattr(mat, "G", None)
        kw = dict(tets=es.conn, tet_mat=code, tet_p=(mat.params()[0], float(np.median(G))
                                                     if np.ndim(G) else mat.params()[1]),
                  tet_G=G if np.ndim(G) else None)
    else:
        kw = dict(trusses=es.conn, truss_strain=es.strain_model.code,
                  truss_E=mat.elasticity_modulus(), truss_A=es.areas())
    return oracle.Model(structure.mesh.coords, fixed_dofs=structure.fixed_dofs, **kw)
