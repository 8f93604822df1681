"""Staged gather assembly (each workgroup copies the states of the elements of
its rows into LDS and sums the blocks from there) against the list-driven
gather that reads every contribution from HBM.  Both sum the same values in
the same order: f_int (FP64 sums of the element forces) agrees bit for bit;
the tangent blocks agree to a few units in the last place of the largest block
entry, as the compiler fuses the multiply-adds of the two kernels differently
(HIP device code contracts across statements); the block-Jacobi inverse is
seen through a PCG solve.  Checked for every rows-per-workgroup choice, both
precisions, the three tet materials, on a structured mesh (the gather order's
translate classes) and on a scrambled one (node and element numbering
shuffled).  Both are also checked against the oracle in test_gpu_parity.py
(the list-driven gather is the default)."""

# tangent agreement, relative to the largest entry: a few ulps of FP32 / FP64
K_TOL = {"mixed": 1e-6, "fp64": 1e-14}
# PCG solutions on the two tangents (relative 2-norm)
X_TOL = {"mixed": 1e-3, "fp64": 1e-9}


def _same(prec, out, ref):
    np.testing.assert_array_equal(out["f"], ref["f"])
    np.testing.assert_allclose(out["K"], ref["K"], rtol=0,
                               atol=K_TOL[prec] * np.abs(ref["K"]).max())
    assert (np.linalg.norm(out["x"] - ref["x"]) <= X_TOL[prec] * np.linalg.norm(ref["x"]))
    assert abs(out["it"] - ref["it"]) <= 2 + 0.03 * ref["it"]
import contextlib
import os

import numpy as np
import pytest

from onsas_amd import _capi, meshes, workloads

pytestmark = pytest.mark.gpu

PREC = {"mixed": _capi.PRECISION_MIXED, "fp64": _capi.PRECISION_FP64}
MATS = {"svk": (_capi.MAT_SVK, (0.6e6, 0.4e6)), "nh": (_capi.MAT_NEOHOOKEAN, (1.3e6, 0.4e6)),
        "linear": (_capi.MAT_LINEAR, (1.0e6, 0.3))}


@contextlib.contextmanager
def _env(**kw):
    """The gather settings are read when a context is finalized."""
    old = {k: os.environ.get(k) for k in kw}
    for k, v in kw.items():
        os.environ[k] = str(v)
    try:
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _mesh(kind):
    sm = meshes.cube(9, jitter=0.1, seed=11)
    coords, tets = sm.coords, sm.tets
    if kind == "scrambled":
        rng = np.random.default_rng(4)
        perm = rng.permutation(len(coords))  # new id of node i
        inv = np.argsort(perm)
        coords = coords[inv]
        tets = perm[tets][rng.permutation(len(tets))].astype(np.int32)
    fixed = np.flatnonzero(np.repeat(coords[:, 2] == coords[:, 2].min(), 3))
    return coords, tets, fixed


def _run(prec, mat, coords, tets, fixed, U, b, **env):
    with _env(**env):
        c = _capi.Context(PREC[prec])
        c.set_nodes(coords)
        code, p = MATS[mat]
        c.add_tet_group(tets, code, p, None)
        c.set_fixed_dofs(fixed)
        c.finalize()
    try:
        c.set_U(U)
        f = c.assemble(want_fint=True)
        K = c.tangent_bsr().data.copy()
        x, it, _ = c.linear_solve(b)
        return dict(f=f, K=K, x=x, it=it, layout=c.gather_layout(),
                    mode=c.timings()["assembly_mode"])
    finally:
        c.close()


@pytest.mark.parametrize("mesh", ["cube", "scrambled"])
@pytest.mark.parametrize("mat", ["svk", "nh", "linear"])
@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_staged_gather_matches_lists(prec, mat, mesh):
    print(f"[gpu] staged gather vs lists {prec} {mat} {mesh}", flush=True)
    coords, tets, fixed = _mesh(mesh)
    rng = np.random.default_rng(6)
    # small strains: a positive definite tangent, so the PCG counts are stable
    U = rng.normal(scale=1e-4, size=coords.size)
    U[fixed] = 0
    b = rng.normal(size=coords.size)
    ref = _run(prec, mat, coords, tets, fixed, U, b)  # the default: lists
    assert ref["mode"] == 1 and ref["layout"]["rows"] == 0
    staged = 0
    for rows in (64, 32, 16, 8):
        out = _run(prec, mat, coords, tets, fixed, U, b, ONSAS_GATHER="staged",
                   ONSAS_GATHER_ROWS=rows, ONSAS_GATHER_LDS_KB=160)
        lay = out["layout"]
        if lay["rows"] == 0:  # this row count's largest element list exceeds the LDS
            assert out["mode"] == 1
            continue
        staged += 1
        print(f"  rows {rows}: {lay}", flush=True)
        assert out["mode"] == 2 and lay["rows"] == rows and lay["lds_bytes"] <= 160 * 1024
        _same(prec, out, ref)
    assert staged >= 1


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_staged_gather_on_compression_cube(prec):
    """The staged gather's own choice of rows (largest row count within
    64 KiB of LDS) on the 32^3 compression cube at a nonuniform state matches
    the lists, twice in a row (the LDS and the step table are rebuilt per
    launch)."""
    print(f"[gpu] staged gather own rows 32^3 {prec}", flush=True)
    s, sm, p = workloads.cube_compression(n=32)
    (mat, es), = list(s.materials)
    rng = np.random.default_rng(1)
    U = rng.normal(scale=1e-3, size=3 * sm.nn)
    U[s.fixed_dofs] = 0
    res = {}
    for mode in ("lists", "staged"):
        with _env(ONSAS_GATHER=mode):
            c = _capi.Context(PREC[prec])
            c.set_nodes(sm.coords)
            c.add_tet_group(es.conn, mat.kind, mat.params(), None)
            c.set_fixed_dofs(s.fixed_dofs)
            c.finalize()
        try:
            c.set_U(U)
            f1 = c.assemble(want_fint=True)
            K1 = c.tangent_bsr().data.copy()
            c.set_U(0.5 * U)
            f2 = c.assemble(want_fint=True)
            K2 = c.tangent_bsr().data.copy()
            res[mode] = (f1, K1, f2, K2, c.gather_layout())
        finally:
            c.close()
    lay = res["staged"][4]
    assert lay["rows"] in (64, 32, 16, 8) and lay["lds_bytes"] <= 64 * 1024
    assert lay["staged"] >= lay["groups"]
    for k in (0, 2):
        np.testing.assert_array_equal(res["staged"][k], res["lists"][k])
    for k in (1, 3):
        ref = res["lists"][k]
        np.testing.assert_allclose(res["staged"][k], ref, rtol=0,
                                   atol=K_TOL[prec] * np.abs(ref).max())
