"""ctypes wrapper of the FP64 CPU oracle (oracle/onsas_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke`` and
by the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product
package never imports it.

Each wrapper names the ONSAS.jl function it restates (paths relative to the
ONSAS.jl v0.4.6 source tree).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libonsas_oracle.so")
# the same source built with OpenMP: the all-cores CPU baseline only (its
# reductions' order depends on the thread count, so it is not the checker)
_LIB_OMP_PATH = os.path.join(_HERE, "_build", "libonsas_oracle_omp.so")

MAT_SVK, MAT_NEOHOOKEAN, MAT_LINEAR = 1, 2, 3
STRAIN_ROTATED, STRAIN_GREEN = 0, 1

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class _Model(C.Structure):
    _fields_ = [
        ("nn", C.c_int64), ("X", C.c_void_p),
        ("ntet", C.c_int64), ("tet", C.c_void_p), ("tet_mat", C.c_int),
        ("tet_p0", C.c_double), ("tet_p1", C.c_double), ("tet_G", C.c_void_p),
        ("ntruss", C.c_int64), ("truss", C.c_void_p), ("truss_strain", C.c_int),
        ("truss_E", C.c_double), ("truss_A", C.c_void_p),
        ("nfree", C.c_int64), ("free_dofs", C.c_void_p),
    ]


class _CSR(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64), ("rowptr", C.POINTER(C.c_int64)),
                ("col", C.POINTER(C.c_int32)), ("val", C.POINTER(C.c_double))]


class StepRecord(C.Structure):
    _fields_ = [("dU_norm", C.c_double), ("dU_rel", C.c_double), ("r_norm", C.c_double),
                ("r_rel", C.c_double), ("iters", C.c_int32), ("criterion", C.c_int32),
                ("cg_iters", C.c_int64)]


_libs = {}


def build():
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(omp=False):
    """The serial oracle (the checker), or with omp=True its OpenMP build."""
    if omp not in _libs:
        path = _LIB_OMP_PATH if omp else _LIB_PATH
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_svk_stress.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp]
        L.or_svk_energy.argtypes = [C.c_double, C.c_double, _dp]
        L.or_svk_energy.restype = C.c_double
        L.or_nh_stress.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp]
        L.or_nh_energy.argtypes = [C.c_double, C.c_double, _dp]
        L.or_nh_energy.restype = C.c_double
        L.or_linear_stress.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp]
        L.or_tet_volume.argtypes = [_dp, C.POINTER(C.c_double)]
        L.or_tet_gradients.argtypes = [_dp, _dp, C.POINTER(C.c_double)]
        L.or_tet_hyperelastic.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                                          _dp, _dp]
        L.or_tet_linear.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp]
        L.or_truss.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_face_pressure.argtypes = [_dp, C.c_double, _dp]
        L.or_assemble.argtypes = [C.POINTER(_Model), _dp, _dp, C.POINTER(_CSR),
                                  C.POINTER(_CSR), C.c_void_p, C.c_void_p]
        L.or_csr_free.argtypes = [C.POINTER(_CSR)]
        L.or_coo_to_csr.argtypes = [C.c_int64, _i32p, _i32p, _dp, C.c_int64, C.POINTER(_CSR)]
        L.or_cg.argtypes = [C.POINTER(_CSR), _dp, _dp, C.c_double, C.c_double, C.c_int64]
        L.or_cg.restype = C.c_int64
        L.or_newton_static.argtypes = [C.POINTER(_Model), C.c_int, _dp, C.c_double, C.c_double,
                                       C.c_int, C.c_double, _dp, C.c_void_p,
                                       C.POINTER(StepRecord)]
        _libs[omp] = L
    return _libs[omp]


class OracleError(RuntimeError):
    pass


def _check(st):
    if st != 0:
        raise OracleError(f"oracle status {st}")


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


# ---- materials ---------------------------------------------------------------

def svk_stress(lam, G, E):
    """cosserat_stress!(::SVK) -- src/Materials/SVKMaterial.jl:89-100."""
    S, D = np.zeros(9), np.zeros(36)
    lib().or_svk_stress(lam, G, _f64(E, 9), S, D)
    return S.reshape(3, 3), D.reshape(6, 6)


def svk_energy(lam, G, E):
    return lib().or_svk_energy(lam, G, _f64(E, 9))


def nh_stress(K, G, E):
    """cosserat_stress!(::NeoHookean) -- src/Materials/NeoHookeanMaterial.jl:94-138."""
    S, D = np.zeros(9), np.zeros(36)
    lib().or_nh_stress(K, G, _f64(E, 9), S, D)
    return S.reshape(3, 3), D.reshape(6, 6)


def nh_energy(K, G, E):
    return lib().or_nh_energy(K, G, _f64(E, 9))


def linear_stress(Emod, nu, eps):
    """stress!(::IsotropicLinearElastic) -- IsotropicLinearElasticMaterial.jl:80-92."""
    S, D = np.zeros(9), np.zeros(36)
    lib().or_linear_stress(Emod, nu, _f64(eps, 9), S, D)
    return S.reshape(3, 3), D.reshape(6, 6)


# ---- elements ------------------------------------------------------------------

def tet_volume(X):
    v = C.c_double()
    st = lib().or_tet_volume(_f64(X, 12), C.byref(v))
    return v.value, st


def tet_gradients(X):
    dN, v = np.zeros(12), C.c_double()
    _check(lib().or_tet_gradients(_f64(X, 12), dN, C.byref(v)))
    return dN.reshape(4, 3), v.value


def tet_hyperelastic(mat, p0, p1, X, u):
    """internal_forces(::AbstractHyperElasticMaterial, ::Tetrahedron) -- Tetrahedrons.jl:186-224.

    Returns (f[12], K[12,12], P[3,3], C[3,3])."""
    f, K, P, Cm = np.zeros(12), np.zeros(144), np.zeros(9), np.zeros(9)
    _check(lib().or_tet_hyperelastic(mat, p0, p1, _f64(X, 12), _f64(u, 12), f, K, P, Cm))
    return f, K.reshape(12, 12), P.reshape(3, 3), Cm.reshape(3, 3)


def tet_linear(Emod, nu, X, u):
    """internal_forces(::IsotropicLinearElastic, ::Tetrahedron) -- Tetrahedrons.jl:240-266."""
    f, K, s, e = np.zeros(12), np.zeros(144), np.zeros(9), np.zeros(9)
    _check(lib().or_tet_linear(Emod, nu, _f64(X, 12), _f64(u, 12), f, K, s, e))
    return f, K.reshape(12, 12), s.reshape(3, 3), e.reshape(3, 3)


def truss(strain_model, Emod, A, X, u):
    """internal_forces(::AbstractHyperElasticMaterial, ::Truss) -- Trusses.jl:126-184."""
    f, K = np.zeros(6), np.zeros(36)
    s, e = C.c_double(), C.c_double()
    _check(lib().or_truss(strain_model, Emod, A, _f64(X, 6), _f64(u, 6), f, K, C.byref(s),
                          C.byref(e)))
    return f, K.reshape(6, 6), s.value, e.value


def face_pressure(X, p):
    """apply(::Pressure, ::TriangularFace, t) -- LocalLoadBoundaryConditions.jl:36-56."""
    f = np.zeros(9)
    lib().or_face_pressure(_f64(X, 9), float(p), f)
    return f.reshape(3, 3)


# ---- global model --------------------------------------------------------------

class Model:
    """Flat description of a structure for the oracle's global routines."""

    def __init__(self, X, tets=None, tet_mat=MAT_SVK, tet_p=(0.0, 0.0), tet_G=None,
                 trusses=None, truss_strain=STRAIN_ROTATED, truss_E=0.0, truss_A=None,
                 fixed_dofs=(), omp=False):
        self.L = lib(omp)  # omp: the OpenMP build (CPU baseline timing only)
        self.X = _f64(X)
        self.nn = self.X.size // 3
        self.n = 3 * self.nn
        self.tets = np.ascontiguousarray(tets if tets is not None else np.zeros((0, 4)),
                                         dtype=np.int32).reshape(-1)
        self.trusses = np.ascontiguousarray(
            trusses if trusses is not None else np.zeros((0, 2)), dtype=np.int32).reshape(-1)
        ntr = self.trusses.size // 2
        self.truss_A = _f64(truss_A if truss_A is not None else np.zeros(ntr))
        if np.ndim(self.truss_A) == 0 or self.truss_A.size == 1 and ntr != 1:
            self.truss_A = np.full(ntr, float(self.truss_A.reshape(-1)[0]) if ntr else 0.0)
        self.tet_G = None if tet_G is None else _f64(tet_G)
        fixed = np.unique(np.asarray(fixed_dofs, dtype=np.int64))
        mask = np.ones(self.n, dtype=bool)
        mask[fixed] = False
        self.free = np.ascontiguousarray(np.nonzero(mask)[0], dtype=np.int64)
        self._m = _Model(
            self.nn, self.X.ctypes.data,
            self.tets.size // 4, self.tets.ctypes.data, tet_mat, float(tet_p[0]),
            float(tet_p[1]), None if self.tet_G is None else self.tet_G.ctypes.data,
            ntr, self.trusses.ctypes.data, truss_strain, float(truss_E),
            self.truss_A.ctypes.data, self.free.size, self.free.ctypes.data)

    @property
    def ntet(self):
        return self.tets.size // 4

    @property
    def ntruss(self):
        return self.trusses.size // 2

    def assemble(self, U, want_full=True, want_ff=False, want_stress=False):
        """assemble!(s, sa) -- src/StructuralAnalyses/StaticAnalyses.jl:99-122."""
        import scipy.sparse as sp
        U = _f64(U, self.n)
        fint = np.zeros(self.n)
        Kf, Kff = _CSR(), _CSR()
        ne = self.ntet + self.ntruss
        stress = np.zeros(9 * ne) if want_stress else None
        strain = np.zeros(9 * ne) if want_stress else None
        st = self.L.or_assemble(C.byref(self._m), U, fint,
                               C.byref(Kf) if want_full else None,
                               C.byref(Kff) if want_ff else None,
                               stress.ctypes.data if want_stress else None,
                               strain.ctypes.data if want_stress else None)
        _check(st)
        out = {"fint": fint}
        for name, M, w in (("K", Kf, want_full), ("Kff", Kff, want_ff)):
            if not w:
                continue
            n, nnz = M.n, M.nnz
            rp = np.ctypeslib.as_array(M.rowptr, shape=(n + 1,)).copy()
            col = np.ctypeslib.as_array(M.col, shape=(max(nnz, 1),))[:nnz].copy()
            val = np.ctypeslib.as_array(M.val, shape=(max(nnz, 1),))[:nnz].copy()
            self.L.or_csr_free(C.byref(M))
            out[name] = sp.csr_matrix((val, col, rp), shape=(n, n))
        if want_stress:
            out["stress"] = stress.reshape(ne, 3, 3)
            out["strain"] = strain.reshape(ne, 3, 3)
        return out

    def newton(self, fext_steps, rel_U_tol=1e-6, rel_res_tol=1e-6, max_iter=20,
               cg_reltol=-1.0, U0=None, keep_steps=True):
        """_solve!(::NonLinearStaticAnalysis, ::NewtonRaphson) --
        src/StructuralAnalyses/NonLinearStaticAnalyses.jl:70-148."""
        F = np.ascontiguousarray(fext_steps, dtype=np.float64)
        nsteps = F.shape[0]
        U = np.zeros(self.n) if U0 is None else _f64(U0, self.n).copy()
        Us = np.zeros((nsteps, self.n)) if keep_steps else None
        recs = (StepRecord * nsteps)()
        st = self.L.or_newton_static(C.byref(self._m), nsteps, F.reshape(-1), rel_U_tol,
                                    rel_res_tol, max_iter, cg_reltol, U,
                                    Us.ctypes.data if keep_steps else None, recs)
        _check(st)
        return U, Us, [dict((k, getattr(r, k)) for k, _ in StepRecord._fields_) for r in recs]


def _csr_to_scipy(M):
    import scipy.sparse as sp
    n, nnz = M.n, M.nnz
    rp = np.ctypeslib.as_array(M.rowptr, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(M.col, shape=(max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(M.val, shape=(max(nnz, 1),))[:nnz].copy()
    lib().or_csr_free(C.byref(M))
    return sp.csr_matrix((val, col, rp), shape=(n, n))


def coo_sum(I, J, V, n):
    """end_assemble!: K[I[k], J[k]] += V[k] in triplet order (Assemblers.jl:84-88), 0-based."""
    I = np.ascontiguousarray(I, dtype=np.int32)
    J = np.ascontiguousarray(J, dtype=np.int32)
    V = _f64(V)
    M = _CSR()
    _check(lib().or_coo_to_csr(V.size, I, J, V, n, C.byref(M)))
    return _csr_to_scipy(M)


def cg(A, b, reltol=1.4901161193847656e-08, abstol=0.0, maxiter=None, omp=False):
    """IterativeSolvers 0.9.4 cg! on a scipy CSR matrix (x0 = 0, no preconditioner)."""
    A = A.tocsr()
    n = A.shape[0]
    rp = np.ascontiguousarray(A.indptr, dtype=np.int64)
    col = np.ascontiguousarray(A.indices, dtype=np.int32)
    val = np.ascontiguousarray(A.data, dtype=np.float64)
    M = _CSR(n, val.size, rp.ctypes.data_as(C.POINTER(C.c_int64)),
             col.ctypes.data_as(C.POINTER(C.c_int32)), val.ctypes.data_as(C.POINTER(C.c_double)))
    x = np.zeros(n)
    it = lib(omp).or_cg(C.byref(M), _f64(b, n), x, abstol, reltol,
                        n if maxiter is None else maxiter)
    return x, it
