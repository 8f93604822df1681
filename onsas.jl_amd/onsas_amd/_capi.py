"""ctypes binding of libonsas_hip.so (include/onsas_hip.h).

The library is loaded from this package directory.  There is no fallback: if
the shared object is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.environ.get("ONSAS_HIP_LIB") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "libonsas_hip.so")

# status codes (include/onsas_hip.h)
OK = 0
ERR_ARG, ERR_HIP, ERR_NEGATIVE_VOLUME, ERR_STATE = -1, -2, -3, -4
ERR_NOT_FINITE, ERR_LINEAR_SOLVE, ERR_ZERO_NORM, ERR_ALLOC = -5, -6, -7, -8
ERR_INJECTED, ERR_UNSUPPORTED = -9, -10

PRECISION_MIXED, PRECISION_FP64 = 0, 1
MAT_SVK, MAT_NEOHOOKEAN, MAT_LINEAR = 1, 2, 3
STRAIN_ROTATED, STRAIN_GREEN = 0, 1
LINSOLVE_AUTO, LINSOLVE_PCG, LINSOLVE_GMRES = 0, 1, 2
PRECOND_BLOCK_JACOBI, PRECOND_AMG = 0, 1
AMG_REFRESH_STEP, AMG_REFRESH_ASSEMBLY, AMG_REFRESH_ONCE = 0, 1, 2

# every symbol declared in include/onsas_hip.h
EXPORTED = (
    "onsas_create", "onsas_destroy", "onsas_last_error", "onsas_version", "onsas_set_nodes",
    "onsas_add_tet_group", "onsas_add_truss_group", "onsas_add_pressure_faces",
    "onsas_set_fixed_dofs", "onsas_finalize", "onsas_set_linear_solver", "onsas_set_loads",
    "onsas_assemble", "onsas_newton_step", "onsas_get_U", "onsas_snapshot_U",
    "onsas_read_snapshot", "onsas_set_U", "onsas_get_fext",
    "onsas_get_fint", "onsas_get_sizes", "onsas_gather_layout", "onsas_get_pattern",
    "onsas_get_tangent",
    "onsas_tangent_matvec", "onsas_tangent_matvec_noalloc", "onsas_element_forces",
    "onsas_get_stress_strain",
    "onsas_linear_solve", "onsas_adjoint_shear", "onsas_set_shear_moduli", "onsas_locate_points",
    "onsas_get_timings", "onsas_reset_timings", "onsas_inject_error", "onsas_amg_levels",
    "onsas_amg_level_info", "onsas_amg_get_operator", "onsas_amg_get_prolongator",
    "onsas_time_cycle_products", "onsas_time_level_products",
)


class LinearSettings(C.Structure):
    _fields_ = [("kind", C.c_int32), ("maxiter", C.c_int32), ("rtol", C.c_double),
                ("gmres_restart", C.c_int32), ("check_every", C.c_int32),
                ("forcing", C.c_int32), ("newton_tol", C.c_double), ("eta_max", C.c_double),
                ("eta_first", C.c_double), ("precond", C.c_int32),
                ("amg_refresh", C.c_int32)]


class StepStats(C.Structure):
    _fields_ = [("dU_norm", C.c_double), ("dU_rel", C.c_double), ("r_norm", C.c_double),
                ("r_rel", C.c_double), ("fext_norm", C.c_double), ("U_norm", C.c_double),
                ("lin_rel_residual", C.c_double), ("lin_iters", C.c_int32),
                ("lin_solver", C.c_int32), ("lin_converged", C.c_int32), ("reserved", C.c_int32),
                ("assemble_ms", C.c_double), ("solve_ms", C.c_double), ("lin_rtol", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class Timings(C.Structure):
    _fields_ = [("assemble_ms", C.c_double), ("scatter_ms", C.c_double), ("solve_ms", C.c_double),
                ("pcg_ms", C.c_double), ("n_assemble", C.c_int64), ("n_steps", C.c_int64),
                ("pcg_iters", C.c_int64), ("gmres_iters", C.c_int64),
                ("spmv_ms", C.c_double), ("spmv_launches", C.c_int64),
                ("elem_ms", C.c_double), ("assembly_mode", C.c_int64),
                ("cycle_spmv_ms", C.c_double), ("cycle_spmv_launches", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OnsasError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None
_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p


def load():
    """Load libonsas_hip.so (raises if it is absent: there is no CPU path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OnsasError(ERR_STATE, f"{LIB_PATH} is missing; build it with onsas.jl_amd/build.py")
    L = C.CDLL(LIB_PATH)
    sig = {
        "onsas_create": ([C.c_int32, C.POINTER(_vp)], C.c_int),
        "onsas_destroy": ([_vp], None),
        "onsas_last_error": ([_vp], C.c_char_p),
        "onsas_version": ([], C.c_char_p),
        "onsas_set_nodes": ([_vp, C.c_int64, _vp, _vp, _vp], C.c_int),
        "onsas_add_tet_group": ([_vp, C.c_int64, _vp, C.c_int32, _vp, _vp, _i32p], C.c_int),
        "onsas_add_truss_group": ([_vp, C.c_int64, _vp, C.c_int32, C.c_double, _vp, _i32p],
                                  C.c_int),
        "onsas_add_pressure_faces": ([_vp, C.c_int64, _vp, _i32p], C.c_int),
        "onsas_set_fixed_dofs": ([_vp, C.c_int64, _vp], C.c_int),
        "onsas_finalize": ([_vp], C.c_int),
        "onsas_set_linear_solver": ([_vp, C.POINTER(LinearSettings)], C.c_int),
        "onsas_set_loads": ([_vp, _vp, C.c_int32, _vp], C.c_int),
        "onsas_assemble": ([_vp, _vp], C.c_int),
        "onsas_newton_step": ([_vp, C.POINTER(StepStats)], C.c_int),
        "onsas_get_U": ([_vp, _vp], C.c_int),
        "onsas_snapshot_U": ([_vp, C.c_int32], C.c_int),
        "onsas_read_snapshot": ([_vp, C.c_int32, _vp], C.c_int),
        "onsas_set_U": ([_vp, _vp], C.c_int),
        "onsas_get_fext": ([_vp, _vp], C.c_int),
        "onsas_get_fint": ([_vp, _vp], C.c_int),
        "onsas_get_sizes": ([_vp, _i64p, _i64p, _i64p], C.c_int),
        "onsas_gather_layout": ([_vp, _i64p], C.c_int),
        "onsas_get_pattern": ([_vp, _vp, _vp], C.c_int),
        "onsas_get_tangent": ([_vp, _vp], C.c_int),
        "onsas_tangent_matvec": ([_vp, _vp, _vp], C.c_int),
        "onsas_tangent_matvec_noalloc": ([_vp, _vp, _vp], C.c_int),
        "onsas_element_forces": ([_vp, C.c_int32, _vp, _vp, _vp, _vp], C.c_int),
        "onsas_get_stress_strain": ([_vp, C.c_int32, C.c_int32, _vp, _vp], C.c_int),
        "onsas_linear_solve": ([_vp, _vp, _vp, _i32p, _dp], C.c_int),
        "onsas_adjoint_shear": ([_vp, C.c_int32, _vp, _vp, _i32p], C.c_int),
        "onsas_set_shear_moduli": ([_vp, C.c_int32, _vp], C.c_int),
        "onsas_locate_points": ([_vp, C.c_int32, C.c_int64, _vp, _vp, _vp, _vp, _vp], C.c_int),
        "onsas_get_timings": ([_vp, C.POINTER(Timings)], C.c_int),
        "onsas_reset_timings": ([_vp], C.c_int),
        "onsas_inject_error": ([_vp, C.c_int32], C.c_int),
        "onsas_amg_levels": ([_vp, _i32p], C.c_int),
        "onsas_amg_level_info": ([_vp, C.c_int32, _i64p, _i32p, _i64p, _i64p, _dp], C.c_int),
        "onsas_amg_get_operator": ([_vp, C.c_int32, _vp, _vp, _vp], C.c_int),
        "onsas_amg_get_prolongator": ([_vp, C.c_int32, _vp, _vp, _vp], C.c_int),
        "onsas_time_cycle_products": ([_vp, C.c_int32, C.POINTER(C.c_double)], C.c_int),
        "onsas_time_level_products": ([_vp, C.c_int32, C.c_int32, C.POINTER(C.c_double)],
                                      C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32).reshape(-1)


class Context:
    """Owner of one onsas_ctx (one HIP stream and its device memory)."""

    def __init__(self, precision=PRECISION_MIXED):
        self.L = load()
        h = _vp()
        st = self.L.onsas_create(int(precision), C.byref(h))
        if st != OK:
            raise OnsasError(st, self.L.onsas_last_error(None).decode())
        self.h = h
        self.precision = precision
        self.nn = 0
        self.ngroups = 0
        self.group_sizes = []
        self.group_nodes = []
        self.npressure = 0

    # -- plumbing
    def _check(self, st):
        if st != OK:
            raise OnsasError(st, self.L.onsas_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.onsas_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- model
    def set_nodes(self, coords):
        X = np.ascontiguousarray(np.asarray(coords, dtype=np.float64).reshape(-1, 3))
        x, y, z = (np.ascontiguousarray(X[:, d]) for d in range(3))
        self._check(self.L.onsas_set_nodes(self.h, X.shape[0], x.ctypes.data, y.ctypes.data,
                                           z.ctypes.data))
        self.nn = X.shape[0]

    def add_tet_group(self, conn, material, params, shear_per_elem=None):
        c = _i32(conn)
        p = _f64(params, 2)
        g = _f64(shear_per_elem) if shear_per_elem is not None else None
        gid = C.c_int32()
        self._check(self.L.onsas_add_tet_group(self.h, c.size // 4, c.ctypes.data, int(material),
                                               p.ctypes.data, _ptr(g), C.byref(gid)))
        self.group_sizes.append(c.size // 4)
        self.group_nodes.append(4)
        return gid.value

    def add_truss_group(self, conn, strain_model, E, area):
        c = _i32(conn)
        ne = c.size // 2
        a = np.broadcast_to(np.asarray(area, dtype=np.float64), (ne,)).copy()
        gid = C.c_int32()
        self._check(self.L.onsas_add_truss_group(self.h, ne, c.ctypes.data, int(strain_model),
                                                 float(E), a.ctypes.data, C.byref(gid)))
        self.group_sizes.append(ne)
        self.group_nodes.append(2)
        return gid.value

    def add_pressure_faces(self, conn):
        c = _i32(conn)
        gid = C.c_int32()
        self._check(self.L.onsas_add_pressure_faces(self.h, c.size // 3, c.ctypes.data,
                                                    C.byref(gid)))
        self.npressure += 1
        return gid.value

    def set_fixed_dofs(self, dofs):
        d = np.ascontiguousarray(dofs, dtype=np.int64).reshape(-1)
        self._check(self.L.onsas_set_fixed_dofs(self.h, d.size, d.ctypes.data))

    def finalize(self):
        self._check(self.L.onsas_finalize(self.h))

    def set_linear_solver(self, kind=LINSOLVE_AUTO, rtol=0.0, maxiter=0, gmres_restart=40,
                          check_every=16, forcing=False, newton_tol=0.0, eta_max=0.0,
                          eta_first=0.0, precond=PRECOND_BLOCK_JACOBI,
                          amg_refresh=AMG_REFRESH_STEP):
        s = LinearSettings(int(kind), int(maxiter), float(rtol), int(gmres_restart),
                           int(check_every), int(bool(forcing)), float(newton_tol),
                           float(eta_max), float(eta_first), int(precond), int(amg_refresh))
        self._check(self.L.onsas_set_linear_solver(self.h, C.byref(s)))

    # -- per step / iteration
    def set_loads(self, nodal=None, pressures=()):
        nd = None if nodal is None else _f64(nodal, 3 * self.nn)
        p = _f64(pressures) if len(pressures) else None
        self._check(self.L.onsas_set_loads(self.h, _ptr(nd), 0 if p is None else p.size, _ptr(p)))

    def assemble(self, want_fint=False):
        f = np.empty(3 * self.nn) if want_fint else None
        self._check(self.L.onsas_assemble(self.h, _ptr(f)))
        return f

    def newton_step(self):
        s = StepStats()
        self._check(self.L.onsas_newton_step(self.h, C.byref(s)))
        return s

    # -- state
    def _get(self, fn, n):
        out = np.empty(n)
        self._check(fn(self.h, out.ctypes.data))
        return out

    def get_U(self):
        return self._get(self.L.onsas_get_U, 3 * self.nn)

    def snapshot_U(self, slot):
        """Enqueue the copy of U into pinned snapshot `slot` (0 or 1); returns at once."""
        self._check(self.L.onsas_snapshot_U(self.h, slot))

    def read_snapshot(self, slot, out):
        """Wait for snapshot `slot` and copy it into `out` (3 nn float64).  Uses
        no stream: safe on another thread while this one drives the context."""
        if out.dtype != np.float64 or out.size != 3 * self.nn or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous float64 array of 3 nn values")
        rc = self.L.onsas_read_snapshot(self.h, slot, out.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"onsas_read_snapshot failed ({rc})")
        return out

    def set_U(self, U):
        u = _f64(U, 3 * self.nn)
        self._check(self.L.onsas_set_U(self.h, u.ctypes.data))

    def get_fext(self):
        return self._get(self.L.onsas_get_fext, 3 * self.nn)

    def get_fint(self):
        return self._get(self.L.onsas_get_fint, 3 * self.nn)

    def sizes(self):
        nn, nnzb, nc = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self.L.onsas_get_sizes(self.h, C.byref(nn), C.byref(nnzb), C.byref(nc)))
        return nn.value, nnzb.value, nc.value

    def gather_layout(self):
        """Staged-gather layout: rows per workgroup (0: not staged), workgroups,
        staged elements, stored contribution entries, LDS bytes per workgroup,
        steps."""
        out = (C.c_int64 * 6)()
        self._check(self.L.onsas_gather_layout(self.h, out))
        return dict(zip(("rows", "groups", "staged", "entries", "lds_bytes", "steps"), list(out)))

    def pattern(self):
        nn, nnzb, _ = self.sizes()
        rp = np.empty(nn + 1, dtype=np.int32)
        col = np.empty(nnzb, dtype=np.int32)
        self._check(self.L.onsas_get_pattern(self.h, rp.ctypes.data, col.ctypes.data))
        return rp, col

    # -- multigrid hierarchy (after a PCG solve with PRECOND_AMG)
    def amg_levels(self):
        n = C.c_int32()
        self._check(self.L.onsas_amg_levels(self.h, C.byref(n)))
        return n.value

    def amg_level_info(self, level):
        n, nnzb, nnzp = C.c_int64(), C.c_int64(), C.c_int64()
        bs, lam = C.c_int32(), C.c_double()
        self._check(self.L.onsas_amg_level_info(self.h, level, C.byref(n), C.byref(bs), C.byref(nnzb),
                                                C.byref(nnzp), C.byref(lam)))
        return dict(n=n.value, bs=bs.value, nnzb=nnzb.value, nnz_p=nnzp.value, lam=lam.value)

    def amg_operator(self, level):
        """Level >= 1 operator as a scipy BSR matrix (6x6 blocks, FP64 copy)."""
        import scipy.sparse as sp
        info = self.amg_level_info(level)
        rp = np.empty(info["n"] + 1, dtype=np.int32)
        col = np.empty(info["nnzb"], dtype=np.int32)
        vals = np.empty(36 * info["nnzb"])
        self._check(self.L.onsas_amg_get_operator(self.h, level, rp.ctypes.data, col.ctypes.data,
                                                  vals.ctypes.data))
        m = 6 * info["n"]
        return sp.bsr_matrix((vals.reshape(-1, 6, 6), col, rp), shape=(m, m))

    def amg_prolongator(self, level):
        """Prolongator from level+1 to level as a scipy BSR matrix (bs x 6 blocks)."""
        import scipy.sparse as sp
        info = self.amg_level_info(level)
        nc = self.amg_level_info(level + 1)["n"]
        bs = info["bs"]
        rp = np.empty(info["n"] + 1, dtype=np.int32)
        col = np.empty(info["nnz_p"], dtype=np.int32)
        vals = np.empty(bs * 6 * info["nnz_p"])
        self._check(self.L.onsas_amg_get_prolongator(self.h, level, rp.ctypes.data, col.ctypes.data,
                                                     vals.ctypes.data))
        return sp.bsr_matrix((vals.reshape(-1, bs, 6), col, rp), shape=(bs * info["n"], 6 * nc))

    def time_cycle_products(self, launches=50):
        """ms per launch of the multigrid cycle's fine-level tangent products,
        `launches` back-to-back launches between two events on the stream."""
        ms = C.c_double(0)
        self._check(self.L.onsas_time_cycle_products(self.h, int(launches), C.byref(ms)))
        return ms.value

    def time_level_products(self, level, launches=50):
        """ms per launch of a multigrid level's cycle products (level 0: the two
        fine-level ones; a sliced-ELL level >= 1: its four, in cycle order),
        `launches` back-to-back launches between two events on the stream."""
        ms = C.c_double(0)
        self._check(self.L.onsas_time_level_products(self.h, int(level), int(launches),
                                                     C.byref(ms)))
        return ms.value

    def tangent_bsr(self):
        """The assembled tangent as a scipy BSR matrix (3x3 blocks, FP64 copy)."""
        import scipy.sparse as sp
        rp, col = self.pattern()
        vals = np.empty(9 * col.size)
        self._check(self.L.onsas_get_tangent(self.h, vals.ctypes.data))
        n = 3 * (rp.size - 1)
        return sp.bsr_matrix((vals.reshape(-1, 3, 3), col, rp), shape=(n, n))

    def matvec(self, x):
        xv = _f64(x, 3 * self.nn)
        y = np.empty_like(xv)
        self._check(self.L.onsas_tangent_matvec(self.h, xv.ctypes.data, y.ctypes.data))
        return y

    def matvec_noalloc(self, x):
        """K x through onsas_tangent_matvec_noalloc (no device allocation; at
        most 2,097,152 dofs; calls overlapping on one GPU race)."""
        xv = _f64(x, 3 * self.nn)
        y = np.empty_like(xv)
        self._check(self.L.onsas_tangent_matvec_noalloc(self.h, xv.ctypes.data, y.ctypes.data))
        return y

    def element_forces(self, group, want_K=True):
        ne, k = self.group_sizes[group], self.group_nodes[group]
        nd = 3 * k
        f = np.empty(ne * nd)
        K = np.empty(ne * nd * nd) if want_K else None
        s = np.empty(ne * 9)
        e = np.empty(ne * 9)
        self._check(self.L.onsas_element_forces(self.h, group, f.ctypes.data, _ptr(K),
                                                s.ctypes.data, e.ctypes.data))
        return (f.reshape(ne, nd), None if K is None else K.reshape(ne, nd, nd),
                s.reshape(ne, 3, 3), e.reshape(ne, 3, 3))

    def stress_strain(self, group, at_last_assembly=True):
        ne = self.group_sizes[group]
        s = np.empty(ne * 9)
        e = np.empty(ne * 9)
        self._check(self.L.onsas_get_stress_strain(self.h, group, int(bool(at_last_assembly)),
                                                   s.ctypes.data, e.ctypes.data))
        return s.reshape(ne, 3, 3), e.reshape(ne, 3, 3)

    def linear_solve(self, b):
        bv = _f64(b, 3 * self.nn)
        x = np.empty_like(bv)
        it, rr = C.c_int32(), C.c_double()
        self._check(self.L.onsas_linear_solve(self.h, bv.ctypes.data, x.ctypes.data, C.byref(it),
                                              C.byref(rr)))
        return x, it.value, rr.value

    def adjoint_shear(self, group, dJdU):
        b = _f64(dJdU, 3 * self.nn)
        out = np.empty(self.group_sizes[group])
        it = C.c_int32()
        self._check(self.L.onsas_adjoint_shear(self.h, group, b.ctypes.data, out.ctypes.data,
                                               C.byref(it)))
        return out, it.value

    def set_shear_moduli(self, group, G):
        g = _f64(G, self.group_sizes[group])
        self._check(self.L.onsas_set_shear_moduli(self.h, group, g.ctypes.data))

    def locate_points(self, group, points):
        P = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
        n = P.shape[0]
        x, y, z = (np.ascontiguousarray(P[:, d]) for d in range(3))
        elem = np.empty(n, dtype=np.int64)
        w = np.empty(4 * n)
        self._check(self.L.onsas_locate_points(self.h, group, n, x.ctypes.data, y.ctypes.data,
                                               z.ctypes.data, elem.ctypes.data, w.ctypes.data))
        return elem, w.reshape(n, 4)

    def timings(self):
        t = Timings()
        self._check(self.L.onsas_get_timings(self.h, C.byref(t)))
        return t.as_dict()

    def reset_timings(self):
        self._check(self.L.onsas_reset_timings(self.h))

    def inject_error(self, code):
        self._check(self.L.onsas_inject_error(self.h, int(code)))
