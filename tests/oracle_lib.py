"""ctypes binding of the CPU oracle (oracle/pd_oracle.h) for the tests.

Test infrastructure only: the oracle is the checker, never the thing measured
or shipped.  Built by ``make -C oracle`` (``__graft_entry__.build()`` does it).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the default build; PD_ORACLE_LIB names another build of it (the sanitized one)
ORACLE_SO = os.environ.get("PD_ORACLE_LIB") or os.path.join(ROOT, "oracle", "build", "liboracle.so")
REF_MATRIX_SO = os.path.join(ROOT, "oracle", "_ref", "libref_matrix.so")


class OrParams(C.Structure):
    _fields_ = [
        ("n1", C.c_double), ("n2", C.c_double), ("horizon", C.c_double),
        ("dx", C.c_double), ("rho", C.c_double), ("dt", C.c_double),
        ("E", C.c_double), ("nu", C.c_double), ("s0", C.c_double),
        ("kappa_tol", C.c_double),
    ]


class DamageParams(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("kappa0", C.c_double), ("At", C.c_double), ("Bt", C.c_double),
        ("Ac", C.c_double), ("Bc", C.c_double),
    ]


DAMAGE_REFERENCE = 1
DAMAGE_MAZARS = 2


def damage_params(mode=DAMAGE_REFERENCE, kappa0=3.0 / 3.5e4, At=0.87, Bt=2.0e4, Ac=0.65,
                  Bc=2150.0):
    """Defaults are the reference's (src/main_seq.cpp:480, :660)."""
    return DamageParams(mode, kappa0, At, Bt, Ac, Bc)


class RefParams(C.Structure):
    _fields_ = [
        ("n1", C.c_double), ("n2", C.c_double), ("horizon", C.c_double),
        ("dx", C.c_double), ("E", C.c_double), ("nu", C.c_double),
        ("rho", C.c_longdouble), ("dt", C.c_longdouble),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        P = C.c_void_p
        L.or_weight.restype = C.c_double
        L.or_weight.argtypes = [C.POINTER(OrParams)] + [C.c_double] * 4
        L.or_build_rev.restype = C.c_int
        L.or_build_rev.argtypes = [C.c_int, P, P, P]
        L.or_forces_fp64.restype = C.c_long
        L.or_forces_fp64.argtypes = [C.POINTER(OrParams), C.c_int] + [P] * 8 + [C.c_int, P, P, P, C.c_int]
        L.or_bond_stretch.restype = C.c_double
        L.or_bond_stretch.argtypes = [C.c_double] * 4
        L.or_bond_breaks.restype = C.c_int
        L.or_bond_breaks.argtypes = [C.c_double] * 5
        L.or_break_bonds.restype = C.c_long
        L.or_break_bonds.argtypes = [C.POINTER(OrParams), C.c_int, P, P, P, P, P]
        L.or_step_fp64.restype = C.c_long
        L.or_step_fp64.argtypes = [C.POINTER(OrParams), C.c_int] + [P] * 13 + [C.c_int, C.c_int]
        L.or_sym2_eig.restype = None
        L.or_sym2_eig.argtypes = [C.c_double] * 3 + [P, P]
        L.or_mazars_update.restype = C.c_int
        L.or_mazars_update.argtypes = [C.POINTER(DamageParams)] + [C.c_double] * 3 + [P, P]
        L.or_forces_damage_fp64.restype = C.c_long
        L.or_forces_damage_fp64.argtypes = ([C.POINTER(OrParams), C.POINTER(DamageParams), C.c_int]
                                            + [P] * 13 + [C.c_int])
        L.or_step_damage_fp64.restype = C.c_long
        L.or_step_damage_fp64.argtypes = ([C.POINTER(OrParams), C.POINTER(DamageParams), C.c_int]
                                          + [P] * 16 + [C.c_int, C.c_int])
        L.ref_cpu_compute_velocity.restype = None
        L.ref_cpu_compute_velocity.argtypes = [C.POINTER(RefParams), C.c_int] + [P] * 8 + [C.c_int] * 3
        L.ref_cpu_update_positions.restype = None
        L.ref_cpu_update_positions.argtypes = [C.POINTER(RefParams), C.c_int, P, P, P, P]
        L.ref_cpu_apply_velocity_bc.restype = None
        L.ref_cpu_apply_velocity_bc.argtypes = [C.c_int, P, P, P, P]
        L.ref_cpu_apply_fixed_bc.restype = None
        L.ref_cpu_apply_fixed_bc.argtypes = [C.c_int, P, P, P]
        L.ref_find_neighbor.restype = C.c_long
        L.ref_find_neighbor.argtypes = [C.c_int, P, C.c_longdouble, C.c_longdouble, P, P, C.c_long]
        L.ref_mat_inverse2d.restype = None
        L.ref_mat_inverse2d.argtypes = [P, P]
        L.ref_mat_mul.restype = None
        L.ref_mat_mul.argtypes = [P, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def or_params(n1=10.0, n2=10.0, horizon=3.015, dx=1e-3, rho=8000.0, dt=1e-8,
              E=192e9, nu=0.3, s0=0.0, kappa_tol=1e-10):
    return OrParams(n1, n2, horizon, dx, rho, dt, E, nu, s0, kappa_tol)


def weight(p, xj, xk):
    return lib().or_weight(C.byref(p), float(xj[0]), float(xj[1]), float(xk[0]), float(xk[1]))


def build_rev(row, col):
    n = len(row) - 1
    rev = np.empty(len(col), dtype=np.int32)
    rc = lib().or_build_rev(n, _p(row), _p(col), _p(rev))
    if rc != 0:
        raise ValueError("asymmetric families")
    return rev


def forces(p, X, V, row, col, u, broken=None, E_arr=None, rev=None, bugcompat=False,
           want_T=False, want_Kinv=False, nthreads=8):
    n = len(V)
    nb = len(col)
    f = np.zeros(2 * n)
    T = np.zeros(2 * nb) if want_T else None
    Ki = np.zeros(3 * nb) if want_Kinv else None
    nsing = lib().or_forces_fp64(C.byref(p), n, _p(X), _p(V), _p(E_arr), _p(row), _p(col),
                                 _p(rev), _p(broken), _p(u), int(bugcompat), _p(f), _p(T),
                                 _p(Ki), nthreads)
    return f, T, Ki, nsing


def bond_stretch(xi, du):
    return lib().or_bond_stretch(float(xi[0]), float(xi[1]), float(du[0]), float(du[1]))


def bond_breaks(xi, du, s0):
    c = (1.0 + s0) * (1.0 + s0)
    return bool(lib().or_bond_breaks(float(xi[0]), float(xi[1]), float(du[0]), float(du[1]), c))


def break_bonds(p, X, row, col, u, broken):
    n = len(row) - 1
    return lib().or_break_bonds(C.byref(p), n, _p(X), _p(row), _p(col), _p(u), _p(broken))


def step(p, X, V, row, col, rev, bc, v0, u, v, a, f, broken, nsteps, E_arr=None, nthreads=8):
    n = len(V)
    return lib().or_step_fp64(C.byref(p), n, _p(X), _p(V), _p(E_arr), _p(row), _p(col), _p(rev),
                              _p(bc), _p(v0), _p(u), _p(v), _p(a), _p(f), _p(broken), nsteps,
                              nthreads)


def sym2_eig(a, b, c):
    l1, l2 = C.c_double(), C.c_double()
    lib().or_sym2_eig(float(a), float(b), float(c), C.byref(l1), C.byref(l2))
    return l1.value, l2.value


def mazars_update(dp, e11, e22, e12, kappa, d=0.0):
    k, dd = C.c_double(kappa), C.c_double(d)
    upd = lib().or_mazars_update(C.byref(dp), float(e11), float(e22), float(e12), C.byref(k),
                                 C.byref(dd))
    return bool(upd), k.value, dd.value


def forces_damage(p, dp, X, V, row, col, u, kappa, dmg, broken=None, E_arr=None, rev=None,
                  nthreads=8):
    """Forces and the damage update at u; kappa, dmg updated in place.
    Returns (f, status, changed)."""
    n = len(V)
    f = np.zeros(2 * n)
    status = np.zeros(n)
    changed = lib().or_forces_damage_fp64(C.byref(p), C.byref(dp), n, _p(X), _p(V), _p(E_arr),
                                          _p(row), _p(col), _p(rev), _p(broken), _p(u), _p(f),
                                          None, _p(kappa), _p(dmg), _p(status), nthreads)
    return f, status, changed


def step_damage(p, dp, X, V, row, col, rev, bc, v0, u, v, a, f, broken, kappa, dmg, status,
                nsteps, E_arr=None, nthreads=8):
    n = len(V)
    return lib().or_step_damage_fp64(C.byref(p), C.byref(dp), n, _p(X), _p(V), _p(E_arr),
                                     _p(row), _p(col), _p(rev), _p(bc), _p(v0), _p(u), _p(v),
                                     _p(a), _p(f), _p(broken), _p(kappa), _p(dmg), _p(status),
                                     nsteps, nthreads)


# ---- literal long-double restatement -------------------------------------------------

def ref_params(n1=10.0, n2=10.0, horizon=3.0, dx=1.0, E=3.5e4, nu=0.2, rho=2.5e-6, dt=1e-9):
    """Defaults are the reference driver's hard-coded values (src/main_seq.cpp:607-662)."""
    return RefParams(n1, n2, horizon, dx, E, nu, rho, dt)


def ref_compute_velocity(p, X, y, V, row, col, v, a, f, i0=0, i1=None, nthreads=8):
    n = len(V)
    if i1 is None:
        i1 = n
    lib().ref_cpu_compute_velocity(C.byref(p), n, _p(X), _p(y), _p(V), _p(row), _p(col), _p(v),
                                   _p(a), _p(f), i0, i1, nthreads)


def ref_update_positions(p, y, v, a, f):
    lib().ref_cpu_update_positions(C.byref(p), len(y) // 2, _p(y), _p(v), _p(a), _p(f))


def ref_apply_velocity_bc(ids, v0, v, a):
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    v0 = np.ascontiguousarray(v0, dtype=np.longdouble)
    lib().ref_cpu_apply_velocity_bc(len(ids), _p(ids), _p(v0), _p(v), _p(a))


def ref_apply_fixed_bc(ids, X, y):
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    lib().ref_cpu_apply_fixed_bc(len(ids), _p(ids), _p(X), _p(y))


def ref_find_neighbor(X, dx, horizon, cap_per=64):
    n = len(X) // 2
    row = np.zeros(n + 1, dtype=np.int32)
    cap = n * cap_per
    col = np.zeros(cap, dtype=np.int32)
    nb = lib().ref_find_neighbor(n, _p(X), dx, horizon, _p(row), _p(col), cap)
    if nb < 0:
        raise ValueError("neighbour capacity exceeded")
    return row, col[:nb].copy()
