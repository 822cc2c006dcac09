"""Python host binding of lib2dpdgpu (include/pd_hip.h) via ctypes.

Mirrors the reference's mechanics entry points (src/main_seq.cpp:15-46,
:386-470; include/mechanics.cuh:867-881): ``Context.compute_velocity``,
``update_positions``, ``apply_velocity_bc``, ``apply_fixed_bc``, plus the fused
``step``.  Arrays are numpy, in the caller's particle order, xy interleaved
(``double[2n]``) as the reference GPU path packs them
(include/mechanics.cuh:1009-1031).  Every call goes to the HIP library; there
is no CPU fallback: a missing library or a HIP failure raises ``PDError``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; PD_HIP_LIB names another build of the same library
LIB_PATH = os.environ.get("PD_HIP_LIB") or os.path.join(_HERE, "lib", "lib2dpdgpu.so")

PD_OK = 0
PD_ERR_INVALID = 1
PD_ERR_HIP = 2
PD_ERR_UNSUPPORTED = 3
PD_ERR_NOMEM = 4
PD_ERR_NOT_FINITE = 5

FP64 = 0
MIXED = 1

BC_FREE = -1
BC_FIXED = -2

OPT_USE_GRAPH = 1
OPT_PROFILE = 2
OPT_WAVES_PER_WG = 3
OPT_XCD_REMAP = 4
OPT_TILE = 5
OPT_PERSISTENT = 6
OPT_CHECK_FINITE = 7
OPT_MIXED_GUARD = 8

DAMAGE_OFF = 0
DAMAGE_REFERENCE = 1
DAMAGE_MAZARS = 2

# every entry point include/pd_hip.h declares
EXPORTS = (
    "pd_last_error", "pd_version", "pd_create", "pd_create_padded", "pd_destroy",
    "pd_set_stream", "pd_synchronize", "pd_set_option", "pd_set_youngs_modulus",
    "pd_set_state", "pd_get_state", "pd_set_positions", "pd_get_positions",
    "pd_apply_velocity_bc", "pd_apply_fixed_bc", "pd_update_positions",
    "pd_compute_velocity", "pd_set_boundary", "pd_step", "pd_autotune",
    "pd_get_broken", "pd_set_broken", "pd_get_bond_forces", "pd_get_stats",
    "pd_reset_profile", "pd_build_families", "pd_notch_mask", "pd_probe_copy",
    "pd_set_damage", "pd_compute_damage", "pd_get_damage", "pd_set_damage_state",
    "pd_read_mesh", "pd_write_velocity", "pd_write_positions", "pd_write_gathered",
    "pd_create_from_points", "pd_get_families", "pd_save_checkpoint", "pd_load_checkpoint",
    "pd_compute_velocity_static",
)


class DamageParams(C.Structure):
    """pd_damage_params.  Defaults: the reference's Mazars constants
    (src/main_seq.cpp:480) and threshold tensileStrength / E = 3 / 3.5e4
    (:660, reference units)."""
    _fields_ = [
        ("mode", C.c_int), ("kappa0", C.c_double), ("At", C.c_double), ("Bt", C.c_double),
        ("Ac", C.c_double), ("Bc", C.c_double),
    ]

    @classmethod
    def make(cls, mode=DAMAGE_REFERENCE, kappa0=3.0 / 3.5e4, At=0.87, Bt=2.0e4, Ac=0.65,
             Bc=2150.0):
        return cls(mode, kappa0, At, Bt, Ac, Bc)


class PDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lib2dpdgpu error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    """pd_params.  Defaults: the SI plate of the configs (E = 192 GPa,
    nu = 0.3, rho = 8000, dx = 1 mm, delta = 3.015 dx, dt = 1e-8 s, s0 = 0.02)."""
    _fields_ = [
        ("ndim", C.c_int), ("precision", C.c_int),
        ("n1", C.c_double), ("n2", C.c_double), ("horizon", C.c_double), ("dx", C.c_double),
        ("rho", C.c_double), ("dt", C.c_double), ("E", C.c_double), ("nu", C.c_double),
        ("s0", C.c_double), ("kappa_tol", C.c_double),
    ]

    @classmethod
    def make(cls, precision=FP64, n1=10.0, n2=10.0, horizon=3.015, dx=1e-3, rho=8000.0,
             dt=1e-8, E=192e9, nu=0.3, s0=0.02, kappa_tol=1e-10):
        return cls(2, precision, n1, n2, horizon, dx, rho, dt, E, nu, s0, kappa_tol)


class Stats(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("nb", C.c_long), ("n_broken", C.c_long), ("n_singular", C.c_long),
        ("max_family", C.c_int), ("lanes_per_family", C.c_int), ("bonds_per_lane", C.c_int),
        ("persistent", C.c_int), ("waves_per_wg", C.c_int), ("steps", C.c_long),
        ("phase_ms", C.c_double * 4), ("phase_calls", C.c_long * 4), ("matrix_cores", C.c_int),
        ("n_ill_conditioned", C.c_long), ("n_guarded_families", C.c_long),
    ]


_lib = None


def lib():
    """Load lib2dpdgpu.so (built by ``__graft_entry__.build()``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PDError(-1, f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    I = C.c_int
    ctx = C.c_void_p
    sig = {
        "pd_last_error": (C.c_char_p, [ctx]),
        "pd_version": (C.c_char_p, []),
        "pd_create": (I, [C.POINTER(Params), I, P, P, P, P, C.POINTER(C.c_void_p)]),
        "pd_create_padded": (I, [C.POINTER(Params), I, P, P, P, P, I, C.POINTER(C.c_void_p)]),
        "pd_destroy": (None, [ctx]),
        "pd_set_stream": (I, [ctx, P]),
        "pd_synchronize": (I, [ctx]),
        "pd_set_option": (I, [ctx, I, C.c_long]),
        "pd_set_youngs_modulus": (I, [ctx, P]),
        "pd_set_state": (I, [ctx, P, P, P, P]),
        "pd_get_state": (I, [ctx, P, P, P, P]),
        "pd_set_positions": (I, [ctx, P]),
        "pd_get_positions": (I, [ctx, P]),
        "pd_apply_velocity_bc": (I, [ctx, I, P, P]),
        "pd_apply_fixed_bc": (I, [ctx, I, P]),
        "pd_update_positions": (I, [ctx]),
        "pd_compute_velocity": (I, [ctx, I]),
        "pd_set_boundary": (I, [ctx, P, I, P]),
        "pd_step": (I, [ctx, I]),
        "pd_autotune": (I, [ctx, I]),
        "pd_get_broken": (I, [ctx, P]),
        "pd_set_broken": (I, [ctx, P]),
        "pd_get_bond_forces": (I, [ctx, P]),
        "pd_get_stats": (I, [ctx, C.POINTER(Stats)]),
        "pd_reset_profile": (None, [ctx]),
        "pd_build_families": (C.c_long, [I, P, C.c_double, C.c_double, P, P, C.c_long]),
        "pd_notch_mask": (C.c_long, [I, P, P, P, I, P, P]),
        "pd_probe_copy": (I, [C.c_long, I, C.POINTER(C.c_double)]),
        "pd_set_damage": (I, [ctx, C.POINTER(DamageParams)]),
        "pd_compute_damage": (I, [ctx]),
        "pd_get_damage": (I, [ctx, P, P, P]),
        "pd_set_damage_state": (I, [ctx, P, P]),
        "pd_read_mesh": (C.c_long, [C.c_char_p, I, P, P, C.c_long]),
        "pd_create_from_points": (I, [C.POINTER(Params), I, P, P, C.POINTER(C.c_void_p)]),
        "pd_get_families": (I, [ctx, P, P]),
        "pd_save_checkpoint": (I, [ctx, C.c_char_p]),
        "pd_load_checkpoint": (I, [ctx, C.c_char_p]),
        "pd_write_velocity": (I, [ctx, C.c_char_p, I]),
        "pd_write_positions": (I, [ctx, C.c_char_p, P, I]),
        "pd_write_gathered": (I, [ctx, C.c_char_p, P, I]),
        "pd_compute_velocity_static": (I, [C.POINTER(Params), I, C.c_long, I] + [P] * 9 + [I, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _i32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.int32)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _u8(a, n):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def build_families(X, dx, horizon):
    """Families with the reference cutoff |xi| < 1.01 dx delta, ascending ids
    (findNeighbor, src/SetupParticleSystem.cpp:100-136).  Returns (row, col)."""
    X = _f64(X)
    n = X.size // 2
    row = np.zeros(n + 1, dtype=np.int32)
    nb = lib().pd_build_families(n, _ptr(X), float(dx), float(horizon), _ptr(row), None, 0)
    if nb < 0:
        raise PDError(PD_ERR_INVALID, "pd_build_families failed")
    col = np.zeros(max(nb, 1), dtype=np.int32)
    nb2 = lib().pd_build_families(n, _ptr(X), float(dx), float(horizon), _ptr(row), _ptr(col), nb)
    if nb2 != nb:
        raise PDError(PD_ERR_INVALID, "pd_build_families failed")
    return row, col[:nb]


def read_mesh(path, reference_notch=True):
    """readMeshFile (src/SetupParticleSystem.cpp:11-60): returns X (2n), V (n)."""
    p = os.fsencode(path)
    n = lib().pd_read_mesh(p, int(bool(reference_notch)), None, None, 0)
    if n < 0:
        raise PDError(PD_ERR_INVALID, lib().pd_last_error(None).decode())
    X = np.zeros(2 * max(n, 1))
    V = np.zeros(max(n, 1))
    if lib().pd_read_mesh(p, int(bool(reference_notch)), _ptr(X), _ptr(V), n) != n:
        raise PDError(PD_ERR_INVALID, lib().pd_last_error(None).decode())
    return X[: 2 * n], V[:n]


def probe_copy(nbytes=1 << 30, iters=20):
    """Achievable HBM rate (GB/s) of a plain copy kernel reading and writing
    `nbytes`; also the kernel of known traffic for counter calibration."""
    g = C.c_double()
    rc = lib().pd_probe_copy(int(nbytes), int(iters), C.byref(g))
    if rc != PD_OK:
        raise PDError(rc, lib().pd_last_error(None).decode())
    return g.value


class Context:
    """One simulation context: mesh, families and state resident on the GPU."""

    def __init__(self, params, X, V, row, col):
        self._h = C.c_void_p()
        self.params = params
        V = _f64(V)
        self.n = V.size
        X = _f64(X, 2 * self.n)
        row = _i32(row, self.n + 1)
        if row[0] != 0 or np.any(np.diff(row) < 0):
            raise ValueError("row_ptr must start at 0 and never decrease")
        self.nb = int(row[-1])
        col = _i32(col, self.nb)
        rc = lib().pd_create(C.byref(params), self.n, _ptr(X), _ptr(V), _ptr(row), _ptr(col),
                             C.byref(self._h))
        if rc != PD_OK:
            raise PDError(rc, lib().pd_last_error(None).decode())

    @classmethod
    def from_points(cls, params, X, V):
        """Families built on the device (pd_create_from_points); their CSR in
        the caller's bond order is ``families()``."""
        self = cls.__new__(cls)
        self._h = C.c_void_p()
        self.params = params
        V = _f64(V)
        self.n = V.size
        X = _f64(X, 2 * self.n)
        rc = lib().pd_create_from_points(C.byref(params), self.n, _ptr(X), _ptr(V), C.byref(self._h))
        if rc != PD_OK:
            raise PDError(rc, lib().pd_last_error(None).decode())
        row = np.zeros(self.n + 1, dtype=np.int32)
        self._check(lib().pd_get_families(self._h, _ptr(row), None))
        self.nb = int(row[-1])
        return self

    def families(self):
        """(row_ptr, col) in the caller's order."""
        row = np.zeros(self.n + 1, dtype=np.int32)
        col = np.zeros(max(self.nb, 1), dtype=np.int32)
        self._check(lib().pd_get_families(self._h, _ptr(row), _ptr(col)))
        return row, col[: self.nb]

    @classmethod
    def from_padded(cls, params, X, V, nbr, nbr_count, cap):
        self = cls.__new__(cls)
        self._h = C.c_void_p()
        self.params = params
        V = _f64(V)
        self.n = V.size
        X = _f64(X, 2 * self.n)
        cnt = _i32(nbr_count, self.n)
        nbr = _i32(nbr, self.n * int(cap))
        self.nb = int(cnt.sum())
        rc = lib().pd_create_padded(C.byref(params), self.n, _ptr(X), _ptr(V), _ptr(nbr),
                                    _ptr(cnt), int(cap), C.byref(self._h))
        if rc != PD_OK:
            raise PDError(rc, lib().pd_last_error(None).decode())
        return self

    # -- plumbing --
    def _check(self, rc):
        if rc != PD_OK:
            raise PDError(rc, lib().pd_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().pd_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, opt, value):
        self._check(lib().pd_set_option(self._h, int(opt), int(value)))

    def synchronize(self):
        self._check(lib().pd_synchronize(self._h))

    # -- material / state --
    def set_youngs_modulus(self, E):
        self._check(lib().pd_set_youngs_modulus(self._h, _ptr(_f64(E, self.n))))

    def set_state(self, u=None, v=None, a=None, f=None):
        arrs = [None if x is None else _f64(x, 2 * self.n) for x in (u, v, a, f)]
        self._check(lib().pd_set_state(self._h, *[_ptr(x) for x in arrs]))

    def get_state(self):
        out = {k: np.zeros(2 * self.n) for k in ("u", "v", "a", "f")}
        self._check(lib().pd_get_state(self._h, _ptr(out["u"]), _ptr(out["v"]), _ptr(out["a"]),
                                       _ptr(out["f"])))
        return out

    def set_positions(self, y):
        self._check(lib().pd_set_positions(self._h, _ptr(_f64(y, 2 * self.n))))

    def get_positions(self):
        y = np.zeros(2 * self.n)
        self._check(lib().pd_get_positions(self._h, _ptr(y)))
        return y

    # -- reference phase calls --
    def apply_velocity_bc(self, ids, v0):
        ids = _i32(ids)
        self._check(lib().pd_apply_velocity_bc(self._h, ids.size, _ptr(ids), _ptr(_f64(v0, 2))))

    def apply_fixed_bc(self, ids):
        ids = _i32(ids)
        self._check(lib().pd_apply_fixed_bc(self._h, ids.size, _ptr(ids)))

    def update_positions(self):
        self._check(lib().pd_update_positions(self._h))

    def compute_velocity(self, write_acceleration=False):
        self._check(lib().pd_compute_velocity(self._h, int(bool(write_acceleration))))

    # -- whole steps --
    def set_boundary(self, code, v0_groups):
        code = _i32(code, self.n)
        v0 = _f64(v0_groups)
        if v0.size % 2:
            raise ValueError("v0_groups holds (vx, vy) per group: an even count")
        self._check(lib().pd_set_boundary(self._h, _ptr(code), v0.size // 2, _ptr(v0)))

    def step(self, nsteps=1):
        self._check(lib().pd_step(self._h, int(nsteps)))

    def autotune(self, steps=5):
        self._check(lib().pd_autotune(self._h, int(steps)))

    # -- bonds / diagnostics --
    def get_broken(self):
        m = np.zeros(max(self.nb, 1), dtype=np.uint8)
        self._check(lib().pd_get_broken(self._h, _ptr(m)))
        return m[: self.nb]

    def set_broken(self, mask):
        m = _u8(mask, self.nb)
        self._check(lib().pd_set_broken(self._h, _ptr(m)))

    def get_bond_forces(self):
        T = np.zeros(2 * max(self.nb, 1))
        self._check(lib().pd_get_bond_forces(self._h, _ptr(T)))
        return T[: 2 * self.nb]

    # -- Mazars bond damage (computeDamageStatus) --
    def set_damage(self, params=None, **kw):
        """Enable (or, with mode=DAMAGE_OFF, stop) the damage update; resets
        kappa = kappa0 and d = 0.  ``params``: DamageParams, or keywords of
        DamageParams.make."""
        dp = params if params is not None else DamageParams.make(**kw)
        self._check(lib().pd_set_damage(self._h, C.byref(dp)))

    def compute_damage(self):
        self._check(lib().pd_compute_damage(self._h))

    def set_damage_state(self, bond_damage=None, kappa=None):
        d = None if bond_damage is None else _f64(bond_damage, self.nb)
        k = None if kappa is None else _f64(kappa, self.nb)
        self._check(lib().pd_set_damage_state(self._h, _ptr(d), _ptr(k)))

    def get_damage(self):
        """dict(bond=d per bond, status=damageStatus per particle, kappa=threshold per bond)."""
        d = np.zeros(max(self.nb, 1))
        k = np.zeros(max(self.nb, 1))
        st = np.zeros(self.n)
        self._check(lib().pd_get_damage(self._h, _ptr(d), _ptr(st), _ptr(k)))
        return {"bond": d[: self.nb], "status": st, "kappa": k[: self.nb]}

    # -- result files in the reference's formats --
    def write_velocity(self, path, digits=0):
        self._check(lib().pd_write_velocity(self._h, os.fsencode(path), int(digits)))

    def write_positions(self, path, ids=None, digits=0):
        ids = None if ids is None else _i32(ids, self.n)
        self._check(lib().pd_write_positions(self._h, os.fsencode(path), _ptr(ids), int(digits)))

    def write_gathered(self, path, ids=None, digits=0):
        ids = None if ids is None else _i32(ids, self.n)
        self._check(lib().pd_write_gathered(self._h, os.fsencode(path), _ptr(ids), int(digits)))

    # -- checkpoint / resume --
    def save_checkpoint(self, path):
        self._check(lib().pd_save_checkpoint(self._h, os.fsencode(path)))

    def load_checkpoint(self, path):
        self._check(lib().pd_load_checkpoint(self._h, os.fsencode(path)))

    def stats(self):
        s = Stats()
        self._check(lib().pd_get_stats(self._h, C.byref(s)))
        return s

    def reset_profile(self):
        lib().pd_reset_profile(self._h)
