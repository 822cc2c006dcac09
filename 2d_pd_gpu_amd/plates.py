"""Synthetic plate workloads (inputs for the bond-associated NOSB step).

A plate is an nx x ny square grid of spacing dx, centred on `origin`, each
coordinate jittered by U(-jitter, +jitter) dx, with volume V = dx^2 h.  Random
numbers come from numpy's counter-based Philox generator keyed by the seed, so
a plate is reproducible from (seed, plate index) on any machine.  Families use
the reference cutoff |xi| < 1.01 delta dx (src/SetupParticleSystem.cpp:113).
The top row (iy = ny - 1) moves at +v, the bottom row (iy = 0) at -v in y, as
prescribed velocity groups (applyVelocityBC, src/main_seq.cpp:28-36).  Notches
are segments; bonds that cross one start broken (the reference instead deletes
particles, src/SetupParticleSystem.cpp:33).

The named workloads are the plates of the reference's configuration list:
  plate_100      100 x 100, central notch 20 dx, v = 20 m/s (CPU-runnable case)
  plate_1024     1024 x 1024, edge notch 0.25 W, v = 20 m/s
  plate_2048     2048 x 2048, per-particle E ~ N(192 GPa, 5 %), central notch, v = 40 m/s
  plates_96      96 plates, sizes U{64..512}^2, notch U[0.1, 0.4] W at U[0, 45] deg,
                 v ~ U[5, 40] m/s, one disjoint-union mesh
  plate_4096     4096 x 4096, delta = 4.015 dx, three notches, 1 % porosity
"""
from dataclasses import dataclass, field

import numpy as np

from . import pdhip

E_STEEL = 192e9
NU_STEEL = 0.3
RHO_STEEL = 8000.0
S0 = 0.02
DT = 1e-8
DX = 1e-3
H = 1e-3


@dataclass
class Workload:
    name: str
    X: np.ndarray            # 2n
    V: np.ndarray            # n
    row: np.ndarray          # n+1
    col: np.ndarray          # nb
    code: np.ndarray         # n, BC code (-1 free, g >= 0 velocity group)
    v0: np.ndarray           # 2 * groups
    broken: np.ndarray       # nb, pre-broken notch bonds
    horizon: float = 3.015
    dx: float = DX
    E: np.ndarray = None     # per-particle E or None
    plate_of: np.ndarray = None
    info: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.V.size

    @property
    def nb(self):
        return int(self.row[-1])

    def params(self, precision=pdhip.FP64, s0=S0, dt=DT):
        return pdhip.Params.make(precision=precision, horizon=self.horizon, dx=self.dx,
                                 rho=RHO_STEEL, dt=dt, E=E_STEEL, nu=NU_STEEL, s0=s0)

    def context(self, precision=pdhip.FP64, s0=S0, dt=DT):
        ctx = pdhip.Context(self.params(precision, s0, dt), self.X, self.V, self.row, self.col)
        if self.E is not None:
            ctx.set_youngs_modulus(self.E)
        ctx.set_boundary(self.code, self.v0)
        if self.broken.any():
            ctx.set_broken(self.broken)
        return ctx


def _rng(seed, stream):
    return np.random.Generator(np.random.Philox(key=[int(seed) & (2**64 - 1), int(stream)]))


def plate_points(nx, ny, dx=DX, jitter=0.01, seed=0, stream=0, origin=(0.0, 0.0), porosity=0.0):
    """Return X (2n), iy (n) for one jittered plate; porous plates drop a
    fraction of the non-edge particles."""
    rng = _rng(seed, stream)
    iy, ix = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ix = ix.ravel()
    iy = iy.ravel()
    x = origin[0] + (ix - (nx - 1) / 2.0) * dx
    y = origin[1] + (iy - (ny - 1) / 2.0) * dx
    jit = rng.uniform(-jitter, jitter, size=(2, ix.size)) * dx
    x = x + jit[0]
    y = y + jit[1]
    keep = np.ones(ix.size, dtype=bool)
    if porosity > 0:
        drop = rng.random(ix.size) < porosity
        drop &= (iy > 0) & (iy < ny - 1)
        keep = ~drop
    X = np.empty(2 * int(keep.sum()))
    X[0::2] = x[keep]
    X[1::2] = y[keep]
    return X, iy[keep]


def notch_segment(cx, cy, length, angle_deg=0.0):
    a = np.deg2rad(angle_deg)
    dxh = 0.5 * length * np.cos(a)
    dyh = 0.5 * length * np.sin(a)
    return [cx - dxh, cy - dyh, cx + dxh, cy + dyh]


def notch_mask(X, row, col, segments):
    seg = np.ascontiguousarray(np.asarray(segments, dtype=np.float64).reshape(-1))
    mask = np.zeros(max(len(col), 1), dtype=np.uint8)
    X = np.ascontiguousarray(X, dtype=np.float64)
    rc = pdhip.lib().pd_notch_mask(len(row) - 1, pdhip._ptr(X), pdhip._ptr(row), pdhip._ptr(col),
                                  seg.size // 4, pdhip._ptr(seg), pdhip._ptr(mask))
    if rc < 0:
        raise pdhip.PDError(pdhip.PD_ERR_INVALID, "pd_notch_mask failed")
    return mask[: len(col)]


def single_plate(name, nx, ny, v, notch_len_cells=None, notch_kind="center", horizon=3.015,
                 seed=0, E_sigma=0.0, porosity=0.0, notches=None, dx=DX):
    X, iy = plate_points(nx, ny, dx=dx, seed=seed, porosity=porosity)
    n = X.size // 2
    V = np.full(n, dx * dx * H)
    row, col = pdhip.build_families(X, dx, horizon)
    # notch line halfway between the two middle rows
    ymid = ((ny // 2 - 1) - (ny - 1) / 2.0 + 0.5) * dx
    segs = []
    W = nx * dx
    if notches is not None:
        segs = notches
    elif notch_len_cells is not None:
        L = notch_len_cells * dx
        if notch_kind == "center":
            segs = [[-0.5 * L, ymid, 0.5 * L, ymid]]
        else:  # edge notch from the left edge
            x0 = -(nx - 1) / 2.0 * dx - dx
            segs = [[x0, ymid, x0 + dx + L, ymid]]
    broken = notch_mask(X, row, col, segs) if segs else np.zeros(len(col), dtype=np.uint8)
    code = np.full(n, pdhip.BC_FREE, dtype=np.int32)
    code[iy == ny - 1] = 0
    code[iy == 0] = 1
    v0 = np.array([0.0, v, 0.0, -v])
    E = None
    if E_sigma > 0:
        E = _rng(seed, 1 << 20).normal(E_STEEL, E_sigma * E_STEEL, size=n)
        E = np.maximum(E, 0.5 * E_STEEL)
    return Workload(name, X, V, row, col, code, v0, broken, horizon=horizon, dx=dx, E=E,
                    info={"nx": nx, "ny": ny, "v": v, "notches": segs, "W": W})


def plate_100(seed=0):
    return single_plate("plate_100", 100, 100, 20.0, notch_len_cells=20, notch_kind="center", seed=seed)


def plate_1024(seed=0):
    return single_plate("plate_1024", 1024, 1024, 20.0, notch_len_cells=256, notch_kind="edge", seed=seed)


def plate_2048(seed=0):
    return single_plate("plate_2048", 2048, 2048, 40.0, notch_len_cells=512, notch_kind="center",
                        seed=seed, E_sigma=0.05)


def plate_4096(seed=0, n=4096):
    rng = _rng(seed, 7)
    W = n * DX
    segs = []
    for _ in range(3):
        cx, cy = rng.uniform(-0.3 * W, 0.3 * W, size=2)
        segs.append(notch_segment(cx, cy, rng.uniform(0.05, 0.15) * W, rng.uniform(0, 180)))
    return single_plate(f"plate_{n}", n, n, 20.0, horizon=4.015, seed=seed, porosity=0.01, notches=segs)


def plates_batched(count=96, seed=0, lo=64, hi=512):
    """Many independent plates as one disjoint-union mesh (plates side by side
    with a gap wider than the horizon, so no bond crosses plates)."""
    rng = _rng(seed, 11)
    sizes = rng.integers(lo, hi + 1, size=(count, 2))
    nl = rng.uniform(0.1, 0.4, size=count)
    ang = rng.uniform(0.0, 45.0, size=count)
    vel = rng.uniform(5.0, 40.0, size=count)
    Xs, codes, plate_of, segs = [], [], [], []
    gap = 8 * DX
    cols = int(np.ceil(np.sqrt(count)))
    cell = (hi + 8) * DX
    for p in range(count):
        nx, ny = int(sizes[p, 0]), int(sizes[p, 1])
        ox = (p % cols) * (cell + gap)
        oy = (p // cols) * (cell + gap)
        X, iy = plate_points(nx, ny, seed=seed, stream=p + 1, origin=(ox, oy))
        c = np.full(X.size // 2, pdhip.BC_FREE, dtype=np.int32)
        c[iy == ny - 1] = 2 * p
        c[iy == 0] = 2 * p + 1
        Xs.append(X)
        codes.append(c)
        plate_of.append(np.full(X.size // 2, p, dtype=np.int32))
        ymid = oy + ((ny // 2 - 1) - (ny - 1) / 2.0 + 0.5) * DX
        segs.append(notch_segment(ox, ymid, nl[p] * nx * DX, ang[p]))
    X = np.concatenate(Xs)
    code = np.concatenate(codes)
    n = X.size // 2
    V = np.full(n, DX * DX * H)
    row, col = pdhip.build_families(X, DX, 3.015)
    broken = notch_mask(X, row, col, segs)
    v0 = np.empty(4 * count)
    v0[0::4] = 0.0
    v0[1::4] = vel
    v0[2::4] = 0.0
    v0[3::4] = -vel
    return Workload("plates_96", X, V, row, col, code, v0, broken, plate_of=np.concatenate(plate_of),
                    info={"sizes": sizes.tolist(), "v": vel.tolist(), "notch_len": nl.tolist(),
                          "angle": ang.tolist()})


WORKLOADS = {
    "plate_100": plate_100,
    "plate_1024": plate_1024,
    "plate_2048": plate_2048,
    "plates_96": plates_batched,
    "plate_4096": plate_4096,
}
