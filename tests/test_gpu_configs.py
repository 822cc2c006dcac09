"""The benchmark configurations at their full sizes.

C1 (100x100 plate, 1000 steps) is small enough for the FP64 oracle, and so is
one step of C2 (1024 x 1024).  For longer runs at the larger sizes the oracle
is too slow, so the tests use properties that hold at any size:
* determinism of the graph-replayed step;
* linear momentum balance sum_i V_i f_i = 0 (pairwise antisymmetric forces);
* symmetry of the broken-bond set;
* the mixed-precision error against the FP64 run of the same plate.
"""
import numpy as np
import pytest

import hassa as H
import oracle_lib as O
from conftest import say

pytestmark = pytest.mark.gpu


def test_c1_plate_100_1000_steps_matches_oracle(pdpkg):
    say("C1: plate_100, 1000 steps, s0 = 0.02, FP64 and mixed vs the FP64 oracle")
    pdh = pdpkg.pdhip
    w = pdpkg.plates.plate_100()
    p = O.or_params(horizon=w.horizon, dx=w.dx, rho=8000.0, dt=1e-8, E=192e9, nu=0.3, s0=0.02)
    rev = O.build_rev(w.row, w.col)
    u, v, a, f = (np.zeros(2 * w.n) for _ in range(4))
    br = w.broken.copy()
    O.step(p, w.X, w.V, w.row, w.col, rev, w.code, w.v0, u, v, a, f, br, 1000, nthreads=16)
    mask = w.code >= 0
    for prec, tol in ((0, 1e-10), (1, 1e-5)):
        ctx = w.context(precision=prec, s0=0.02)
        ctx.step(1000)
        st = ctx.get_state()
        brg = ctx.get_broken()
        ctx.close()
        np.testing.assert_array_equal(brg, br)
        _, nonbc, _ = H.parity_norms(st["v"], v, mask)
        assert nonbc < tol, (prec, nonbc)
        assert np.linalg.norm(st["u"] - u) <= tol * np.linalg.norm(u)


def test_c2_plate_1024_step_1_matches_oracle(pdpkg):
    """The C2 check of the configuration: step-1 velocities of the full plate
    against the FP64 oracle, within 1e-12 (non-prescribed particles)."""
    say("C2: plate_1024, step 1, FP64 vs the FP64 oracle (16 threads)")
    w = pdpkg.plates.plate_1024()
    p = O.or_params(horizon=w.horizon, dx=w.dx, rho=8000.0, dt=1e-8, E=192e9, nu=0.3, s0=0.02)
    rev = O.build_rev(w.row, w.col)
    u, v, a, f = (np.zeros(2 * w.n) for _ in range(4))
    br = w.broken.copy()
    O.step(p, w.X, w.V, w.row, w.col, rev, w.code, w.v0, u, v, a, f, br, 1, nthreads=16)
    ctx = w.context(precision=0)
    ctx.step(1)
    st = ctx.get_state()
    brg = ctx.get_broken()
    ctx.close()
    np.testing.assert_array_equal(brg, br)
    np.testing.assert_array_equal(st["u"], u)
    _, nonbc, _ = H.parity_norms(st["v"], v, w.code >= 0)
    assert nonbc < 1e-12, nonbc
    assert np.linalg.norm(st["f"] - f) <= 1e-12 * np.linalg.norm(f)


def test_c2_plate_1024_full_size_properties(pdpkg):
    say("C2: plate_1024 (1.05 M particles, 29.3 M bonds): determinism, momentum, "
        "broken symmetry, mixed vs FP64")
    pdh = pdpkg.pdhip
    w = pdpkg.plates.plate_1024()
    rev = O.build_rev(w.row, w.col)
    runs = []
    for prec in (0, 0, 1):
        ctx = w.context(precision=prec)
        ctx.step(40)
        st = ctx.get_state()
        runs.append((st, ctx.get_broken()))
        ctx.close()
    (a, ba), (b, bb), (m, bm) = runs
    for k in ("u", "v", "f"):
        np.testing.assert_array_equal(a[k], b[k])
    np.testing.assert_array_equal(ba, bb)
    np.testing.assert_array_equal(ba, ba[rev])
    np.testing.assert_array_equal(bm, bm[rev])
    Vf = a["f"].reshape(-1, 2) * w.V[:, None]
    assert np.all(np.abs(Vf.sum(axis=0)) <= 1e-12 * np.abs(Vf).sum(axis=0))
    mask = w.code >= 0
    _, nonbc, _ = H.parity_norms(m["v"], a["v"], mask)
    assert nonbc < 1e-5
    np.testing.assert_array_equal(bm, ba)


def test_c3_plate_2048_per_particle_E_mixed_vs_fp64(pdpkg):
    say("C3: plate_2048 (4.2 M particles, per-particle E): mixed vs FP64, 30 steps")
    w = pdpkg.plates.plate_2048()
    assert w.E is not None and np.std(w.E) > 0.04 * np.mean(w.E)
    out = []
    for prec in (0, 1):
        ctx = w.context(precision=prec)
        ctx.step(30)
        out.append((ctx.get_state(), ctx.get_broken()))
        ctx.close()
    (a, ba), (m, bm) = out
    mask = w.code >= 0
    _, nonbc, _ = H.parity_norms(m["v"], a["v"], mask)
    assert nonbc < 1e-5
    for k in ("u", "f"):
        assert np.linalg.norm(m[k] - a[k]) <= 1e-4 * np.linalg.norm(a[k])
    np.testing.assert_array_equal(bm, ba)


@pytest.mark.parametrize("name,steps", [("plate_1024", 5000), ("plate_2048", 2000)])
def test_full_length_runs_mixed_vs_fp64(pdpkg, name, steps):
    """C2 and C3 at their configured lengths: every pd_step checks the state
    for non-finite values; mixed precision keeps the FP64 broken-bond set and
    stays within 1e-5 of the FP64 velocities; forces stay pair-balanced."""
    say(f"{name}: {steps} steps, FP64 and mixed")
    w = getattr(pdpkg.plates, name)()
    out = []
    for prec in (0, 1):
        ctx = w.context(precision=prec)
        for _ in range(steps // 500):
            ctx.step(500)
        assert ctx.stats().n_ill_conditioned == 0   # every K within mixed precision's range
        out.append((ctx.get_state(), ctx.get_broken()))
        ctx.close()
    (a, ba), (m, bm) = out
    np.testing.assert_array_equal(bm, ba)
    np.testing.assert_array_equal(ba, ba[O.build_rev(w.row, w.col)])
    _, nonbc, _ = H.parity_norms(m["v"], a["v"], w.code >= 0)
    assert nonbc < 1e-5, nonbc
    Vf = a["f"].reshape(-1, 2) * w.V[:, None]
    assert np.all(np.abs(Vf.sum(axis=0)) <= 1e-12 * np.abs(Vf).sum(axis=0))
