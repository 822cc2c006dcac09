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


def max_alive_stretch(w, u, broken, chunk=1 << 25):
    """Largest stretch s = |xi + du| / |xi| - 1 over the alive bonds at
    displacement u (FP64, numpy, in chunks of bonds)."""
    X = w.X.reshape(-1, 2)
    U = u.reshape(-1, 2)
    owner = np.repeat(np.arange(w.n, dtype=np.int64), np.diff(w.row))
    smax = -np.inf
    for b0 in range(0, w.nb, chunk):
        b1 = min(w.nb, b0 + chunk)
        i, j = owner[b0:b1], w.col[b0:b1]
        xi = X[j] - X[i]
        eta = xi + (U[j] - U[i])
        s = np.sqrt((eta * eta).sum(axis=1) / (xi * xi).sum(axis=1)) - 1.0
        alive = broken[b0:b1] == 0
        if alive.any():
            smax = max(smax, float(s[alive].max()))
    return smax


def broken_symmetric(w, mask):
    """Whether (j -> i) is broken exactly when (i -> j) is: the broken bonds'
    ordered pairs equal their reversed pairs as sets."""
    b = np.nonzero(mask)[0]
    owner = np.searchsorted(w.row, b, side="right") - 1
    fwd = owner.astype(np.int64) * w.n + w.col[b]
    bwd = w.col[b].astype(np.int64) * w.n + owner
    return np.array_equal(np.sort(fwd), np.sort(bwd))


def check_pair_balanced(w, f, tol=1e-12):
    """Linear momentum balance sum_i V_i f_i = 0: each pair term enters once
    with each sign.  FP64 holds it to round-off; in mixed precision the own-bond
    sum is formed in FP32 from the same FP32 terms the neighbours subtract
    singly, so it holds to the FP32 summation error (measured 4e-11 .. 2e-10 of
    sum |V f| on the 96 plates)."""
    Vf = f.reshape(-1, 2) * w.V[:, None]
    assert np.all(np.abs(Vf.sum(axis=0)) <= tol * np.abs(Vf).sum(axis=0))


@pytest.mark.parametrize("name,steps", [("plate_1024", 5000), ("plate_2048", 2000)])
def test_full_length_runs_mixed_vs_fp64(pdpkg, name, steps):
    """C2 and C3 at their configured lengths: every pd_step checks the state
    for non-finite values; mixed precision keeps the FP64 broken-bond set and
    stays within 1e-5 of the FP64 velocities; forces stay pair-balanced.
    The load waves from the moving rows do not reach the notch in these runs:
    the final broken set is exactly the pre-broken notch set, which holds
    because every alive bond's stretch at the final displacement stays below
    s0 (the breaking rule evaluated independently, in numpy)."""
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
    np.testing.assert_array_equal(ba, w.broken)
    smax = max_alive_stretch(w, a["u"], ba)
    say(f"{name}: largest alive-bond stretch {smax:.3e} (s0 = 0.02)")
    assert 0.0 < smax < 0.02
    _, nonbc, _ = H.parity_norms(m["v"], a["v"], w.code >= 0)
    assert nonbc < 1e-5, nonbc
    check_pair_balanced(w, a["f"])


def run_fp64_and_mixed(w, steps, chunk, tune=0):
    """Both precisions for `steps` steps in pd_step calls of `chunk` steps
    (each call checks the state for non-finite values)."""
    out = []
    for prec in (0, 1):
        ctx = w.context(precision=prec)
        if tune:
            ctx.autotune(tune)
        for _ in range(steps // chunk):
            ctx.step(chunk)
        st = ctx.stats()
        out.append((ctx.get_state(), ctx.get_broken(), st))
        say(f"  {w.name} precision {prec}: {steps} steps, broken {st.n_broken} "
            f"(pre-broken {int(w.broken.sum())}), tile {st.lanes_per_family}x{st.bonds_per_lane} "
            f"mf {st.matrix_cores}, guarded families {st.n_guarded_families}")
        ctx.close()
    return out


def test_c4_plates_96_configured_run(pdpkg):
    """C4 as configured: 96 plates (64..512 per side, notch length, angle and
    edge velocity drawn per plate) in one disjoint-union mesh, 1000 steps,
    autotuned tile, graph-replayed steps, FP64 and mixed precision.  Mixed
    stays within 1e-5 of FP64 on the velocities and keeps its broken-bond set
    (identical and symmetric), forces stay pair-balanced, every 500 steps the
    state is checked finite, and every alive bond ends below the critical
    stretch."""
    say("C4: plates_96, 1000 steps, autotuned, FP64 and mixed")
    w = pdpkg.plates.plates_batched()
    assert len(np.unique(w.plate_of)) == 96
    (a, ba, _), (m, bm, _) = run_fp64_and_mixed(w, 1000, 500, tune=2)
    np.testing.assert_array_equal(bm, ba)
    assert broken_symmetric(w, ba)
    assert np.all(ba >= w.broken)
    assert max_alive_stretch(w, a["u"], ba) < 0.02
    _, nonbc, _ = H.parity_norms(m["v"], a["v"], w.code >= 0)
    say(f"  mixed vs FP64 velocity error {nonbc:.3e}")
    assert nonbc < 1e-5, nonbc
    check_pair_balanced(w, a["f"])
    check_pair_balanced(w, m["f"], tol=1e-8)


def test_c5_plate_4096_configured_run(pdpkg):
    """C5 as configured: 4096 x 4096, delta = 4.015 dx (families of up to 48,
    beyond the reference's padded cap of 30), three notches, 1 % porosity,
    500 steps in FP64 and mixed precision, with the same property checks as
    C4."""
    say("C5: plate_4096, 500 steps, FP64 and mixed")
    w = pdpkg.plates.plate_4096()
    assert np.diff(w.row).max() == 48 and w.n < 4096 * 4096
    (a, ba, sa), (m, bm, _) = run_fp64_and_mixed(w, 500, 250)
    assert sa.lanes_per_family * sa.bonds_per_lane >= 48
    np.testing.assert_array_equal(bm, ba)
    assert broken_symmetric(w, ba)
    assert np.all(ba >= w.broken)
    smax = max_alive_stretch(w, a["u"], ba)
    say(f"  largest alive-bond stretch {smax:.3e}")
    assert smax < 0.02
    _, nonbc, _ = H.parity_norms(m["v"], a["v"], w.code >= 0)
    say(f"  mixed vs FP64 velocity error {nonbc:.3e}")
    assert nonbc < 1e-5, nonbc
    check_pair_balanced(w, a["f"])
