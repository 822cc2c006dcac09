"""Mazars bond damage on the GPU (pd_set_damage / pd_compute_damage /
pd_get_damage, fused into the bond kernel) against the oracle's restatement of
computeDamageStatus (src/main_seq.cpp:473-598).  Parity with the reference
itself is unpinned (no reference file holds damage values); the oracle is
pinned by tests/test_oracle_damage.py."""
import numpy as np
import pytest

import hassa as H
import oracle_lib as O
from conftest import say

pytestmark = pytest.mark.gpu

KAPPA0 = 1e-4  # SI plate: strains of 1e-4 .. 1e-3 under the edge velocities


@pytest.fixture(scope="module")
def plate(pdpkg):
    return pdpkg.plates.single_plate("p48x40", 48, 40, 30.0, notch_len_cells=16, seed=3)


def smooth_u(w, amp, seed):
    """A smooth displacement with strains of order amp."""
    rng = np.random.default_rng(seed)
    P = w.X.reshape(-1, 2)
    L = np.ptp(P, axis=0).max()
    k = rng.normal(size=(3, 2)) * 2 * np.pi / L
    ph = rng.uniform(0, 2 * np.pi, size=3)
    c = rng.normal(size=(3, 2))
    u = sum(c[m] * np.sin(P @ k[m] + ph[m])[:, None] / np.linalg.norm(k[m]) for m in range(3))
    return (amp * u).reshape(-1)


def oracle_params(w):
    return O.or_params(horizon=w.horizon, dx=w.dx, rho=8000.0, dt=1e-8, E=192e9, nu=0.3)


@pytest.mark.parametrize("mode", [1, 2])
def test_compute_damage_matches_oracle(pdpkg, plate, mode):
    say(f"damage at a prescribed displacement, mode {mode}")
    pdh = pdpkg.pdhip
    w = plate
    u = smooth_u(w, 5 * KAPPA0, seed=mode)
    dp = O.damage_params(mode=mode, kappa0=KAPPA0)
    kap = np.full(w.nb, KAPPA0)
    dm = np.zeros(w.nb)
    broken = w.broken.copy()
    rev = O.build_rev(w.row, w.col)
    _, st_o, changed = O.forces_damage(oracle_params(w), dp, w.X, w.V, w.row, w.col, u, kap, dm,
                                       broken=broken, rev=rev)
    assert changed > w.nb // 10
    for prec, tol in ((0, 1e-10), (1, 1e-4)):
        ctx = w.context(precision=prec, s0=0.0)
        ctx.set_state(u=u)
        ctx.set_damage(pdh.DamageParams.make(mode=mode, kappa0=KAPPA0))
        ctx.compute_damage()
        g = ctx.get_damage()
        ctx.compute_damage()  # same positions: nothing changes
        g2 = ctx.get_damage()
        ctx.close()
        for key in ("bond", "kappa", "status"):
            np.testing.assert_array_equal(g[key], g2[key])
        np.testing.assert_array_equal(g["bond"][broken == 1], 1.0)
        assert np.linalg.norm(g["kappa"] - kap) <= tol * np.linalg.norm(kap)
        assert np.linalg.norm(g["status"] - st_o) <= tol * np.linalg.norm(st_o)
        if prec == 0:
            assert np.linalg.norm(g["bond"] - dm) <= tol * np.linalg.norm(dm)
        else:
            # a bond within FP32 round-off of its threshold may take the other
            # branch; the rest agree to the FP32 strain error
            close = np.abs(g["bond"] - dm) <= 1e-3
            assert close.mean() > 0.999


@pytest.mark.parametrize("mode", [1, 2])
def test_damage_during_steps_matches_oracle(pdpkg, plate, mode):
    """300 steps with bond breaking: the fused damage update (after each force
    evaluation) against or_step_damage_fp64."""
    say(f"plate 48x40 fp64, 300 steps, damage mode {mode}")
    pdh = pdpkg.pdhip
    w = plate
    p = oracle_params(w)
    p.s0 = 0.005
    rev = O.build_rev(w.row, w.col)
    u, v, a, f = (np.zeros(2 * w.n) for _ in range(4))
    br = w.broken.copy()
    kap = np.full(w.nb, KAPPA0)
    dm = np.zeros(w.nb)
    st = np.zeros(w.n)
    dp = O.damage_params(mode=mode, kappa0=KAPPA0)
    O.step_damage(p, dp, w.X, w.V, w.row, w.col, rev, w.code, w.v0, u, v, a, f, br, kap, dm, st, 300)
    ctx = w.context(s0=0.005)
    ctx.set_damage(pdh.DamageParams.make(mode=mode, kappa0=KAPPA0))
    ctx.step(300)
    g = ctx.get_damage()
    np.testing.assert_array_equal(ctx.get_broken(), br)
    ctx.close()
    assert (dm > 0).sum() > w.nb // 20
    # the threshold is the largest equivalent strain seen: robust to round-off
    assert np.linalg.norm(g["kappa"] - kap) <= 1e-9 * np.linalg.norm(kap)
    if mode == 2:
        assert np.linalg.norm(g["bond"] - dm) <= 1e-8 * np.linalg.norm(dm)
        assert np.linalg.norm(g["status"] - st) <= 1e-8 * np.linalg.norm(st)
    else:
        # the reference law restarts from each new threshold, so a bond whose
        # strain grew by a round-off amount in one step has d ~ 0 on one side
        # and its previous value on the other: compare bond by bond
        agree = np.abs(g["bond"] - dm) <= 1e-8 * np.maximum(1.0, np.abs(dm))
        assert agree.mean() > 0.99


@pytest.mark.parametrize("precision,tile", [(0, 13201), (0, 16401), (0, 11603), (1, 11602),
                                            (1, 13201), (0, 21603), (0, 23201), (0, 31603), (1, 41603),
                                            (0, 31602)])
def test_damage_with_matrix_core_contractions(pdpkg, plate, precision, tile):
    """The fused damage update on the matrix-core tiles, 300 steps against the
    oracle (mode 2, the published law)."""
    say(f"damage on matrix-core tile {tile}, precision {precision}, 300 steps")
    pdh = pdpkg.pdhip
    w = plate
    p = oracle_params(w)
    p.s0 = 0.005
    rev = O.build_rev(w.row, w.col)
    u, v, a, f = (np.zeros(2 * w.n) for _ in range(4))
    br = w.broken.copy()
    kap = np.full(w.nb, KAPPA0)
    dm = np.zeros(w.nb)
    st = np.zeros(w.n)
    dp = O.damage_params(mode=2, kappa0=KAPPA0)
    O.step_damage(p, dp, w.X, w.V, w.row, w.col, rev, w.code, w.v0, u, v, a, f, br, kap, dm, st, 300)
    ctx = w.context(precision=precision, s0=0.005)
    ctx.set_option(pdh.OPT_TILE, tile)
    ctx.set_damage(pdh.DamageParams.make(mode=2, kappa0=KAPPA0))
    ctx.step(300)
    g = ctx.get_damage()
    stats = ctx.stats()
    brg = ctx.get_broken()
    ctx.close()
    assert stats.matrix_cores == tile // 10000
    np.testing.assert_array_equal(brg, br)
    tol = 1e-8 if precision == 0 else 1e-3
    assert np.linalg.norm(g["kappa"] - kap) <= tol * np.linalg.norm(kap)
    assert np.linalg.norm(g["status"] - st) <= tol * np.linalg.norm(st)
    if precision == 0:
        assert np.linalg.norm(g["bond"] - dm) <= tol * np.linalg.norm(dm)


def test_damage_persistent_and_off(pdpkg, plate):
    say("damage: persistent kernel bitwise, mode off stops updates")
    pdh = pdpkg.pdhip
    w = plate
    out = []
    for persist in (0, 2):
        ctx = w.context(s0=0.005)
        ctx.set_option(pdh.OPT_PERSISTENT, persist)
        ctx.set_option(pdh.OPT_WAVES_PER_WG, 2)
        ctx.set_damage(pdh.DamageParams.make(mode=pdh.DAMAGE_MAZARS, kappa0=KAPPA0))
        ctx.step(80)
        out.append(ctx.get_damage())
        if persist == 0:
            ctx.set_damage(pdh.DamageParams.make(mode=pdh.DAMAGE_OFF))
            before = ctx.get_damage()
            ctx.step(40)
            after = ctx.get_damage()
            np.testing.assert_array_equal(before["bond"], after["bond"])
        ctx.close()
    for key in ("bond", "kappa", "status"):
        np.testing.assert_array_equal(out[0][key], out[1][key])
    ctx = w.context()
    with pytest.raises(pdh.PDError):
        ctx.compute_damage()  # damage never enabled
    ctx.close()


def test_adapter_compute_damage_status(tmp_path):
    """A driver calling computeDamageStatus with the reference signature
    (include/pd_mechanics_adapter.hpp) against the oracle, reference mode and
    constants, at the driver's prescribed displacement."""
    import os
    import subprocess
    from conftest import ROOT
    say("adapter computeDamageStatus (reference signature)")
    X, V, row, col, sets = H.case()
    X = X.astype(np.float64)
    x, y = X[0::2], X[1::2]
    s = 4e-4
    u = np.empty_like(X)
    u[0::2] = s * (0.7 * x + 0.2 * y + 0.01 * y * y)
    u[1::2] = s * (0.4 * y - 0.1 * x)
    kap0 = 3.0 / 3.5e4
    kap = np.full(len(col), kap0)
    dm = np.zeros(len(col))
    p = O.or_params(horizon=H.HORIZON, dx=H.DX, rho=H.RHO, dt=H.DT, E=H.E, nu=H.NU)
    _, st, changed = O.forces_damage(p, O.damage_params(kappa0=kap0), X, V.astype(np.float64), row,
                                     col, u, kap, dm)
    assert changed > len(col) // 4
    exe = os.path.join(ROOT, "tests", "adapter", "build", "hassa_driver")
    out = tmp_path / "dmg.txt"
    r = subprocess.run([exe, H.MESH, str(out), "damage"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    status = np.loadtxt(out)
    bonds = np.loadtxt(str(out) + ".bonds")
    assert np.linalg.norm(bonds[:, 1] - kap) <= 1e-10 * np.linalg.norm(kap)
    assert np.linalg.norm(bonds[:, 0] - dm) <= 1e-9 * np.linalg.norm(dm)
    assert np.linalg.norm(status - st) <= 1e-9 * np.linalg.norm(st)


def test_autotune_restores_damage_state(pdpkg, plate):
    say("autotune leaves damage state as it was")
    pdh = pdpkg.pdhip
    ctx = plate.context(s0=0.005)
    ctx.set_damage(pdh.DamageParams.make(mode=pdh.DAMAGE_MAZARS, kappa0=KAPPA0))
    ctx.step(40)
    before = ctx.get_damage()
    st0 = ctx.get_state()
    ctx.autotune(2)
    after = ctx.get_damage()
    st1 = ctx.get_state()
    ctx.close()
    for key in ("bond", "kappa", "status"):
        np.testing.assert_array_equal(before[key], after[key])
    for key in ("u", "v", "a", "f"):
        np.testing.assert_array_equal(st0[key], st1[key])


def test_mixed_guarded_families_damage_matches_oracle(pdpkg):
    """Damage on a fracturing plate in mixed precision with guarded families
    (delta = 4.015 dx, families of 48, s0 = 0.004, 300 steps, over 10,000 bonds
    broken): the families with an ill-conditioned K take their forces and their
    damage update from the FP64 fix-up, so per-bond kappa and d and the
    particles' damageStatus stay within the mixed-precision bar of the FP64
    oracle, with the oracle's broken set."""
    say("mixed precision, guarded families, damage mode 2: delta = 4.015 plate, 300 steps")
    pdh = pdpkg.pdhip
    w = pdpkg.plates.single_plate("d4", 48, 40, 30.0, horizon=4.015, seed=3, notch_len_cells=16)
    p = oracle_params(w)
    p.s0 = 0.004
    rev = O.build_rev(w.row, w.col)
    u, v, a, f = (np.zeros(2 * w.n) for _ in range(4))
    br = w.broken.copy()
    kap = np.full(w.nb, KAPPA0)
    dm = np.zeros(w.nb)
    st = np.zeros(w.n)
    dp = O.damage_params(mode=2, kappa0=KAPPA0)
    O.step_damage(p, dp, w.X, w.V, w.row, w.col, rev, w.code, w.v0, u, v, a, f, br, kap, dm, st, 300)
    assert br.sum() > 10000
    for tile in (1603, 11603):
        ctx = w.context(precision=1, s0=0.004)
        ctx.set_option(pdh.OPT_TILE, tile)
        ctx.set_damage(pdh.DamageParams.make(mode=2, kappa0=KAPPA0))
        ctx.step(300)
        g = ctx.get_damage()
        stats = ctx.stats()
        np.testing.assert_array_equal(ctx.get_broken(), br)
        ctx.close()
        assert stats.n_guarded_families > 0
        ek = np.linalg.norm(g["kappa"] - kap) / np.linalg.norm(kap)
        ed = np.linalg.norm(g["bond"] - dm) / np.linalg.norm(dm)
        es = np.linalg.norm(g["status"] - st) / np.linalg.norm(st)
        worst = np.max(np.abs(g["kappa"] - kap) / kap)
        say(f"  tile {tile}: guarded families {stats.n_guarded_families}, kappa {ek:.2e}, "
            f"d {ed:.2e}, status {es:.2e}, worst bond kappa {worst:.2e}")
        assert ek < 1e-5 and ed < 1e-5 and es < 1e-5
        assert worst < 1e-3
