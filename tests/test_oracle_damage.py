"""Checks that pin the oracle's restatement of the reference's Mazars bond damage
(computeDamageStatus, src/main_seq.cpp:473-598).

No file of the reference holds damage values and the reference routine needs
Eigen (absent here), so parity with the reference itself is unpinned.  The
restatement is pinned instead by:

* principal strains against numpy.linalg.eigvalsh, on rotated tensors;
* the equivalent strain of uniaxial, biaxial, mixed-sign and compressive states
  in closed form, independent of orientation;
* the limits of the damage law (0 at the threshold, 1 - (1 - A_t) kappa0 / eps
  far above it) and alpha_t = 1 / alpha_c = 0 for the strain split;
* the tension law's rate constants at interior points: ln(1 - (1 - A_t)
  kappa_r / eps - d) is linear in eps with slope -B_t and intercept ln A_t;
* the compression law is unreachable (alpha_c = 0 for every strain), so A_c
  and B_c never change d;
* idempotence at a fixed strain, and the reference mode's running-threshold
  behaviour against the published (initial-threshold) law;
* the affine patch test: u = eps0 X gives F_ik - I = eps0 for every bond, so
  every bond's damage equals the single-strain value.
"""
import numpy as np
import pytest

import oracle_lib as O

KAPPA0 = 3.0 / 3.5e4  # src/main_seq.cpp:660


def rot(theta):
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[c, -s], [s, c]])


def test_principal_strains_match_eigvalsh():
    rng = np.random.default_rng(11)
    for _ in range(500):
        lam = np.sort(rng.normal(size=2)) * 10.0 ** rng.uniform(-6, 0)
        if rng.random() < 0.1:
            lam[1] = lam[0]  # degenerate
        R = rot(rng.uniform(0, 2 * np.pi))
        eps = R @ np.diag(lam) @ R.T
        l1, l2 = O.sym2_eig(eps[0, 0], 0.5 * (eps[0, 1] + eps[1, 0]), eps[1, 1])
        ref = np.linalg.eigvalsh(0.5 * (eps + eps.T))
        scale = np.abs(ref).max()
        assert l1 <= l2
        assert abs(l1 - ref[0]) <= 1e-14 * scale and abs(l2 - ref[1]) <= 1e-14 * scale
    assert O.sym2_eig(0.0, 0.0, 0.0) == (0.0, 0.0)


@pytest.mark.parametrize("theta", [0.0, 0.3, 1.1, 2.5])
def test_equivalent_strain_closed_forms(theta):
    dp = O.damage_params()
    R = rot(theta)
    e = 4.0 * KAPPA0
    cases = [
        (np.diag([e, 0.0]), e),                 # uniaxial tension
        (np.diag([e, e]), np.sqrt(2.0) * e),    # equibiaxial tension
        (np.diag([-e, e]), e),                  # pure shear: only the positive part counts
        (np.diag([-e, -2 * e]), 0.0),           # compression: no tensile strain
    ]
    for lam, eq in cases:
        eps = R @ lam @ R.T
        upd, kap, d = O.mazars_update(dp, eps[0, 0], eps[1, 1], eps[0, 1], 1e-300)
        if eq == 0.0:
            assert not upd and kap == 1e-300 and d == 0.0
        else:
            assert upd
            assert kap == pytest.approx(eq, rel=1e-13)


def test_damage_law_limits_and_split():
    """alpha_t = 1 and alpha_c = 0 for the positive / negative strain split, so
    the first loading past kappa0 follows the tension law; it starts at 0 and
    tends to 1 - (1 - A_t) kappa0 / eps once the exponential has died out."""
    dp = O.damage_params()
    _, _, d0 = O.mazars_update(dp, KAPPA0 * (1 + 1e-9), 0.0, 0.0, KAPPA0)
    assert 0.0 <= d0 < 1e-6
    far = 100.0 * KAPPA0  # B_t (eps - kappa0) ~ 170: exp term below 1e-73
    _, _, d_far = O.mazars_update(dp, far, 0.0, 0.0, KAPPA0)
    assert d_far == pytest.approx(1.0 - (1.0 - dp.At) * KAPPA0 / far, rel=1e-14)
    # a compressive principal strain leaves alpha_t at 1: same damage as the
    # uniaxial state with the same tensile strain
    _, _, d_u = O.mazars_update(dp, 3 * KAPPA0, 0.0, 0.0, KAPPA0)
    _, _, d_m = O.mazars_update(dp, 3 * KAPPA0, -5 * KAPPA0, 0.0, KAPPA0)
    assert d_m == pytest.approx(d_u, rel=1e-14)
    # the law increases with the equivalent strain
    ds = [O.mazars_update(dp, s * KAPPA0, 0.0, 0.0, KAPPA0)[2] for s in (1.5, 2, 4, 8, 16)]
    assert all(b > a for a, b in zip(ds, ds[1:]))


def test_reference_running_threshold_versus_published_law():
    dpr = O.damage_params(mode=O.DAMAGE_REFERENCE)
    dpm = O.damage_params(mode=O.DAMAGE_MAZARS)
    out = {}
    for name, dp in (("ref", dpr), ("maz", dpm)):
        kap, d, hist = KAPPA0, 0.0, []
        for s in (2.0, 2.1, 2.1, 1.5, 3.0):
            upd, kap, d = O.mazars_update(dp, s * KAPPA0, 0.0, 0.0, kap, d)
            hist.append((upd, kap, d))
        out[name] = hist
    for name in ("ref", "maz"):
        h = out[name]
        assert h[0][0] and h[1][0]
        assert not h[2][0] and h[2][1:] == h[1][1:]   # same strain again: no change
        assert not h[3][0] and h[3][1:] == h[1][1:]   # unloading: no change
        assert h[4][0] and h[4][1] == pytest.approx(3.0 * KAPPA0)
    # the published law is monotone in the loading history ...
    assert out["maz"][1][2] > out["maz"][0][2]
    assert out["maz"][4][2] > out["maz"][1][2]
    # ... the reference's running threshold restarts the law from each new
    # threshold (src/main_seq.cpp:579-582), so a small increment gives less
    assert out["ref"][1][2] < out["ref"][0][2]
    assert out["ref"][0][2] == out["maz"][0][2]   # identical on first loading


def grid(nx, ny, jitter, seed):
    rng = np.random.default_rng(seed)
    x, y = np.meshgrid(np.arange(nx, dtype=float), np.arange(ny, dtype=float), indexing="xy")
    X = np.stack([x.ravel(), y.ravel()], axis=1)
    X += rng.uniform(-jitter, jitter, size=X.shape)
    return X.reshape(-1)


def families(X, horizon):
    P = X.reshape(-1, 2)
    d = np.linalg.norm(P[:, None, :] - P[None, :, :], axis=2)
    adj = (d < 1.01 * horizon) & (d > 0)
    row = np.concatenate([[0], np.cumsum(adj.sum(axis=1))]).astype(np.int32)
    col = np.nonzero(adj)[1].astype(np.int32)
    return row, col


@pytest.mark.parametrize("mode", [O.DAMAGE_REFERENCE, O.DAMAGE_MAZARS])
def test_affine_patch_damage_is_uniform(mode):
    X = grid(10, 9, 0.03, 4)
    n = X.size // 2
    row, col = families(X, 3.015)
    V = np.ones(n)
    p = O.or_params(horizon=3.015, dx=1.0, E=3.5e4, nu=0.2)
    dp = O.damage_params(mode=mode)
    eps = np.array([[5e-4, 1e-4], [3e-4, -2e-4]])
    u = (X.reshape(-1, 2) @ eps.T).reshape(-1)
    kappa = np.full(len(col), KAPPA0)
    dmg = np.zeros(len(col))
    f, status, changed = O.forces_damage(p, dp, X, V, row, col, u, kappa, dmg)
    e12 = 0.5 * (eps[0, 1] + eps[1, 0])
    _, k1, d1 = O.mazars_update(dp, eps[0, 0], eps[1, 1], e12, KAPPA0)
    assert changed == len(col)
    np.testing.assert_allclose(kappa, k1, rtol=1e-10)
    np.testing.assert_allclose(dmg, d1, rtol=1e-8)
    np.testing.assert_allclose(status, d1, rtol=1e-8)
    # the same positions again change nothing
    k_before, d_before = kappa.copy(), dmg.copy()
    _, status2, changed2 = O.forces_damage(p, dp, X, V, row, col, u, kappa, dmg)
    assert changed2 == 0
    np.testing.assert_array_equal(kappa, k_before)
    np.testing.assert_array_equal(dmg, d_before)
    np.testing.assert_array_equal(status2, status)


def test_broken_bonds_count_as_fully_damaged():
    X = grid(8, 8, 0.03, 5)
    n = X.size // 2
    row, col = families(X, 3.015)
    V = np.ones(n)
    rev = O.build_rev(row, col)
    p = O.or_params(horizon=3.015, dx=1.0, E=3.5e4, nu=0.2)
    dp = O.damage_params()
    broken = np.zeros(len(col), np.uint8)
    broken[row[20]:row[20] + 3] = 1
    broken[rev[row[20]:row[20] + 3]] = 1
    kappa = np.full(len(col), KAPPA0)
    dmg = np.zeros(len(col))
    _, status, changed = O.forces_damage(p, dp, X, V, row, col, np.zeros(X.size), kappa, dmg,
                                         broken=broken, rev=rev)
    assert changed == broken.sum()
    np.testing.assert_array_equal(dmg, broken.astype(float))
    m = np.diff(row)
    np.testing.assert_allclose(status, np.add.reduceat(broken.astype(float), row[:-1]) / m)


@pytest.mark.parametrize("mode,kappa_prev", [(O.DAMAGE_REFERENCE, 1.0), (O.DAMAGE_MAZARS, 1.0),
                                             (O.DAMAGE_REFERENCE, 1.3), (O.DAMAGE_MAZARS, 1.3)])
def test_tension_law_rate_constants_at_interior_points(mode, kappa_prev):
    """Away from its limits the tension law pins A_t and B_t: for a loading
    past the threshold kappa_r used in the law (kappa0 in the published law,
    the running threshold in the reference's, src/main_seq.cpp:579-584),
    ln(1 - (1 - A_t) kappa_r / eps - d) = ln A_t - B_t (eps - kappa_r) is a
    straight line in eps with slope -B_t.  Sampled where the exponential is
    between 0.3 and 0.9 of A_t, so a swapped B_t / B_c (slope -2150) or a wrong
    A_t fails."""
    dp = O.damage_params(mode=mode)
    kap = kappa_prev * KAPPA0
    kr = KAPPA0 if mode == O.DAMAGE_MAZARS else kap
    eps = kap * np.linspace(1.02, 1.6, 12)
    y = []
    for e in eps:
        upd, k, d = O.mazars_update(dp, e, 0.0, 0.0, kap)
        assert upd and k == e
        y.append(np.log(1.0 - (1.0 - dp.At) * kr / e - d))
    slope, icpt = np.polyfit(eps - kr, np.array(y), 1)
    assert slope == pytest.approx(-dp.Bt, rel=1e-9)
    assert icpt == pytest.approx(np.log(dp.At), rel=1e-9)
    # a planted swap of the tension and compression rates moves the line
    bad = O.damage_params(mode=mode, Bt=dp.Bc, Bc=dp.Bt)
    yb = [np.log(1.0 - (1.0 - bad.At) * kr / e - O.mazars_update(bad, e, 0.0, 0.0, kap)[2]) for e in eps]
    assert np.polyfit(eps - kr, np.array(yb), 1)[0] == pytest.approx(-dp.Bc, rel=1e-9)


def test_compression_law_is_unreachable():
    """strain_c holds the non-positive principal parts (src/main_seq.cpp:556-562),
    so its eigenvalues are never > 0 and alpha_c stays 0 (:571-577): the
    compression constants A_c, B_c never enter d.  For tension-compression
    states alpha_t = 1 (the ascending spectra pair l2 with 0 and 0 with l1),
    so d is the tension law of the positive part alone."""
    rng = np.random.default_rng(3)
    dp = O.damage_params()
    odd = O.damage_params(Ac=0.05, Bc=7.0e5)
    for _ in range(300):
        lam = np.sort(rng.uniform(-6, 6, size=2)) * KAPPA0
        R = rot(rng.uniform(0, np.pi))
        eps = R @ np.diag(lam) @ R.T
        a = O.mazars_update(dp, eps[0, 0], eps[1, 1], eps[0, 1], KAPPA0)
        b = O.mazars_update(odd, eps[0, 0], eps[1, 1], eps[0, 1], KAPPA0)
        assert a == b
        eq = np.sqrt(np.sum(np.maximum(lam, 0.0) ** 2))
        if eq > KAPPA0:
            assert a[0]
            d_t = 1.0 - KAPPA0 * (1 - dp.At) / eq - dp.At * np.exp(-dp.Bt * (eq - KAPPA0))
            assert a[2] == pytest.approx(d_t, rel=1e-12, abs=1e-15)
