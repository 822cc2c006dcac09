"""Parity of the HIP path (through the C ABI) with the CPU oracle.

FP64: <= 1e-12 norm-wise on velocities after one step (Hassa mesh, reference
parameters), a stated drift bound after many steps with bond breaking, and
identical broken-bond sets.  Mixed precision: <= 1e-5 norm-wise velocity error
against the FP64 oracle.  Norm-wise metric as in tests/hassa.parity_norms
(per-element relative error is meaningless where velocities are round-off).
"""
import os
import subprocess

import numpy as np
import pytest

import hassa as H
import oracle_lib as O
from conftest import ROOT, say

pytestmark = pytest.mark.gpu


def hassa_params(pdh, precision=0):
    return pdh.Params.make(precision=precision, n1=H.N1, n2=H.N2, horizon=H.HORIZON, dx=H.DX,
                           rho=H.RHO, dt=H.DT, E=H.E, nu=H.NU, s0=0.0)


@pytest.fixture(scope="module")
def hassa_case():
    return H.case()


@pytest.fixture(scope="module")
def hassa_oracle(hassa_case):
    X, V, row, col, sets = hassa_case
    n = V.size
    po = O.or_params(horizon=H.HORIZON, dx=H.DX, rho=H.RHO, dt=H.DT, E=H.E, nu=H.NU)
    bc = np.full(n, -1, np.int32)
    bc[sets[3]] = 0
    bc[sets[2]] = -2
    rev = O.build_rev(row, col)
    u, v, a, f = (np.zeros(2 * n) for _ in range(4))
    O.step(po, X.astype(float), V.astype(float), row, col, rev, bc, np.array(H.V_TOP), u, v, a,
           f, None, 1)
    T = O.forces(po, X.astype(float), V.astype(float), row, col, u, rev=rev, want_T=True)[1]
    return u, v, a, f, T


def hassa_gpu(pdpkg, hassa_case, via="phases", precision=0, tile=None):
    pdh = pdpkg.pdhip
    X, V, row, col, sets = hassa_case
    ctx = pdh.Context(hassa_params(pdh, precision), X.astype(float), V.astype(float), row, col)
    if tile is not None:
        ctx.set_option(pdh.OPT_TILE, tile)
    if via == "phases":
        # src/main_seq.cpp:686-699
        ctx.apply_velocity_bc(sets[3], H.V_TOP)
        ctx.update_positions()
        ctx.apply_fixed_bc(sets[2])
        ctx.compute_velocity()
    else:
        code = np.full(V.size, pdh.BC_FREE, np.int32)
        code[sets[3]] = 0
        code[sets[2]] = pdh.BC_FIXED
        ctx.set_boundary(code, H.V_TOP)
        ctx.step(1)
    st = ctx.get_state()
    st["T"] = ctx.get_bond_forces()
    ctx.close()
    return st


def bc_mask(hassa_case):
    X, V, row, col, sets = hassa_case
    m = np.zeros(V.size, bool)
    m[sets[3]] = True
    return m


def test_hassa_step_fp64_matches_oracle(pdpkg, hassa_case, hassa_oracle):
    say("hassa fp64 single step, n=4064")
    st = hassa_gpu(pdpkg, hassa_case)
    u, v, a, f, T = hassa_oracle
    np.testing.assert_array_equal(st["u"], u)
    all_rel, nonbc, elem = H.parity_norms(st["v"], v, bc_mask(hassa_case))
    assert all_rel < 1e-14
    assert nonbc < 1e-12
    assert elem < 1e-12
    assert np.linalg.norm(st["f"] - f) < 1e-12 * np.linalg.norm(f)
    assert np.linalg.norm(st["T"] - T) < 1e-12 * np.linalg.norm(T)


def test_hassa_step_matches_reference_fixture(pdpkg, hassa_case):
    say("hassa fp64 vs velocity_step_0_after_seq_ref.txt")
    st = hassa_gpu(pdpkg, hassa_case)
    vfix = H.load_velocity_fixture("velocity_step_0_after_seq_ref.txt").reshape(-1)
    _, nonbc, elem = H.parity_norms(st["v"], vfix, bc_mask(hassa_case))
    assert nonbc < 5e-6 and elem < 5e-6


def test_hassa_fused_step_equals_phase_calls(pdpkg, hassa_case):
    say("hassa fused step vs phase calls")
    a = hassa_gpu(pdpkg, hassa_case, via="phases")
    b = hassa_gpu(pdpkg, hassa_case, via="step")
    for k in ("u", "v", "f"):
        np.testing.assert_array_equal(a[k], b[k])


def test_hassa_mixed_precision_error(pdpkg, hassa_case, hassa_oracle):
    say("hassa mixed single step")
    st = hassa_gpu(pdpkg, hassa_case, precision=1)
    _, v, _, f, _ = hassa_oracle
    _, nonbc, _ = H.parity_norms(st["v"], v, bc_mask(hassa_case))
    assert nonbc < 1e-5


def test_reference_signature_adapter_driver(hassa_case, tmp_path):
    """A main.cu-style driver built against include/pd_mechanics_adapter.hpp
    (compute_velocity_GPU_host / computeVelocity signatures) reproduces the
    oracle given the same long-double position update."""
    say("adapter driver (reference signatures)")
    X, V, row, col, sets = hassa_case
    exe = os.path.join(ROOT, "tests", "adapter", "build", "hassa_driver")
    n = V.size
    p = O.ref_params()
    y = X.copy()
    vl = np.zeros(2 * n, np.longdouble)
    al = np.zeros_like(vl)
    fl = np.zeros_like(vl)
    O.ref_apply_velocity_bc(sets[3], H.V_TOP, vl, al)
    O.ref_update_positions(p, y, vl, al, fl)
    O.ref_apply_fixed_bc(sets[2], X, y)
    u = (y - X).astype(np.float64)
    po = O.or_params(horizon=H.HORIZON, dx=H.DX, rho=H.RHO, dt=H.DT, E=H.E, nu=H.NU)
    f, _, _, _ = O.forces(po, X.astype(float), V.astype(float), row, col, u)
    v0 = vl.astype(np.float64)
    v = v0 + (0.5 * (al.astype(np.float64) + f / H.RHO)) * H.DT
    for entry in ("gpu", "seq"):
        out = tmp_path / f"v_{entry}.txt"
        r = subprocess.run([exe, H.MESH, str(out), entry], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        vg = np.loadtxt(out).reshape(-1)
        _, nonbc, elem = H.parity_norms(vg, v, bc_mask(hassa_case))
        assert nonbc < 1e-12 and elem < 1e-12


def test_padded_layout_equals_csr(pdpkg, hassa_case):
    say("padded reference layout (cap 30, INT_MAX pad)")
    pdh = pdpkg.pdhip
    X, V, row, col, sets = hassa_case
    n = V.size
    cap = 30
    nbr = np.full(n * cap, np.iinfo(np.int32).max, np.int32)
    cnt = np.diff(row).astype(np.int32)
    for i in range(n):
        nbr[i * cap: i * cap + cnt[i]] = col[row[i]: row[i + 1]]
    ctx = pdh.Context.from_padded(hassa_params(pdh), X.astype(float), V.astype(float), nbr, cnt, cap)
    ctx.apply_velocity_bc(sets[3], H.V_TOP)
    ctx.update_positions()
    ctx.apply_fixed_bc(sets[2])
    ctx.compute_velocity(write_acceleration=True)
    st = ctx.get_state()
    ctx.close()
    ref = hassa_gpu(pdpkg, hassa_case)
    np.testing.assert_array_equal(st["v"], ref["v"])
    # the reference GPU kernel's a = f / rho write-back
    np.testing.assert_array_equal(st["a"], st["f"] / H.RHO)


def test_state_roundtrip_and_positions(pdpkg):
    pdh = pdpkg.pdhip
    say("state round trip")
    w = pdpkg.plates.single_plate("rt", 20, 17, 10.0, seed=2)
    ctx = pdh.Context(w.params(), w.X, w.V, w.row, w.col)
    rng = np.random.default_rng(0)
    arrs = [rng.normal(size=2 * w.n) for _ in range(4)]
    ctx.set_state(*arrs)
    st = ctx.get_state()
    for k, x in zip("uvaf", arrs):
        np.testing.assert_array_equal(st[k], x)
    y = w.X + 1e-6 * arrs[0]
    ctx.set_positions(y)
    np.testing.assert_allclose(ctx.get_positions(), y, rtol=0, atol=1e-18)
    m = np.zeros(w.nb, np.uint8)
    ctx.set_broken(m)
    assert ctx.get_broken().sum() == 0
    with pytest.raises(pdh.PDError):
        bad = m.copy()
        bad[0] = 1  # asymmetric
        ctx.set_broken(bad)
    ctx.close()


def run_plate_oracle(w, s0, steps, E=None):
    p = O.or_params(horizon=w.horizon, dx=w.dx, rho=8000.0, dt=1e-8, E=192e9, nu=0.3, s0=s0)
    rev = O.build_rev(w.row, w.col)
    u, v, a, f = (np.zeros(2 * w.n) for _ in range(4))
    br = w.broken.copy()
    O.step(p, w.X, w.V, w.row, w.col, rev, w.code, w.v0, u, v, a, f, br, steps, E_arr=E)
    return u, v, f, br


def run_plate_gpu(pdh, w, s0, steps, precision=0, E=None, chunks=1, graph=True, tile=None, profile=False):
    ctx = w.context(precision=precision, s0=s0)
    if tile is not None:
        ctx.set_option(pdh.OPT_TILE, tile)
    if E is not None:
        ctx.set_youngs_modulus(E)
    ctx.set_option(pdh.OPT_USE_GRAPH, int(graph))
    ctx.set_option(pdh.OPT_PROFILE, int(profile))
    for _ in range(chunks):
        ctx.step(steps // chunks)
    st = ctx.get_state()
    br = ctx.get_broken()
    stats = ctx.stats()
    ctx.close()
    return st, br, stats


@pytest.fixture(scope="module")
def small_plate(pdpkg):
    return pdpkg.plates.single_plate("p48x40", 48, 40, 30.0, notch_len_cells=16, seed=3)


@pytest.fixture(scope="module")
def small_plate_oracle(small_plate):
    return run_plate_oracle(small_plate, 0.005, 300)


def test_plate_fp64_breaking_matches_oracle(pdpkg, small_plate, small_plate_oracle):
    say("plate 48x40 fp64, 300 steps, s0 = 0.005")
    u, v, f, br = small_plate_oracle
    assert br.sum() > small_plate.broken.sum() + 1000  # bonds did break
    st, brg, stats = run_plate_gpu(pdpkg.pdhip, small_plate, 0.005, 300)
    np.testing.assert_array_equal(brg, br)
    assert stats.n_broken == br.sum()
    mask = small_plate.code >= 0
    _, nonbc, _ = H.parity_norms(st["v"], v, mask)
    # drift bound after 300 steps (round-off of ~1e-16 per step amplified by the
    # explicit dynamics; measured ~1e-13)
    assert nonbc < 1e-10
    assert np.linalg.norm(st["u"] - u) < 1e-10 * np.linalg.norm(u)


def test_plate_mixed_breaking_vs_oracle(pdpkg, small_plate, small_plate_oracle):
    say("plate 48x40 mixed, 300 steps")
    u, v, f, br = small_plate_oracle
    st, brg, _ = run_plate_gpu(pdpkg.pdhip, small_plate, 0.005, 300, precision=1)
    mask = small_plate.code >= 0
    _, nonbc, _ = H.parity_norms(st["v"], v, mask)
    assert nonbc < 1e-5
    np.testing.assert_array_equal(brg, br)


@pytest.mark.parametrize("precision", [0, 1])
def test_plate_graph_and_chunks_bitwise(pdpkg, small_plate, precision):
    """Graph replay, eager launches, chunked pd_step calls and the per-phase
    profiling path (no fused gather + position update) give the same bits."""
    say(f"graph replay vs eager vs per-phase, chunked steps, precision {precision}")
    pdh = pdpkg.pdhip
    runs = [run_plate_gpu(pdh, small_plate, 0.005, 60, precision=precision, graph=True),
            run_plate_gpu(pdh, small_plate, 0.005, 60, precision=precision, graph=False),
            run_plate_gpu(pdh, small_plate, 0.005, 60, precision=precision, graph=True, chunks=3),
            run_plate_gpu(pdh, small_plate, 0.005, 60, precision=precision, profile=True)]
    a, ba, _ = runs[0]
    assert ba.sum() > small_plate.broken.sum()  # bonds broke during the run
    for b, bb, _ in runs[1:]:
        for k in ("u", "v", "a", "f"):
            np.testing.assert_array_equal(a[k], b[k])
        np.testing.assert_array_equal(ba, bb)


def test_launch_shapes_and_autotune_bitwise(pdpkg, small_plate):
    say("launch-shape variants and autotune")
    pdh = pdpkg.pdhip
    ref, bref, _ = run_plate_gpu(pdh, small_plate, 0.005, 40)
    mask = small_plate.code >= 0
    for tile in (3201, 6401, 1602, 1603, 13201, 16401, 11602, 11603, 23201, 26401, 21603, 31602, 31603):
        tile_ref = None
        for wpb, xcd, persist in ((1, 0, 0), (2, 1, 0), (8, 0, 0), (8, 1, 0), (4, 1, 1), (2, 0, 3),
                                  (8, 1, 2)):
            if True:
                ctx = small_plate.context(s0=0.005)
                ctx.set_option(pdh.OPT_WAVES_PER_WG, wpb)
                ctx.set_option(pdh.OPT_XCD_REMAP, xcd)
                ctx.set_option(pdh.OPT_TILE, tile)
                ctx.set_option(pdh.OPT_PERSISTENT, persist)
                ctx.step(40)
                v = ctx.get_state()["v"]
                np.testing.assert_array_equal(ctx.get_broken(), bref)
                ctx.close()
                # the workgroup shape and XCD mapping never change the bits
                if tile_ref is None:
                    tile_ref = v
                np.testing.assert_array_equal(v, tile_ref)
        # two bonds per lane and the matrix-core contractions reorder the FP64
        # family sums: round-off only
        _, nonbc, _ = H.parity_norms(tile_ref, ref["v"], mask)
        assert nonbc < 1e-13
    mref = None
    for tile in (3201, 6401, 1602, 3202, 1603, 13201, 16401, 11602, 13202, 11603, 31602, 31603, 41602, 41603):
        ctx = small_plate.context(precision=1, s0=0.005)
        ctx.set_option(pdh.OPT_TILE, tile)
        ctx.step(40)
        v = ctx.get_state()["v"]
        ctx.close()
        if mref is None:
            mref = v
        _, nonbc, _ = H.parity_norms(v, mref, mask)
        assert nonbc < 1e-6
    with pytest.raises(pdh.PDError):
        ctx = small_plate.context(s0=0.005)
        try:
            ctx.set_option(pdh.OPT_TILE, 3202)  # FP64 has no (32, 2) tile
        finally:
            ctx.close()
    with pytest.raises(pdh.PDError):
        ctx = small_plate.context(s0=0.005)
        try:
            ctx.set_option(pdh.OPT_TILE, 13202)  # nor a matrix-core (32, 2) tile
        finally:
            ctx.close()
    with pytest.raises(pdh.PDError):
        ctx = small_plate.context(precision=1, s0=0.005)
        try:
            ctx.set_option(pdh.OPT_TILE, 21603)  # matrix-core weights are FP64 only
        finally:
            ctx.close()
    ctx = small_plate.context(s0=0.005)
    ctx.autotune(3)
    st = ctx.stats()
    assert st.steps == 0  # the autotuner restores the state it timed on
    ctx.step(40)
    _, nonbc, _ = H.parity_norms(ctx.get_state()["v"], ref["v"], mask)
    assert nonbc < 1e-13
    np.testing.assert_array_equal(ctx.get_broken(), bref)
    ctx.close()


def test_hassa_matrix_core_contractions(pdpkg, hassa_case, hassa_oracle):
    """Family contractions on the matrix cores (v_mfma_f64_4x4x4 in FP64,
    v_mfma_f32_4x4x1 in mixed precision) meet the same bars as the VALU path."""
    say("hassa single step, matrix-core contractions")
    u, v, a, f, T = hassa_oracle
    for tile in (13201, 16401, 11602, 11603, 23201, 26401, 21603, 31602, 31603):
        st = hassa_gpu(pdpkg, hassa_case, tile=tile)
        np.testing.assert_array_equal(st["u"], u)
        _, nonbc, elem = H.parity_norms(st["v"], v, bc_mask(hassa_case))
        assert nonbc < 1e-12 and elem < 1e-12
        assert np.linalg.norm(st["T"] - T) < 1e-12 * np.linalg.norm(T)
    for tile in (13201, 16401, 11602, 13202, 31602, 41602, 41603):
        st = hassa_gpu(pdpkg, hassa_case, precision=1, tile=tile)
        _, nonbc, _ = H.parity_norms(st["v"], v, bc_mask(hassa_case))
        assert nonbc < 1e-5


@pytest.mark.parametrize("tile", [13201, 16401, 11602, 23201, 21603, 31602, 31603])
def test_plate_matrix_core_breaking_matches_oracle(pdpkg, small_plate, small_plate_oracle, tile):
    say(f"plate 48x40, 300 steps, s0 = 0.005, matrix-core tile {tile}")
    pdh = pdpkg.pdhip
    u, v, f, br = small_plate_oracle
    mask = small_plate.code >= 0
    st, brg, stats = run_plate_gpu(pdh, small_plate, 0.005, 300, tile=tile)
    assert stats.matrix_cores == tile // 10000
    np.testing.assert_array_equal(brg, br)
    _, nonbc, _ = H.parity_norms(st["v"], v, mask)
    assert nonbc < 1e-10
    assert np.linalg.norm(st["u"] - u) < 1e-10 * np.linalg.norm(u)
    if tile // 10000 == 2:
        return  # matrix-core weights run in FP64 only
    if tile // 10000 == 3:
        tile += 10000  # rank-form weights with matrix-core contractions in mixed precision
    st, brg, _ = run_plate_gpu(pdh, small_plate, 0.005, 300, precision=1, tile=tile)
    np.testing.assert_array_equal(brg, br)
    _, nonbc, _ = H.parity_norms(st["v"], v, mask)
    assert nonbc < 1e-5


@pytest.mark.parametrize("precision,tile", [(0, 0), (1, 0), (0, 13201), (1, 13201), (0, 1603), (1, 11603),
                                            (0, 11603), (0, 21603), (0, 23201), (0, 31603), (1, 41603),
                                            (1, 31602)])
def test_persistent_bond_kernel_bitwise(pdpkg, precision, tile):
    """The persistent bond kernel (each wave walks many family groups and
    prefetches the next one) gives the same bits as one wave per group."""
    say(f"persistent bond kernel, 200 x 200 plate, precision {precision}, tile {tile}")
    pdh = pdpkg.pdhip
    w = pdpkg.plates.single_plate("p200", 200, 200, 40.0, notch_len_cells=40, seed=8)
    out = []
    for persist, wpb in ((0, 4), (1, 4), (2, 2), (1, 8)):
        ctx = w.context(precision=precision, s0=0.003)
        ctx.set_option(pdh.OPT_TILE, tile)
        ctx.set_option(pdh.OPT_PERSISTENT, persist)
        ctx.set_option(pdh.OPT_WAVES_PER_WG, wpb)
        ctx.step(30)
        st = ctx.get_state()
        out.append((st["v"], st["f"], ctx.get_broken()))
        ctx.close()
    assert out[0][2].sum() > w.broken.sum()  # bonds broke during the run
    for v, f, b in out[1:]:
        np.testing.assert_array_equal(v, out[0][0])
        np.testing.assert_array_equal(f, out[0][1])
        np.testing.assert_array_equal(b, out[0][2])


def test_per_particle_youngs_modulus(pdpkg):
    say("per-particle E, 50 steps")
    w = pdpkg.plates.single_plate("E", 30, 30, 20.0, notch_len_cells=8, seed=4, E_sigma=0.05)
    u, v, f, br = run_plate_oracle(w, 0.0, 50, E=w.E)
    st, _, _ = run_plate_gpu(pdpkg.pdhip, w, 0.0, 50)
    _, nonbc, _ = H.parity_norms(st["v"], v, w.code >= 0)
    assert nonbc < 1e-11


def test_delta_4015_family_of_48(pdpkg):
    say("delta = 4.015 dx, families of 48: 64 lanes per family, or 16 lanes x 3 bonds")
    w = pdpkg.plates.single_plate("d4", 24, 22, 20.0, horizon=4.015, seed=5, porosity=0.01)
    assert np.diff(w.row).max() == 48
    u, v, f, br = run_plate_oracle(w, 0.0, 30)
    st, _, stats = run_plate_gpu(pdpkg.pdhip, w, 0.0, 30)
    assert stats.lanes_per_family * stats.bonds_per_lane >= 48
    _, nonbc, _ = H.parity_norms(st["v"], v, w.code >= 0)
    assert nonbc < 1e-11
    for tile in (16401, 1603, 11603, 21603, 26401, 31603):
        st, _, _ = run_plate_gpu(pdpkg.pdhip, w, 0.0, 30, tile=tile)
        _, nonbc, _ = H.parity_norms(st["v"], v, w.code >= 0)
        assert nonbc < 1e-11
    for tile in (6401, 16401, 1603, 11603, 31603, 41603):
        st, _, _ = run_plate_gpu(pdpkg.pdhip, w, 0.0, 30, precision=1, tile=tile)
        _, nonbc, _ = H.parity_norms(st["v"], v, w.code >= 0)
        assert nonbc < 1e-5
    # K^-1 assembled again by the (16, 3) shape kernel (pd_set_broken with the
    # tile set) gives the same step
    pdh = pdpkg.pdhip
    for prec, tol in ((0, 1e-11), (1, 1e-5)):
        ctx = w.context(precision=prec, s0=0.0)
        ctx.set_option(pdh.OPT_TILE, 1603)
        ctx.set_broken(ctx.get_broken())
        ctx.step(30)
        _, nonbc, _ = H.parity_norms(ctx.get_state()["v"], v, w.code >= 0)
        ctx.close()
        assert nonbc < tol


def test_families_of_60_in_64_slot_tiles(pdpkg):
    """The largest families the tiles take: delta = 4.4 dx gives 60 bonds per
    interior family (64-slot tiles) against the oracle, with bonds breaking for
    200 steps, until 7784 of the 21872 bonds are broken.

    From about step 100 on a few bonds are left with nearly collinear
    neighbours and K condition numbers up to 4.3e9, which FP32 sums cannot
    resolve (their error grows as 6e-8 x cond(K)).  Mixed precision keeps the
    1e-5 bound and the FP64 broken set there through its FP64 guard (the
    families holding such a bond are evaluated again in FP64 every step); with
    the guard off it is checked over the first 50 steps (largest cond(K) of an
    alive bond 8.4e3), where it holds without it."""
    say("delta = 4.4 dx, families of 60, 64-slot tiles")
    pdh = pdpkg.pdhip
    w = pdpkg.plates.single_plate("h44", 22, 20, 30.0, horizon=4.4, seed=6, notch_len_cells=8)
    assert np.diff(w.row).max() == 60
    mask = w.code >= 0
    cases = ((0, 200, (6401, 16401), 1e-10, 4, None),
             (1, 200, (6401, 16401, 3202, 13202), 1e-5, 4, 1000),
             (1, 50, (6401, 16401, 3202, 13202), 1e-5, 0, 0))
    for prec, steps, tiles, tol, n_ill, guard in cases:
        u, v, f, br = run_plate_oracle(w, 0.004, steps)
        assert br.sum() > w.broken.sum() + 500   # 898 broken by the notch
        assert ill_conditioned_oracle(w, u, br) == n_ill
        for tile in tiles:
            ctx = w.context(precision=prec, s0=0.004)
            ctx.set_option(pdh.OPT_TILE, tile)
            if guard is not None:
                ctx.set_option(pdh.OPT_MIXED_GUARD, guard)
            ctx.step(steps)
            st, brg, stats = ctx.get_state(), ctx.get_broken(), ctx.stats()
            ctx.close()
            np.testing.assert_array_equal(brg, br)
            _, nonbc, _ = H.parity_norms(st["v"], v, mask)
            assert nonbc < tol, (prec, steps, tile, nonbc)
            assert stats.n_ill_conditioned == n_ill, (prec, tile, stats.n_ill_conditioned)
            if guard:
                # the families holding the ill-conditioned bonds, and those with
                # 1e3 <= cond(K) < 1e4
                assert stats.n_guarded_families >= n_ill


def test_mixed_guard_off_loses_the_fragmented_plate(pdpkg):
    """The same plate at 200 steps with the guard off: the FP32 error exceeds
    the bound, which is what the guard is for (a control for the test above)."""
    say("delta = 4.4 dx, mixed precision without its FP64 guard")
    pdh = pdpkg.pdhip
    w = pdpkg.plates.single_plate("h44", 22, 20, 30.0, horizon=4.4, seed=6, notch_len_cells=8)
    u, v, f, br = run_plate_oracle(w, 0.004, 200)
    ctx = w.context(precision=1, s0=0.004)
    ctx.set_option(pdh.OPT_TILE, 6401)
    ctx.set_option(pdh.OPT_MIXED_GUARD, 0)
    ctx.step(200)
    st, stats = ctx.get_state(), ctx.stats()
    ctx.close()
    _, nonbc, _ = H.parity_norms(st["v"], v, w.code >= 0)
    assert nonbc > 1e-4
    assert stats.n_guarded_families == 0
    # the option's error paths: FP64 contexts have no guard, bounds are >= 0
    ctx = w.context(precision=0, s0=0.004)
    ctx.set_option(pdh.OPT_MIXED_GUARD, 0)
    with pytest.raises(pdh.PDError):
        ctx.set_option(pdh.OPT_MIXED_GUARD, 1000)
    assert ctx.stats().n_guarded_families == 0
    ctx.close()
    ctx = w.context(precision=1, s0=0.004)
    with pytest.raises(pdh.PDError):
        ctx.set_option(pdh.OPT_MIXED_GUARD, -1)
    ctx.close()


def test_mixed_guard_graph_eager_chunks_bitwise(pdpkg):
    """With guarded families, the graph-replayed step, the eager one and the
    same steps in chunks give identical bits (the FP64 fix-up of each family is
    deterministic and does not depend on the order the families were collected
    in), and so does the per-phase profiled path."""
    say("mixed guard: graph vs eager vs chunks vs profiled, bitwise")
    pdh = pdpkg.pdhip
    w = pdpkg.plates.single_plate("h44", 22, 20, 30.0, horizon=4.4, seed=6, notch_len_cells=8)
    runs = []
    for graph, chunks, profile in ((1, 1, 0), (0, 1, 0), (1, 4, 0), (1, 1, 1)):
        ctx = w.context(precision=1, s0=0.004)
        ctx.set_option(pdh.OPT_TILE, 3202)
        ctx.set_option(pdh.OPT_USE_GRAPH, graph)
        ctx.set_option(pdh.OPT_PROFILE, profile)
        for _ in range(chunks):
            ctx.step(200 // chunks)
        runs.append((ctx.get_state(), ctx.get_broken(), ctx.stats().n_guarded_families))
        ctx.close()
    assert runs[0][2] > 0
    for st, b, ng in runs[1:]:
        for k in ("u", "v", "a", "f"):
            np.testing.assert_array_equal(st[k], runs[0][0][k])
        np.testing.assert_array_equal(b, runs[0][1])
        assert ng == runs[0][2]


def test_fracturing_plate_mixed_tracks_fp64_until_it_shatters(pdpkg):
    """A 192 x 192 plate (centre notch, +-20 m/s, s0 = 0.006, 5 % E scatter)
    whose load waves meet at the centre line after about 2000 steps and shatter
    it.  Up to step 1500 the guarded mixed run keeps the FP64 velocities within
    1e-5 and the FP64 broken set; after that the fracture amplifies any
    perturbation (two FP64 tiles part there too), so the guarded run is only
    required to stay finite through step 3000 (every pd_step checks)."""
    say("fracturing 192 x 192 plate: guarded mixed vs FP64")
    pdh = pdpkg.pdhip
    w = pdpkg.plates.single_plate("frac192", 192, 192, 20.0, notch_len_cells=48, seed=11, E_sigma=0.05)
    out = []
    for prec in (0, 1):
        ctx = w.context(precision=prec, s0=0.006)
        ctx.set_option(pdh.OPT_TILE, 1602)
        ctx.step(1500)
        out.append((ctx, ctx.get_state(), ctx.get_broken()))
    (c64, a, ba), (cm, m, bm) = out
    assert ba.sum() > w.broken.sum() + 1000
    np.testing.assert_array_equal(bm, ba)
    _, nonbc, _ = H.parity_norms(m["v"], a["v"], w.code >= 0)
    assert nonbc < 1e-5, nonbc
    c64.close()
    cm.step(1500)                      # through the shattering: stays finite
    st = cm.stats()
    assert st.n_broken > 10 * ba.sum() and st.n_guarded_families > 0
    cm.close()


def test_guard_delta_4015_heavy_fracture_against_oracle(pdpkg):
    """delta = 4.015 dx (families of 48) with s0 = 0.004 for 300 steps, 13046
    bonds broken: the guarded mixed tiles, including the three-bonds-per-lane
    ones, keep the oracle's broken set and 1e-5; without the guard the same run
    leaves both."""
    say("delta = 4.015 dx heavy fracture: guarded mixed tiles vs the oracle")
    pdh = pdpkg.pdhip
    w = pdpkg.plates.single_plate("d4", 48, 40, 30.0, horizon=4.015, seed=3, notch_len_cells=16)
    u, v, f, br = run_plate_oracle(w, 0.004, 300)
    assert br.sum() > 10000
    mask = w.code >= 0
    for tile in (1603, 11603, 3202, 6401, 41603):
        st, brg, stats = run_plate_gpu(pdh, w, 0.004, 300, precision=1, tile=tile)
        np.testing.assert_array_equal(brg, br)
        _, nonbc, _ = H.parity_norms(st["v"], v, mask)
        assert nonbc < 1e-5, (tile, nonbc)
        assert stats.n_guarded_families > 0
    ctx = w.context(precision=1, s0=0.004)
    ctx.set_option(pdh.OPT_TILE, 1603)
    ctx.set_option(pdh.OPT_MIXED_GUARD, 0)
    ctx.step(300)
    st = ctx.get_state()
    ctx.close()
    _, nonbc, _ = H.parity_norms(st["v"], v, mask)
    assert nonbc > 1e-4


def test_guard_persistent_kernels_bitwise(pdpkg):
    """The persistent bond kernels collect the guarded families from their
    prefetched groups: on the fracturing 192 x 192 plate (over 200 guarded
    families at step 2600) they give the plain launch's bits."""
    say("guarded families: persistent vs plain bond kernel, bitwise")
    pdh = pdpkg.pdhip
    w = pdpkg.plates.single_plate("frac192", 192, 192, 20.0, notch_len_cells=48, seed=11, E_sigma=0.05)

    def run(tile, persist, wpb):
        ctx = w.context(precision=1, s0=0.006)
        ctx.set_option(pdh.OPT_TILE, tile)
        ctx.set_option(pdh.OPT_WAVES_PER_WG, wpb)
        ctx.set_option(pdh.OPT_PERSISTENT, persist)
        ctx.step(2600)
        r = ctx.get_state(), ctx.get_broken(), ctx.stats().n_guarded_families
        ctx.close()
        return r

    for tile in (1602, 1603):
        a = run(tile, 0, 4)
        assert a[2] > 100
        for persist, wpb in ((1, 4), (8, 1)):
            b = run(tile, persist, wpb)
            for k in ("u", "v", "a", "f"):
                np.testing.assert_array_equal(b[0][k], a[0][k])
            np.testing.assert_array_equal(b[1], a[1])


def test_autotune_restores_guard_flags(pdpkg):
    """pd_autotune times every variant on the context's own state, whose trial
    steps break bonds and re-assemble K (and with it the guard flags): the state,
    the broken set and the guarded families are as before the call."""
    say("autotune restores state and guard flags")
    w = pdpkg.plates.single_plate("h44", 22, 20, 30.0, horizon=4.4, seed=6, notch_len_cells=8)
    ctx = w.context(precision=1, s0=0.004)
    ctx.step(150)
    st0, br0, ng0 = ctx.get_state(), ctx.get_broken(), ctx.stats().n_guarded_families
    assert ng0 > 0
    ctx.autotune(5)
    st1, br1, ng1 = ctx.get_state(), ctx.get_broken(), ctx.stats().n_guarded_families
    ctx.close()
    for k in ("u", "v", "a", "f"):
        np.testing.assert_array_equal(st1[k], st0[k])
    np.testing.assert_array_equal(br1, br0)
    assert ng1 == ng0


def ill_conditioned_oracle(w, u, br, cmin=1e4):
    """Alive bonds whose K has condition number >= cmin, from the oracle's K^-1
    (eigenvalues by numpy)."""
    p = O.or_params(horizon=w.horizon, dx=w.dx, rho=8000.0, dt=1e-8, E=192e9, nu=0.3, s0=0.0)
    _, _, Ki, _ = O.forces(p, w.X, w.V, w.row, w.col, u, broken=br, rev=O.build_rev(w.row, w.col),
                           want_Kinv=True)
    Ki = Ki.reshape(-1, 3)
    keep = (br == 0) & np.any(Ki != 0.0, axis=1)
    m = np.stack([Ki[keep, 0], Ki[keep, 1], Ki[keep, 1], Ki[keep, 2]], axis=1).reshape(-1, 2, 2)
    lam = np.abs(np.linalg.eigvalsh(m))
    return int(np.sum(lam.max(axis=1) >= cmin * lam.min(axis=1)))


def test_isolated_particles_and_tiny_families(pdpkg):
    """Ragged input next to a plate: particles with no family, a two-particle
    body (one bond each way: K is rank one, so singular, and carries no force)
    and a three-particle body (families of two bonds, one corner driven); every
    tile against the oracle."""
    say("isolated particles, one- and two-bond families")
    pl = pdpkg.plates
    base = pl.single_plate("rag", 20, 16, 20.0, notch_len_cells=6, seed=9)
    dx = base.dx
    far = 40 * dx
    extra = np.array([far, far, far + 10 * dx, far, far, far + 10 * dx,            # isolated
                      -far, -far, -far + 1.5 * dx, -far,                            # pair
                      far, -far, far + 1.5 * dx, -far, far, -far + 1.5 * dx])       # triangle
    X = np.concatenate([base.X, extra])
    n = X.size // 2
    V = np.full(n, base.V[0])
    row, col = pdpkg.pdhip.build_families(X, dx, base.horizon)
    lens = np.diff(row)[base.n:]
    np.testing.assert_array_equal(lens, [0, 0, 0, 1, 1, 2, 2, 2])
    code = np.concatenate([base.code, np.array([-1, -1, -1, 0, -1, 0, -1, -1], dtype=np.int32)])
    broken = np.zeros(len(col), dtype=np.uint8)
    broken[:len(base.col)] = base.broken  # the plate's bonds come first (same families)
    w = pl.Workload("rag+", X, V, row, col, code, base.v0, broken, horizon=base.horizon, dx=dx)
    u, v, f, br = run_plate_oracle(w, 0.0, 40)
    fx = f.reshape(-1, 2)[base.n:]
    assert np.all(fx[:5] == 0.0) and np.abs(fx[5:]).max() > 0
    mask = w.code >= 0
    for prec, tiles, tol in ((0, (3201, 6401, 1602, 1603, 13201, 16401, 11602, 11603, 23201, 26401, 21603, 31602,
                                  31603), 1e-11),
                             (1, (3201, 1602, 3202, 11602, 1603, 41602, 41603), 1e-5)):
        for tile in tiles:
            st, _, stats = run_plate_gpu(pdpkg.pdhip, w, 0.0, 40, precision=prec, tile=tile)
            _, nonbc, _ = H.parity_norms(st["v"], v, mask)
            assert nonbc < tol, (prec, tile, nonbc)
            g = st["f"].reshape(-1, 2)[base.n:]
            assert np.all(g[:5] == 0.0)                 # no family, or a singular one
            assert np.abs(g[5:]).max() > 0
            assert stats.n_singular >= 2                # the pair's two bonds


def test_batched_plates_fp64_match_oracle(pdpkg):
    say("batched plates (disjoint union), FP64, no breaking, vs the oracle")
    pdh = pdpkg.pdhip
    wb = pdpkg.plates.plates_batched(count=4, seed=1, lo=24, hi=40)
    st, brb, _ = run_plate_gpu(pdh, wb, 0.0, 40)
    u, v, f, _ = run_plate_oracle(wb, 0.0, 40)
    _, nonbc, _ = H.parity_norms(st["v"], v, wb.code >= 0)
    assert nonbc < 1e-11


def sub_plate(pl, wb, k):
    """Plate k of a disjoint-union mesh as a mesh of its own (its particles and
    bonds are contiguous; its two velocity groups become groups 0 and 1)."""
    idx = np.nonzero(wb.plate_of == k)[0]
    p0, p1 = int(idx[0]), int(idx[-1]) + 1
    b0, b1 = int(wb.row[p0]), int(wb.row[p1])
    code = wb.code[p0:p1].copy()
    code[code >= 0] -= 2 * k
    return pl.Workload(f"plate{k}", wb.X[2 * p0:2 * p1].copy(), wb.V[p0:p1].copy(),
                       (wb.row[p0:p1 + 1] - b0).astype(np.int32), (wb.col[b0:b1] - p0).astype(np.int32),
                       code, wb.v0[4 * k:4 * k + 4].copy(), wb.broken[b0:b1].copy(),
                       horizon=wb.horizon, dx=wb.dx), (p0, p1, b0, b1)


def test_batched_plates_breaking_match_oracle_and_single_runs(pdpkg):
    """The batched configuration's arithmetic at a size the oracle runs: six
    plates of 24..48 per side with their own notch, angle and edge velocity,
    critical stretch 0.008 (8,800 bonds break in four of the six plates within
    300 steps, none in the other two).  FP64 keeps the oracle's broken set with
    velocities within 1e-10; with the tile fixed, each plate of the batched FP64
    run is bit-identical to that plate run alone (the batch changes nothing but
    the launch).  Mixed precision, stepped in lockstep with the FP64 run, has
    the FP64 broken set after every step through step 200 and is within 1e-5
    of it there, and ends with the oracle's broken set.  (Later a few bonds
    whose stretch reaches s0 within the mixed-precision error of u break one
    step apart; see test_batched_plates_mixed_moderate_breaking for a run
    whose crack histories coincide step by step to the end.)"""
    say("batched plates with breaking: 6 plates, s0 = 0.008, 300 steps, FP64 and mixed vs the oracle")
    pdh = pdpkg.pdhip
    wb = pdpkg.plates.plates_batched(count=6, seed=2, lo=24, hi=48)
    s0, steps = 0.008, 300
    u, v, f, br = run_plate_oracle(wb, s0, steps)
    assert br.sum() > wb.broken.sum() + 5000
    mask = wb.code >= 0
    st, brg, _ = run_plate_gpu(pdh, wb, s0, steps, tile=1602)
    np.testing.assert_array_equal(brg, br)
    _, nonbc, _ = H.parity_norms(st["v"], v, mask)
    assert nonbc < 1e-10, nonbc
    for k in (2, 5):
        w1, (p0, p1, b0, b1) = sub_plate(pdpkg.plates, wb, k)
        s1, br1, _ = run_plate_gpu(pdh, w1, s0, steps, tile=1602)
        np.testing.assert_array_equal(s1["v"], st["v"][2 * p0:2 * p1])
        np.testing.assert_array_equal(s1["u"], st["u"][2 * p0:2 * p1])
        np.testing.assert_array_equal(br1, brg[b0:b1])
        assert br1.sum() > w1.broken.sum()
    for tile in (11602, 1603):
        c64 = wb.context(precision=0, s0=s0)
        c64.set_option(pdh.OPT_TILE, 1602)
        cm = wb.context(precision=1, s0=s0)
        cm.set_option(pdh.OPT_TILE, tile)
        differ = []
        for step in range(1, steps + 1):
            c64.step(1)
            cm.step(1)
            if not np.array_equal(c64.get_broken(), cm.get_broken()):
                differ.append(step)
            if step <= 200 and step % 50 == 0:
                assert not differ, (tile, differ)
                _, nonbc, _ = H.parity_norms(cm.get_state()["v"], c64.get_state()["v"], mask)
                assert nonbc < 1e-5, (tile, step, nonbc)
        bm = cm.get_broken()
        c64.close()
        cm.close()
        say(f"  mixed tile {tile}: steps with differing broken sets {differ[:10]}{'...' if len(differ) > 10 else ''}")
        np.testing.assert_array_equal(bm, br)


def test_batched_plates_mixed_moderate_breaking(pdpkg):
    """The same six plates at critical stretch 0.012 (several hundred bonds
    break in the faster plates): the mixed-precision crack history equals the
    FP64 one after every step, and the velocities stay within 1e-5 of the
    oracle's to the end of the 300 steps, in the autotuned and in fixed tiles."""
    say("batched plates, s0 = 0.012, 300 steps, mixed vs the oracle, step by step")
    pdh = pdpkg.pdhip
    wb = pdpkg.plates.plates_batched(count=6, seed=2, lo=24, hi=48)
    s0, steps = 0.012, 300
    u, v, f, br = run_plate_oracle(wb, s0, steps)
    assert br.sum() > wb.broken.sum() + 300
    mask = wb.code >= 0
    for tune, tile in ((2, None), (0, 11602), (0, 1603), (0, 3202)):
        c64 = wb.context(precision=0, s0=s0)
        c64.set_option(pdh.OPT_TILE, 1602)
        cm = wb.context(precision=1, s0=s0)
        if tune:
            cm.autotune(tune)
        else:
            cm.set_option(pdh.OPT_TILE, tile)
        for step in range(steps):
            c64.step(1)
            cm.step(1)
            np.testing.assert_array_equal(cm.get_broken(), c64.get_broken())
        sm = cm.get_state()
        bm = cm.get_broken()
        c64.close()
        cm.close()
        _, nonbc, _ = H.parity_norms(sm["v"], v, mask)
        say(f"  mixed tile {tile or 'autotuned'}: v error {nonbc:.3e}")
        assert nonbc < 1e-5, (tile, nonbc)
        np.testing.assert_array_equal(bm, br)


def test_non_finite_state_is_reported(pdpkg, small_plate):
    say("non-finite state -> PD_ERR_NOT_FINITE naming the particle")
    pdh = pdpkg.pdhip
    ctx = small_plate.context(s0=0.0)
    u = np.zeros(2 * small_plate.n)
    u[2 * 17 + 1] = np.nan
    ctx.set_state(u=u)
    with pytest.raises(pdh.PDError) as e:
        ctx.step(2)
    assert e.value.code == pdh.PD_ERR_NOT_FINITE
    # the named particle (the NaN has spread through the families by now) is bad
    bad = int(str(e.value).split("particle ")[1].split()[0])
    st = ctx.get_state()
    assert not all(np.isfinite(st[k][2 * bad:2 * bad + 2]).all() for k in ("u", "v", "f"))
    ctx.set_option(pdh.OPT_CHECK_FINITE, 0)
    ctx.step(1)  # unchecked: runs, the NaN spreads
    assert not np.isfinite(ctx.get_state()["v"]).all()
    ctx.close()


def test_collinear_families_are_singular_not_nan(pdpkg):
    """A single row of particles: every K_ik is rank one, so every bond is
    singular (K^-1 = 0), no stress, no NaN."""
    say("collinear families: singular K guard")
    pdh = pdpkg.pdhip
    n = 40
    X = np.zeros(2 * n)
    X[0::2] = np.arange(n) * 1e-3
    V = np.full(n, 1e-9)
    row, col = pdh.build_families(X, 1e-3, 3.015)
    ctx = pdh.Context(pdh.Params.make(s0=0.0), X, V, row, col)
    st = ctx.stats()
    assert st.n_singular == len(col)
    code = np.full(n, pdh.BC_FREE, np.int32)
    code[-1] = 0
    ctx.set_boundary(code, np.array([1.0, 0.0]))
    ctx.step(10)
    s = ctx.get_state()
    ctx.close()
    assert np.isfinite(s["v"]).all() and np.abs(s["f"]).max() == 0.0
