"""Result files in the reference's text formats, written from the GPU state,
against the files the reference's own run wrote (tests/golden/hassa)."""
import os

import numpy as np
import pytest

import hassa as H
from conftest import say

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hassa")


def hassa_step_context(pdh):
    X, V = pdh.read_mesh(H.MESH)
    row, col = pdh.build_families(X, H.DX, H.HORIZON)
    p = pdh.Params.make(precision=pdh.FP64, n1=H.N1, n2=H.N2, horizon=H.HORIZON, dx=H.DX, rho=H.RHO,
                        dt=H.DT, E=H.E, nu=H.NU, s0=0.0)
    ctx = pdh.Context(p, X, V, row, col)
    sets = H.boundary_sets(X, row)
    # src/main_seq.cpp:686-699
    ctx.apply_velocity_bc(sets[3], H.V_TOP)
    ctx.update_positions()
    ctx.apply_fixed_bc(sets[2])
    ctx.compute_velocity()
    return ctx


def test_positions_file_is_byte_identical(pdpkg, tmp_path):
    """outputLocalParticlesPositions after step 0 (the reference's
    LocalParticlesPositions_rank_0step_0.txt), from mesh file to text."""
    say("positions writer vs LocalParticlesPositions_rank_0step_0.txt")
    ctx = hassa_step_context(pdpkg.pdhip)
    out = tmp_path / "pos.txt"
    ctx.write_positions(str(out))
    ctx.close()
    with open(os.path.join(GOLD, "LocalParticlesPositions_rank_0step_0.txt"), "rb") as fh:
        ref = fh.read()
    assert out.read_bytes() == ref


def test_velocity_file_matches_seq_fixture(pdpkg, tmp_path):
    """storeVelocity after step 0 against velocity_step_0_after_seq_ref.txt.
    Where the velocity is physical (the loaded rows and the wave front below
    them) the text is the same, except where the 7th digit of the FP64 and
    long-double values straddles a rounding boundary.  Elsewhere the reference
    prints the round-off of its position-form strain (~1e-19); the
    displacement form gives exact zeros or round-off of the same size."""
    say("velocity writer vs velocity_step_0_after_seq_ref.txt")
    ctx = hassa_step_context(pdpkg.pdhip)
    out = tmp_path / "v.txt"
    ctx.write_velocity(str(out))
    out17 = tmp_path / "v17.txt"
    ctx.write_velocity(str(out17), digits=17)
    st = ctx.get_state()
    ctx.close()
    mine = out.read_text().splitlines()
    with open(os.path.join(GOLD, "velocity_step_0_after_seq_ref.txt")) as fh:
        ref = fh.read().splitlines()
    assert len(mine) == len(ref)
    ta = np.array([ln.split() for ln in mine])
    tb = np.array([ln.split() for ln in ref])
    a, b = ta.astype(float), tb.astype(float)
    significant = np.abs(b) > 1e-12 * np.abs(b).max()
    assert 100 < significant.sum() < b.size // 2
    differ = np.count_nonzero(ta[significant] != tb[significant])
    assert differ <= significant.sum() // 50
    np.testing.assert_allclose(a[significant], b[significant], rtol=1.0001e-5, atol=0)
    assert np.abs(a[~significant]).max() <= 1e-12 * np.abs(b).max()
    np.testing.assert_array_equal(np.loadtxt(out17).reshape(-1), st["v"])


def test_gathered_file_with_damage(pdpkg, tmp_path):
    say("gathered positions writer with damageStatus")
    pdh = pdpkg.pdhip
    ctx = hassa_step_context(pdh)
    ctx.set_damage(pdh.DamageParams.make())
    u = ctx.get_state()["u"]
    ctx.set_state(u=u * 2e4)  # strains past the threshold
    ctx.compute_damage()
    status = ctx.get_damage()["status"]
    out = tmp_path / "g.txt"
    ctx.write_gathered(str(out), digits=17)
    X, _ = pdh.read_mesh(H.MESH)
    ctx.close()
    lines = out.read_text().splitlines()
    assert lines[0] == "ID, x, y, ux, uy, damage"
    tab = np.array([list(map(float, s.split(","))) for s in lines[1:]])
    np.testing.assert_array_equal(tab[:, 0], np.arange(len(lines) - 1))
    np.testing.assert_array_equal(tab[:, 1:3].reshape(-1), X)
    np.testing.assert_array_equal(tab[:, 3:5].reshape(-1), u * 2e4)
    np.testing.assert_array_equal(tab[:, 5], status)
    assert status.max() > 0


def test_checkpoint_resume_with_guarded_families_is_bitwise(pdpkg, tmp_path):
    """Mixed precision on the fragmenting delta = 4.4 dx plate: checkpoint at
    step 100 (families already guarded), 100 more steps, against a fresh
    context that loads it (its guard flags rebuilt from the broken set) and
    runs the same 100: identical state and broken set."""
    say("checkpoint / resume with guarded families")
    w = pdpkg.plates.single_plate("h44", 22, 20, 30.0, horizon=4.4, seed=6, notch_len_cells=8)
    ck = tmp_path / "guard.ckpt"
    ctx = w.context(precision=1, s0=0.004)
    ctx.step(100)
    assert ctx.stats().n_guarded_families > 0
    ctx.save_checkpoint(str(ck))
    ctx.step(100)
    a_state, a_br, a_ng = ctx.get_state(), ctx.get_broken(), ctx.stats().n_guarded_families
    ctx.close()
    ctx = w.context(precision=1, s0=0.004)
    ctx.load_checkpoint(str(ck))
    assert ctx.stats().n_guarded_families > 0
    ctx.step(100)
    b_state, b_br, b_ng = ctx.get_state(), ctx.get_broken(), ctx.stats().n_guarded_families
    ctx.close()
    for k in ("u", "v", "a", "f"):
        np.testing.assert_array_equal(a_state[k], b_state[k])
    np.testing.assert_array_equal(a_br, b_br)
    assert a_ng == b_ng


@pytest.mark.parametrize("precision", [0, 1])
def test_checkpoint_resume_is_bitwise(pdpkg, tmp_path, precision):
    """30 steps, checkpoint, 30 more, against a fresh context that loads the
    checkpoint and runs the last 30: identical state, broken set and damage."""
    say(f"checkpoint / resume, precision {precision}")
    pdh = pdpkg.pdhip
    w = pdpkg.plates.single_plate("p48x40", 48, 40, 30.0, notch_len_cells=16, seed=3)
    ck = tmp_path / "state.ckpt"
    ctx = w.context(precision=precision, s0=0.004)
    ctx.set_damage(pdh.DamageParams.make(mode=pdh.DAMAGE_MAZARS, kappa0=1e-4))
    ctx.step(30)
    ctx.save_checkpoint(str(ck))
    ctx.step(30)
    a_state, a_br, a_dmg, a_steps = ctx.get_state(), ctx.get_broken(), ctx.get_damage(), ctx.stats().steps
    ctx.close()
    ctx = w.context(precision=precision, s0=0.004)
    ctx.load_checkpoint(str(ck))
    assert ctx.stats().steps == 30
    ctx.step(30)
    b_state, b_br, b_dmg = ctx.get_state(), ctx.get_broken(), ctx.get_damage()
    assert ctx.stats().steps == a_steps
    ctx.close()
    assert a_br.sum() > w.broken.sum()
    for k in ("u", "v", "a", "f"):
        np.testing.assert_array_equal(a_state[k], b_state[k])
    np.testing.assert_array_equal(a_br, b_br)
    for k in ("bond", "kappa", "status"):
        np.testing.assert_array_equal(a_dmg[k], b_dmg[k])
    # a checkpoint of another mesh is refused
    other = pdpkg.plates.single_plate("o", 40, 40, 30.0, seed=4)
    ctx = other.context(precision=precision)
    with pytest.raises(pdh.PDError):
        ctx.load_checkpoint(str(ck))
    ctx.close()
