"""Device-built set-up (pd_create_from_points): families, order and state equal
those of pd_create with the host families of pd_build_families (findNeighbor's
rule, src/SetupParticleSystem.cpp:100-136)."""
import numpy as np
import pytest

import hassa as H
from conftest import say

pytestmark = pytest.mark.gpu


def meshes(pdpkg):
    P = pdpkg.plates
    X, V = pdpkg.pdhip.read_mesh(H.MESH)
    yield "hassa", X, V, H.DX, H.HORIZON
    w = P.single_plate("j", 60, 45, 20.0, seed=2)
    yield "jittered", w.X, w.V, w.dx, w.horizon
    w = P.single_plate("d4", 40, 33, 20.0, horizon=4.015, seed=5, porosity=0.01)
    yield "porous-4.015", w.X, w.V, w.dx, w.horizon
    w = P.plates_batched(count=5, seed=3, lo=20, hi=40)
    yield "batched", w.X, w.V, w.dx, w.horizon


def test_device_families_equal_host_families(pdpkg):
    pdh = pdpkg.pdhip
    for name, X, V, dx, hor in meshes(pdpkg):
        say(f"device set-up families: {name}")
        row, col = pdh.build_families(X, dx, hor)
        p = pdh.Params.make(horizon=hor, dx=dx)
        ctx = pdh.Context.from_points(p, X, V)
        r2, c2 = ctx.families()
        st = ctx.stats()
        ctx.close()
        np.testing.assert_array_equal(r2, row)
        np.testing.assert_array_equal(c2, col)
        assert st.nb == len(col) and st.max_family == np.diff(row).max()


def test_device_setup_steps_bitwise_equal(pdpkg):
    say("device set-up vs host set-up: 50 steps, breaking")
    pdh = pdpkg.pdhip
    w = pdpkg.plates.single_plate("p48x40", 48, 40, 30.0, notch_len_cells=16, seed=3)
    out = []
    for how in ("host", "device"):
        p = w.params(s0=0.005)
        ctx = pdh.Context(p, w.X, w.V, w.row, w.col) if how == "host" else pdh.Context.from_points(p, w.X, w.V)
        ctx.set_boundary(w.code, w.v0)
        ctx.set_broken(w.broken)
        ctx.step(50)
        out.append((ctx.get_state(), ctx.get_broken(), ctx.get_bond_forces()))
        ctx.close()
    for k in ("u", "v", "f"):
        np.testing.assert_array_equal(out[0][0][k], out[1][0][k])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][2], out[1][2])
