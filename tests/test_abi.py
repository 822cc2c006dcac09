"""CPU-side checks of the C ABI: the library loads, exports every entry point
include/pd_hip.h declares, rejects bad input before touching a device, and its
host-side setup helpers (families, notch mask) agree with the reference
semantics they restate."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import hassa as H
import oracle_lib as O
from conftest import ROOT


def declared_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if not name.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", name)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        syms |= set(re.findall(r"\b(pd_[a-z0-9_]+)\s*\(", text))
    return syms


def test_library_exports_every_declared_symbol(pdpkg):
    L = pdpkg.pdhip.lib()
    declared = declared_symbols()
    assert {"pd_create", "pd_compute_velocity", "pd_step", "pd_update_positions"} <= declared
    for s in declared:
        assert hasattr(L, s), s
    assert set(pdpkg.pdhip.EXPORTS) <= declared
    assert b"gfx950" in L.pd_version()


def test_adapter_header_compiles(tmp_path):
    """include/pd_mechanics_adapter.hpp keeps the reference signatures; it must
    compile as ordinary C++ against the C ABI header."""
    src = tmp_path / "a.cpp"
    src.write_text('#include "pd_mechanics_adapter.hpp"\nint main(){return 0;}\n')
    rc = os.system(f"g++ -std=c++17 -fsyntax-only -I{ROOT}/include {src}")
    assert rc == 0


def _mesh(nx=6, ny=5):
    iy, ix = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    X = np.empty(2 * nx * ny)
    X[0::2] = ix.ravel()
    X[1::2] = iy.ravel()
    return X


def test_create_rejects_bad_input_without_device(pdpkg):
    pdh = pdpkg.pdhip
    X = _mesh()
    n = X.size // 2
    V = np.ones(n)
    row, col = pdh.build_families(X, 1.0, 3.015)
    p = pdh.Params.make(dx=1.0)
    p.ndim = 3
    with pytest.raises(pdh.PDError) as e:
        pdh.Context(p, X, V, row, col)
    assert e.value.code == pdh.PD_ERR_UNSUPPORTED
    p = pdh.Params.make(dx=1.0)
    bad = col.copy()
    bad[0] = n + 5
    with pytest.raises(pdh.PDError) as e:
        pdh.Context(p, X, V, row, bad)
    assert e.value.code == pdh.PD_ERR_INVALID
    # asymmetric families: drop one direction of one bond
    row2 = row.copy()
    col2 = np.delete(col, 0)
    row2[1:] -= 1
    with pytest.raises(pdh.PDError) as e:
        pdh.Context(p, X, V, row2, col2)
    assert e.value.code == pdh.PD_ERR_INVALID and "symmetric" in str(e.value)
    # family larger than 64 bonds
    Xb = _mesh(12, 12)
    rowb, colb = pdh.build_families(Xb, 1.0, 4.6)
    assert np.diff(rowb).max() > 64
    with pytest.raises(pdh.PDError) as e:
        pdh.Context(pdh.Params.make(dx=1.0, horizon=4.6), Xb, np.ones(Xb.size // 2), rowb, colb)
    assert e.value.code == pdh.PD_ERR_UNSUPPORTED
    p = pdh.Params.make(dx=1.0)
    p.dt = -1.0
    with pytest.raises(pdh.PDError):
        pdh.Context(p, X, V, row, col)


def test_build_families_matches_reference_find_neighbor(pdpkg):
    X, V, row_ref, col_ref, sets = H.case()
    row, col = pdpkg.build_families(X.astype(np.float64), 1.0, 3.0)
    np.testing.assert_array_equal(row, row_ref)
    np.testing.assert_array_equal(col, col_ref)


def test_build_families_jittered_plate(pdpkg):
    Xp, _ = pdpkg.plates.plate_points(40, 30, dx=1e-3, seed=5)
    row, col = pdpkg.build_families(Xp, 1e-3, 3.015)
    row_ref, col_ref = O.ref_find_neighbor(Xp.astype(np.longdouble), np.longdouble(1e-3),
                                           np.longdouble(3.015))
    np.testing.assert_array_equal(row, row_ref)
    np.testing.assert_array_equal(col, col_ref)
    m = np.diff(row)
    assert m.max() == 28  # the 1.01 factor keeps jittered interior families at 28


def _segments_cross(p, q, a, b):
    def orient(o, s, t):
        return (s[0] - o[0]) * (t[1] - o[1]) - (s[1] - o[1]) * (t[0] - o[0])
    d1, d2 = orient(a, b, p), orient(a, b, q)
    d3, d4 = orient(p, q, a), orient(p, q, b)
    return d1 * d2 < 0 and d3 * d4 < 0


def test_notch_mask_brute_force(pdpkg):
    w = pdpkg.plates.plate_100()
    X = w.X.reshape(-1, 2)
    seg = w.info["notches"][0]
    a, b = seg[:2], seg[2:]
    expect = np.zeros(len(w.col), np.uint8)
    for i in range(w.n):
        for e in range(w.row[i], w.row[i + 1]):
            j = w.col[e]
            if _segments_cross(X[i], X[j], a, b):
                expect[e] = 1
    np.testing.assert_array_equal(w.broken, expect)
    rev = O.build_rev(w.row, w.col)
    np.testing.assert_array_equal(w.broken, w.broken[rev])
    assert 600 < w.broken.sum() < 800


def test_workload_counts_match_unjittered_counts(pdpkg):
    """Nb of the 100 x 100 plate equals the exact count of the unjittered lattice
    (272,836): the 1.01 cutoff factor absorbs the +-0.01 dx jitter."""
    w = pdpkg.plates.plate_100()
    assert w.n == 10000 and w.nb == 272836


def test_read_mesh_matches_reference_reader(pdpkg, tmp_path):
    """pd_read_mesh keeps readMeshFile's rules (src/SetupParticleSystem.cpp:11-60):
    header skipped, notch particles |x| > 17.5 and |y| < 2 dropped."""
    pdh = pdpkg.pdhip
    X, V = pdh.read_mesh(H.MESH)
    Xr, Vr = H.read_mesh()
    assert V.size == 4064
    np.testing.assert_array_equal(X, Xr.astype(np.float64))
    np.testing.assert_array_equal(V, Vr.astype(np.float64))
    Xa, Va = pdh.read_mesh(H.MESH, reference_notch=False)
    with open(H.MESH) as fh:
        assert Va.size == sum(1 for line in fh.readlines()[1:] if len(line.split()) >= 3)
    bad = tmp_path / "missing.dat"
    with pytest.raises(pdh.PDError):
        pdh.read_mesh(str(bad))
