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


def test_create_rejects_duplicate_neighbours_without_device(pdpkg):
    """A family that lists a neighbour twice would map both copies to one
    reverse slot (one force term written twice, another never): rejected,
    naming the family, the neighbour and the caller bond, whether the
    duplicate is one-sided or mirrored in the neighbour's family."""
    pdh = pdpkg.pdhip
    X = _mesh()
    n = X.size // 2
    V = np.ones(n)
    row, col = pdh.build_families(X, 1.0, 3.015)
    p = pdh.Params.make(dx=1.0)
    i = 7
    j = int(col[row[i]])
    fams = [list(col[row[k]:row[k + 1]]) for k in range(n)]
    for mirrored in (False, True):
        f2 = [list(f) for f in fams]
        f2[i].append(j)
        if mirrored:
            f2[j].append(i)
        row2 = np.concatenate([[0], np.cumsum([len(f) for f in f2])]).astype(np.int32)
        col2 = np.concatenate(f2).astype(np.int32)
        with pytest.raises(pdh.PDError) as e:
            pdh.Context(p, X, V, row2, col2)
        assert e.value.code == pdh.PD_ERR_INVALID
        msg = str(e.value)
        assert "more than once" in msg and f"neighbour {j}" in msg, msg
        assert f"family {i}" in msg or f"family {j}" in msg, msg


def test_binding_checks_array_sizes(pdpkg):
    """The Python binding refuses arrays of the wrong length before the C
    library could read past their end."""
    import ctypes

    pdh = pdpkg.pdhip
    X = _mesh()
    n = X.size // 2
    V = np.ones(n)
    row, col = pdh.build_families(X, 1.0, 3.015)
    p = pdh.Params.make(dx=1.0)
    for args in ((X[:-2], V, row, col), (X, V, row[:-1], col), (X, V, row, col[:-1])):
        with pytest.raises(ValueError):
            pdh.Context(p, *args)
    with pytest.raises(ValueError):
        pdh.Context.from_points(p, X[:-1], V)
    with pytest.raises(ValueError):
        pdh.Context.from_padded(p, X, V, np.zeros(n * 30 - 1, np.int32), np.zeros(n, np.int32), 30)
    ctx = pdh.Context.__new__(pdh.Context)   # no device context: the checks come first
    ctx._h = ctypes.c_void_p()
    ctx.n, ctx.nb = n, len(col)
    with pytest.raises(ValueError):
        ctx.set_broken(np.zeros(len(col) - 1, np.uint8))
    with pytest.raises(ValueError):
        ctx.set_boundary(np.full(n - 1, -1, np.int32), np.zeros(2))
    with pytest.raises(ValueError):
        ctx.set_boundary(np.full(n, -1, np.int32), np.zeros(3))
    for fn in (ctx.set_youngs_modulus,):
        with pytest.raises(ValueError):
            fn(np.ones(n + 1))
    with pytest.raises(ValueError):
        ctx.set_state(u=np.zeros(2 * n - 1))
    with pytest.raises(ValueError):
        ctx.set_damage_state(bond_damage=np.zeros(len(col) + 1))
    with pytest.raises(ValueError):
        ctx.write_positions("/dev/null", ids=np.arange(n - 1))


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


def test_adapter_mesh_signature_sees_y_and_volume(tmp_path):
    """The adapter reuses its device context only for an identical mesh: the
    signature changes with a particle's y position or volume, not only with its
    families and x."""
    src = tmp_path / "sig.cpp"
    src.write_text(r'''
#include "pd_mechanics_adapter.hpp"
#include <cstdio>
#include <set>
struct Particle { int globalID; long double volume; std::vector<long double> initialPositions;
                  std::vector<int> neighbors; };
int main() {
    std::vector<Particle> ps(3);
    std::unordered_map<int, int> ids;
    for (int i = 0; i < 3; ++i) {
        ps[i].globalID = 10 + i; ps[i].volume = 1.0L; ps[i].initialPositions = {1.0L * i, 0.0L};
        ids[10 + i] = i;
    }
    ps[0].neighbors = {11}; ps[1].neighbors = {10, 12}; ps[2].neighbors = {11};
    std::set<uint64_t> seen;
    seen.insert(pd_adapter::mesh_signature(ps, ids));
    auto q = ps; q[1].initialPositions[1] = 1e-12L; seen.insert(pd_adapter::mesh_signature(q, ids));
    q = ps; q[2].volume = 1.0L + 1e-15L; seen.insert(pd_adapter::mesh_signature(q, ids));
    q = ps; q[0].initialPositions[0] = 1e-300L; seen.insert(pd_adapter::mesh_signature(q, ids));
    auto m = ids; m[12] = 0; seen.insert(pd_adapter::mesh_signature(ps, m));
    bool same = pd_adapter::mesh_signature(ps, ids) == pd_adapter::mesh_signature(ps, ids);
    std::printf("%zu %d\n", seen.size(), (int)same);
    return (seen.size() == 5 && same) ? 0 : 1;
}
''')
    exe = tmp_path / "sig"
    rc = os.system(f"g++ -std=c++17 -I{ROOT}/include -o {exe} {src}")
    assert rc == 0
    assert os.system(str(exe)) == 0
