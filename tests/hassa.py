"""The reference's Hassa test case, rebuilt from the fixture mesh.

Mesh reading follows readMeshFile (src/SetupParticleSystem.cpp:11-60): a header
line, then "Center_X Center_Y Area"; particles with |x| > 17.5 and |y| < 2 are
deleted to form two edge notches, leaving 4064 particles.  Boundary sets follow
defineBoundarySet (src/SetupParticleSystem.cpp:513-529) with the box limits set
at :53-56.  Driver parameters are those of src/main_seq.cpp:607-662.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "hassa")
MESH = os.path.join(GOLDEN, "2D_Hassa_test_mesh.dat")

BOX = (-34.5, 34.5, -29.5, 29.5)
DX = 1.0
HORIZON = 3.0
N1 = N2 = 10.0
E = 3.5e4
NU = 0.2
RHO = 2.5e-6
DT = 1e-9
V_TOP = (0.0, 5.0)


def read_mesh():
    """Return X (2n long double), V (n long double) in reference order."""
    xs, ys, vs = [], [], []
    with open(MESH) as fh:
        fh.readline()
        for line in fh:
            parts = line.split()
            if len(parts) < 3:
                continue
            x, y, a = (np.longdouble(t) for t in parts[:3])
            if abs(x) > 17.5 and abs(y) < 2.0:
                continue
            xs.append(x)
            ys.append(y)
            vs.append(a)
    X = np.empty(2 * len(xs), dtype=np.longdouble)
    X[0::2] = xs
    X[1::2] = ys
    return X, np.array(vs, dtype=np.longdouble)


def boundary_sets(X, row):
    """defineBoundarySet: only particles with |H| <= 17 qualify."""
    n = len(row) - 1
    sets = [[] for _ in range(4)]
    half = DX / 2.0
    for k in range(n):
        if row[k + 1] - row[k] > 17:
            continue
        for d in range(2):
            c = X[2 * k + d]
            if c < BOX[2 * d] + half:
                sets[2 * d].append(k)
            elif c > BOX[2 * d + 1] - half:
                sets[2 * d + 1].append(k)
    return [np.array(s, dtype=np.int32) for s in sets]


def load_velocity_fixture(name):
    return np.loadtxt(os.path.join(GOLDEN, name), dtype=np.float64)


def load_positions_fixture():
    """LocalParticlesPositions_rank_0step_0.txt: ID, x, y, ux, uy."""
    return np.loadtxt(os.path.join(GOLDEN, "LocalParticlesPositions_rank_0step_0.txt"),
                      delimiter=",", skiprows=1, dtype=np.float64)


def case():
    """Mesh, CSR families (reference neighbour order) and BC sets."""
    from oracle_lib import ref_find_neighbor
    X, V = read_mesh()
    row, col = ref_find_neighbor(X, np.longdouble(DX), np.longdouble(HORIZON))
    sets = boundary_sets(X, row)
    return X, V, row, col, sets


def parity_norms(v, v_ref, mask_bc):
    """Norm-wise parity metric: (i) all particles, (ii) non-prescribed only,
    (iii) worst per-element excess over rtol*|v_ref| + atol with
    atol = rtol * max_nonBC |v_ref| (returned as the ratio to that bound at
    rtol = 1)."""
    d = v - v_ref
    all_rel = np.linalg.norm(d) / np.linalg.norm(v_ref)
    nb = ~np.repeat(mask_bc, 2)
    nonbc_rel = np.linalg.norm(d[nb]) / np.linalg.norm(v_ref[nb])
    scale = np.max(np.abs(v_ref[nb]))
    elem = np.max(np.abs(d) / (np.abs(v_ref) + scale))
    return all_rel, nonbc_rel, elem
