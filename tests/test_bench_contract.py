"""bench.py's contract, on the CPU: the multi-GPU refusal, and its byte and
operation counts against the per-step budgets stated for the workloads:
B_step = 125 N + 89 Nb (FP64) / 125 N + 65 Nb (mixed), i.e. 2.738 GB at
1024 x 1024 and 8.148 GB at 2048 x 2048 mixed, and 7,503,364 (k, l) pairs
per step on the 100 x 100 plate without jitter."""
import os
import subprocess
import sys

import numpy as np

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_reason():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""


def test_step_bytes_match_the_stated_budgets():
    # 1024 x 1024, delta = 3.015: N = 1,048,576, Nb = 29,286,436 (unjittered counts)
    _, step = bench.kernel_bytes(1048576, 29286436, mixed=False)
    assert abs(step / 1e9 - 2.738) < 0.0005
    # 2048 x 2048 mixed: N = 4,194,304, Nb = 117,293,092
    _, step = bench.kernel_bytes(4194304, 117293092, mixed=True)
    assert abs(step / 1e9 - 8.148) < 0.0005


def test_compute_roofline_counts_pairs_of_each_family():
    # families of 3 and 5 bonds: 9 + 25 pairs, 28 operations each
    row = np.array([0, 3, 8])
    r = bench.compute_roofline(row, bond_ms=1.0, mixed=False)
    assert r["pairs"] == 34 and r["flops_per_pair"] == 28
    assert abs(r["achieved"] - 34 * 28 / 1e-3 / 1e12) < 1e-18
    assert r["peak"] == 78.6 and bench.compute_roofline(row, 1.0, True)["peak"] == 157.3


def test_pairs_of_the_100_plate(pdpkg):
    X, _ = pdpkg.plates.plate_points(100, 100, dx=1e-3, jitter=0.0)
    row, col = pdpkg.pdhip.build_families(X, 1e-3, 3.015)
    assert len(col) == 272836
    assert bench.compute_roofline(row, 1.0, False)["pairs"] == 7503364


def test_cpu_window_sample_is_a_symmetric_sub_mesh(pdpkg):
    """The strong-CPU baseline's sample: the bonds between the particles of a
    central window, so families are symmetric and those of particles further
    than a horizon from the window's edge are the full plate's families."""
    import oracle_lib as O

    w = pdpkg.plates.plate_100()
    Xw, Vw, roww, colw, ids = bench.window(w, 40)
    assert len(Vw) == 1600 and np.all(np.diff(ids) > 0)
    O.build_rev(roww, colw)  # raises on an asymmetric family
    X = w.X.reshape(-1, 2)
    c = np.median(X, axis=0)
    deep = np.nonzero(np.all(np.abs(X[ids] - c) < (20 - 4) * w.dx, axis=1))[0]
    assert len(deep) > 500
    for q in deep:
        i = ids[q]
        np.testing.assert_array_equal(ids[colw[roww[q]:roww[q + 1]]], w.col[w.row[i]:w.row[i + 1]])
    model, ncpu, usable = bench.host_cpu()
    assert model and ncpu >= 1 and 1 <= usable <= ncpu
