#!/usr/bin/env python3
"""Throughput of the explicit bond-associated NOSB peridynamics step on one MI355X.

Workload (default): the 1024 x 1024 plain-FP64 plate (1,048,576 particles,
~29.3 M directed bonds, delta = 3.015 dx, E = 192 GPa, nu = 0.3, rho = 8000,
s0 = 0.02, dt = 1e-8 s, +-20 m/s on the top / bottom rows, edge notch 0.25 W).
A "step" is one full explicit step (velocity BC, position update, fixed BC,
bond breaking, K re-assembly where bonds broke, force state, pairwise force,
velocity update) over the whole plate, with every input resident in HBM.

value  = bond-evals/s = (unbroken directed bonds) x steps / time, FP64.
Also reported: particle-steps/s, the same plate in mixed precision (FP32 bond
arithmetic, FP64 K and state) with its per-field error against the FP64 run,
per-kernel times from HIP events recorded on the launch stream inside the timed
region, the roofline of the dominant (bond) kernel (HBM bytes, and beside it its
arithmetic: the (k, l) pairs' FP64 / FP32 operations), and the CPU baseline: the
literal single-threaded restatement of the reference's computeVelocity
(src/main_seq.cpp:396-470) timed on a bounded sample of the same plate.
"""
import argparse
import importlib
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def kernel_bytes(n, nb, mixed):
    """Algorithmic bytes per launch: the whole step moves 125 N + b Nb
    (b = 89 FP64 / 65 mixed); the bond kernel's share is X 16 + V 8 + row_ptr 4
    + u 16 per particle and col 4 + rev 4 + K^-1 32 + broken 1 + T write
    (16 FP64 / 8 FP32) per bond; the rest belongs to the gather/update kernels."""
    tb = 8 if mixed else 16
    bond = 44 * n + (41 + tb) * nb
    step = 125 * n + (89 - (24 if mixed else 0)) * nb
    return bond, step


def phase_bytes(n, nb, mixed):
    """Algorithmic bytes of each kernel phase as timed by the events (eager,
    unfused): the bond kernel as above; the gather reads row_ptr 4, f_own 16,
    a 16, v 16 and writes v 16, f 16 per particle and reads its W row (16 FP64 /
    8 FP32 per bond); the position update reads v, f, u 16 each and the BC code
    4, and writes a and u 16 each."""
    bond, _ = kernel_bytes(n, nb, mixed)
    return {"bond": bond, "gather": 84 * n + (8 if mixed else 16) * nb, "positions": 84 * n}


FP64_PEAK_TFLOPS = 78.6    # MI355X FP64, vector and matrix
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 (dense)
FLOPS_PER_PAIR = 28        # per (i, k, l): weight 12, D += w P 8, S += w Q 8


def compute_roofline(row, bond_ms, mixed):
    """The bond kernel's arithmetic side: sum_i |H_i|^2 (k, l) pairs per launch,
    FLOPS_PER_PAIR operations each (FMA = 2), over the kernel time, against the
    precision's peak."""
    pairs = float(np.sum(np.diff(row).astype(np.float64) ** 2))
    ach = pairs * FLOPS_PER_PAIR / (bond_ms * 1e-3) / 1e12
    peak = FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS
    return {"bound": "valu", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
            "pairs": pairs, "flops_per_pair": FLOPS_PER_PAIR}


def phase_fracs(pb, ms):
    return {k: pb[k] / (ms[k] * 1e-3) / (HBM_PEAK_GBS * 1e9) if ms[k] > 0 else None for k in pb}


def cpu_baseline(pkg, w, state, budget_s=12.0):
    """Literal long-double restatement of computeVelocity, 1 thread, on the
    first families of the plate (in caller order) until ~budget_s of work."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O  # noqa: E402  (checker / CPU baseline only)

    p = O.RefParams(10.0, 10.0, w.horizon, w.dx, 192e9, 0.3, 8000.0, 1e-8)
    X = w.X.astype(np.longdouble)
    y = (w.X.astype(np.longdouble) + state["u"].astype(np.longdouble))
    V = w.V.astype(np.longdouble)
    v = state["v"].astype(np.longdouble)
    a = state["a"].astype(np.longdouble)
    f = np.zeros(2 * w.n, np.longdouble)
    # calibrate on a few families, then size the sample to the budget
    i0 = w.n // 2
    t = time.perf_counter()
    O.ref_compute_velocity(p, X, y, V, w.row, w.col, v, a, f, i0, i0 + 8, nthreads=1)
    per = (time.perf_counter() - t) / 8
    count = int(max(8, min(w.n - i0, budget_s / max(per, 1e-9))))
    t = time.perf_counter()
    O.ref_compute_velocity(p, X, y, V, w.row, w.col, v, a, f, i0, i0 + count, nthreads=1)
    dt = time.perf_counter() - t
    bonds = int(w.row[i0 + count] - w.row[i0])
    # the FP64 O(N M^2) oracle on all host cores, as a separate "strong CPU" line
    po = O.or_params(horizon=w.horizon, dx=w.dx, rho=8000.0, dt=1e-8, E=192e9, nu=0.3, s0=0.0)
    cores = os.cpu_count() or 1
    t = time.perf_counter()
    O.forces(po, w.X, w.V, w.row, w.col, state["u"], nthreads=cores)
    dt_o = time.perf_counter() - t
    return {
        "value": bonds / dt, "unit": "bond-evals/s", "cores": 1, "kind": "port",
        "sample": f"computeVelocity (literal O(M^3) long-double restatement of src/main_seq.cpp:396-470) "
                  f"for {count} families ({bonds} bonds) of {w.name}, {dt:.1f} s on 1 thread",
        "particle_steps_per_s": count / dt,
        "cpu_model": platform.processor() or platform.machine(),
        "strong_cpu_oracle_fp64": {"value": w.nb / dt_o, "unit": "bond-evals/s", "cores": cores,
                                   "sample": f"one force evaluation of all {w.nb} bonds of {w.name}, "
                                             f"O(N M^2) FP64 oracle, OpenMP {cores} threads"},
    }


def tile_info(t):
    return {"lanes_per_family": t[0], "bonds_per_lane": t[1], "matrix_cores": bool(t[2]),
            "waves_per_wg": t[3], "persistent_wg_per_cu": t[4]}


def read_traffic(workload, prec):
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}_{prec}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("bytes_per_launch")


def run(pkg, w, precision, steps, warmup, tune, tile=None, wpb=4):
    pdh = pkg.pdhip
    ctx = w.context(precision=precision)
    if tile is not None:
        ctx.set_option(pdh.OPT_TILE, tile)
        ctx.set_option(pdh.OPT_WAVES_PER_WG, wpb)
    elif tune:
        # pick the family tile / workgroup shape on this plate (state restored)
        t0 = time.perf_counter()
        ctx.autotune(tune)
        print(f"# autotune {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    ctx.step(warmup)
    ctx.synchronize()
    # timed region 1: graph replay (the production path)
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.synchronize()
    t_graph = time.perf_counter() - t0
    # timed region 2: eager launches with per-kernel HIP events on the launch stream
    ctx.set_option(pdh.OPT_PROFILE, 1)
    ctx.reset_profile()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.synchronize()
    t_prof = time.perf_counter() - t0
    st = ctx.stats()
    state = ctx.get_state()
    ctx.close()
    return t_graph, t_prof, st, state


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="plate_1024")
    ap.add_argument("--tune", type=int, default=3, help="autotuner steps per variant (0: off)")
    ap.add_argument("--tile", type=int, default=None,
                    help="pin the family tile (PD_OPT_TILE value) instead of autotuning")
    ap.add_argument("--wpb", type=int, default=4, help="waves per workgroup with --tile")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--only", choices=("fp64", "mixed"), default=None,
                    help="run one precision only (profiling runs); prints no JSON line")
    ap.add_argument("--probe-copy", type=int, default=1 << 30,
                    help="bytes of the plain copy probe (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the step would shard by space across GPUs "
              f"(Morton slabs with a 2-horizon halo), but no multi-GPU path is built; "
              f"got --gpus {args.gpus}", file=sys.stderr)
        sys.exit(2)
    pkg = importlib.import_module("2d_pd_gpu_amd")
    t = time.perf_counter()
    w = pkg.plates.WORKLOADS[args.workload]()
    print(f"# workload {w.name}: n={w.n} nb={w.nb} pre-broken={int(w.broken.sum())} "
          f"setup {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)

    copy_gbs = None
    if args.probe_copy:
        copy_gbs = pkg.pdhip.probe_copy(args.probe_copy, 20)
        print(f"# copy probe {args.probe_copy} B: {copy_gbs:.1f} GB/s", file=sys.stderr, flush=True)

    res = {}
    for name, prec in (("fp64", pkg.FP64), ("mixed", pkg.MIXED)):
        if name == "mixed" and args.no_mixed:
            continue
        if args.only and name != args.only:
            continue
        t_graph, t_prof, st, state = run(pkg, w, prec, args.steps, args.warmup, args.tune, args.tile,
                                         args.wpb)
        nb_alive = w.nb - st.n_broken
        calls = max(1, st.phase_calls[1])
        bond_ms = st.phase_ms[1] / calls
        gather_ms = st.phase_ms[2] / max(1, st.phase_calls[2])
        pos_ms = st.phase_ms[0] / max(1, st.phase_calls[0])
        kb, sb = kernel_bytes(w.n, w.nb, prec == pkg.MIXED)
        res[name] = dict(t=t_graph, t_prof=t_prof, nb_alive=nb_alive, bond_ms=bond_ms,
                         gather_ms=gather_ms, pos_ms=pos_ms, kb=kb, sb=sb, state=state,
                         n_broken=st.n_broken, guarded=st.n_guarded_families,
                         tile=(st.lanes_per_family, st.bonds_per_lane, st.matrix_cores,
                               st.waves_per_wg, st.persistent))
        print(f"# {name}: {args.steps} steps graph {t_graph:.3f}s eager+events {t_prof:.3f}s "
              f"bond {bond_ms:.4f} ms gather {gather_ms:.4f} ms positions {pos_ms:.4f} ms "
              f"broken {st.n_broken}", file=sys.stderr, flush=True)

    if args.only:
        return
    r = res["fp64"]
    ms_per_step = 1e3 * r["t"] / args.steps
    value = r["nb_alive"] * args.steps / r["t"]
    achieved = r["kb"] / (r["bond_ms"] * 1e-3) / 1e9
    out = {
        "metric": "bond-evals/s (FP64, 1 MI355X)",
        "value": value,
        "unit": "bond-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic jittered plate (counter-based Philox, seed 0)",
        "config": {"workload": w.name, "particles": w.n, "bonds": w.nb, "horizon_dx": w.horizon,
                   "family_max": int(np.diff(w.row).max()), "steps": args.steps},
        "particle_steps_per_s": w.n * args.steps / r["t"],
        "step_hbm_frac": r["sb"] / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "copy_probe_gbs": copy_gbs,
        "kernel_ms": {"bond": r["bond_ms"], "gather": r["gather_ms"], "positions": r["pos_ms"]},
        "phase_hbm_frac": phase_fracs(phase_bytes(w.n, w.nb, False),
                                      {"bond": r["bond_ms"], "gather": r["gather_ms"], "positions": r["pos_ms"]}),
        "tile": tile_info(r["tile"]),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "kernel": "k_bond<double>",
                     "traffic": read_traffic(w.name, "fp64")},
        "compute_roofline": compute_roofline(w.row, r["bond_ms"], False),
    }
    if "mixed" in res:
        m = res["mixed"]
        ms_m = 1e3 * m["t"] / args.steps
        err = {}
        free = np.repeat(w.code < 0, 2)
        for k in ("u", "v", "f"):
            a_, b_ = m["state"][k][free], r["state"][k][free]
            err[k] = float(np.linalg.norm(a_ - b_) / max(np.linalg.norm(b_), 1e-300))
        ach_m = m["kb"] / (m["bond_ms"] * 1e-3) / 1e9
        out["mixed"] = {
            "value": m["nb_alive"] * args.steps / m["t"], "unit": "bond-evals/s", "dtype": "f32/f64",
            "ms_per_step": ms_m, "speedup_vs_fp64": (m["nb_alive"] / m["t"]) / (r["nb_alive"] / r["t"]),
            "particle_steps_per_s": w.n * args.steps / m["t"],
            "step_hbm_frac": m["sb"] / (ms_m * 1e-3) / (HBM_PEAK_GBS * 1e9),
            "kernel_ms": {"bond": m["bond_ms"], "gather": m["gather_ms"], "positions": m["pos_ms"]},
            "phase_hbm_frac": phase_fracs(phase_bytes(w.n, w.nb, True),
                                          {"bond": m["bond_ms"], "gather": m["gather_ms"], "positions": m["pos_ms"]}),
            "tile": tile_info(m["tile"]),
            "roofline": {"bound": "hbm", "achieved": ach_m, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach_m / HBM_PEAK_GBS, "kernel": "k_bond<float>",
                         "traffic": read_traffic(w.name, "mixed")},
            "compute_roofline": compute_roofline(w.row, m["bond_ms"], True),
            "rel_err_vs_fp64": err,
            "broken_equal_count": int(m["n_broken"]) == int(r["n_broken"]),
            "guarded_families": int(m["guarded"]),
        }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pkg, w, r["state"])
        out["vs_cpu_port"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
