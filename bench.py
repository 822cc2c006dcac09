#!/usr/bin/env python3
"""Throughput of the explicit bond-associated NOSB peridynamics step on one MI355X.

Workload (default): the large single-domain plate, 4096 x 4096 particles on a
dx = 1 mm grid (jittered), horizon delta = 4.015 dx (families of up to 48
bonds), three randomly placed notches, 1 % of the particles removed
(16.6 M particles, 789 M directed bonds); E = 192 GPa, nu = 0.3, rho = 8000,
s0 = 0.02, dt = 1e-8 s, +-20 m/s on the top / bottom rows.  A "step" is one
full explicit step (velocity BC, position update, fixed BC, bond breaking, K
re-assembly where bonds broke, force state, pairwise force, velocity update)
over the whole plate, with every input resident in HBM.

value  = bond-evals/s = (unbroken directed bonds) x steps / time, FP64.
Also reported: particle-steps/s; the same plate in mixed precision (FP32 bond
arithmetic, FP64 K and state) with its per-field error against the FP64 run
(u, v, f, and the force state T of every bond) and whether the two broken-bond
sets are identical; per-kernel times from HIP events recorded on the launch
stream inside the timed region and each phase's fraction of the HBM peak; the
roofline of the dominant (bond) kernel (algorithmic HBM bytes, and beside it
its arithmetic: the (k, l) pairs' operations); the 1024 x 1024 plate and the
96-plate batch in both precisions as extra lines; and the CPU baseline: the
literal single-threaded restatement of the reference's computeVelocity
(src/main_seq.cpp:396-470) timed on a bounded sample of the same plate, with
the restructured FP64 oracle on all host cores on a window of it beside it.
"""
import argparse
import hashlib
import importlib
import json
import os
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


def host_cpu():
    """CPU model (from /proc/cpuinfo), logical CPUs and the CPUs this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.lower().startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model or "unknown", os.cpu_count() or 1, usable


def window(w, side):
    """The particles of a side x side (in dx) window at the plate's centre, with
    the bonds between them: a sample of the same plate for the CPU oracle
    (families cut at the window's edge)."""
    X = w.X.reshape(-1, 2)
    c = np.median(X, axis=0)
    inside = np.all(np.abs(X - c) < 0.5 * side * w.dx, axis=1)
    new = np.full(w.n, -1, np.int64)
    new[inside] = np.arange(int(inside.sum()))
    owner = np.repeat(np.arange(w.n, dtype=np.int32), np.diff(w.row))
    keep = inside[owner] & inside[w.col]
    cnt = np.bincount(new[owner[keep]], minlength=int(inside.sum()))
    row = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    col = new[w.col[keep]].astype(np.int32)
    return np.ascontiguousarray(X[inside].reshape(-1)), w.V[inside].copy(), row, col, np.nonzero(inside)[0]


def cpu_baseline(w, state, budget_s=12.0, window_side=512):
    """Literal long-double restatement of computeVelocity, 1 thread, on
    consecutive families from the middle of the plate (caller order) until
    ~budget_s of work; beside it the O(N M^2) FP64 oracle on all host cores,
    one force evaluation of a window of the plate."""
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
    model, ncpu, usable = host_cpu()
    # the FP64 O(N M^2) oracle on the host cores, as a separate "strong CPU" line
    Xw, Vw, roww, colw, ids = window(w, window_side)
    uw = np.ascontiguousarray(state["u"].reshape(-1, 2)[ids].reshape(-1))
    po = O.or_params(horizon=w.horizon, dx=w.dx, rho=8000.0, dt=1e-8, E=192e9, nu=0.3, s0=0.0)
    cores = min(usable, 16)
    O.forces(po, Xw, Vw, roww, colw, uw, nthreads=cores)  # warm-up
    t = time.perf_counter()
    O.forces(po, Xw, Vw, roww, colw, uw, nthreads=cores)
    dt_o = time.perf_counter() - t
    return {
        "value": bonds / dt, "unit": "bond-evals/s", "cores": 1, "kind": "port",
        "sample": f"computeVelocity (literal O(M^3) long-double restatement of src/main_seq.cpp:396-470) "
                  f"for {count} consecutive families ({bonds} bonds) from the middle of {w.name}, "
                  f"{dt:.1f} s on 1 thread",
        "particle_steps_per_s": count / dt,
        "cpu_model": model, "nproc": ncpu, "usable_cpus": usable,
        "strong_cpu_oracle_fp64": {
            "value": len(colw) / dt_o, "unit": "bond-evals/s", "cores": cores,
            "particle_steps_per_s": len(Vw) / dt_o,
            "sample": f"one force evaluation of the central {window_side} x {window_side} window of {w.name} "
                      f"({len(Vw)} particles, {len(colw)} bonds, families cut at its edge), "
                      f"O(N M^2) FP64 oracle, OpenMP {cores} threads, {dt_o:.2f} s"},
    }


def tile_code(t):
    return 10000 * int(t[2]) + 100 * int(t[0]) + int(t[1])


def tile_info(t):
    return {"lanes_per_family": t[0], "bonds_per_lane": t[1], "matrix_cores": bool(t[2]),
            "waves_per_wg": t[3], "persistent_wg_per_cu": t[4], "code": tile_code(t)}


def read_traffic(workload, prec, tile):
    """HBM bytes per bond-kernel launch from the rocprofv3 counter runs kept in
    profiles/ (FETCH_SIZE and WRITE_SIZE, calibrated on the copy probe), when
    they were taken on the same tile; otherwise None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}_{prec}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    if d.get("tile") not in (None, tile_code(tile)):
        return None
    return d.get("bytes_per_launch")


def mask_digest(m):
    return hashlib.sha1(np.ascontiguousarray(m, dtype=np.uint8).tobytes()).hexdigest()[:16]


def run(pkg, w, precision, steps, warmup, tune, tile=None, wpb=4, want_T=False):
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
    broken = ctx.get_broken()
    T = ctx.get_bond_forces() if want_T else None
    ctx.close()
    return dict(t=t_graph, t_prof=t_prof, st=st, state=state, broken=broken, T=T)


def measure(pkg, w, prec, args, want_T=False, tune=None, steps=None, warmup=None, tile=None):
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    r = run(pkg, w, prec, steps, warmup, args.tune if tune is None else tune, tile, args.wpb, want_T)
    st = r["st"]
    calls = max(1, st.phase_calls[1])
    r.update(nb_alive=w.nb - st.n_broken,
             bond_ms=st.phase_ms[1] / calls,
             gather_ms=st.phase_ms[2] / max(1, st.phase_calls[2]),
             pos_ms=st.phase_ms[0] / max(1, st.phase_calls[0]),
             tile=(st.lanes_per_family, st.bonds_per_lane, st.matrix_cores, st.waves_per_wg, st.persistent),
             steps=steps)
    r["kb"], r["sb"] = kernel_bytes(w.n, w.nb, prec == pkg.MIXED)
    r["ms_per_step"] = 1e3 * r["t"] / steps
    r["value"] = r["nb_alive"] * steps / r["t"]
    print(f"# {w.name} {'mixed' if prec == pkg.MIXED else 'fp64'}: {steps} steps graph {r['t']:.3f}s "
          f"eager+events {r['t_prof']:.3f}s bond {r['bond_ms']:.4f} ms gather {r['gather_ms']:.4f} ms "
          f"positions {r['pos_ms']:.4f} ms broken {st.n_broken} tile {tile_code(r['tile'])}",
          file=sys.stderr, flush=True)
    return r


def precision_block(w, r, mixed, dtype):
    kms = {"bond": r["bond_ms"], "gather": r["gather_ms"], "positions": r["pos_ms"]}
    ach = r["kb"] / (r["bond_ms"] * 1e-3) / 1e9
    return {
        "value": r["value"], "unit": "bond-evals/s", "dtype": dtype,
        "ms_per_step": r["ms_per_step"],
        "particle_steps_per_s": w.n * r["steps"] / r["t"],
        "step_hbm_frac": r["sb"] / (r["ms_per_step"] * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "kernel_ms": kms,
        "phase_hbm_frac": phase_fracs(phase_bytes(w.n, w.nb, mixed), kms),
        "tile": tile_info(r["tile"]),
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS,
                     "kernel": "k_bond<float>" if mixed else "k_bond<double>",
                     "traffic": read_traffic(w.name, "mixed" if mixed else "fp64", r["tile"])},
        "compute_roofline": compute_roofline(w.row, r["bond_ms"], mixed),
    }


def rel(a, b, sel=None):
    if sel is not None:
        a, b = a[sel], b[sel]
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="plate_4096")
    ap.add_argument("--extra", default="plate_1024,plates_96",
                    help="comma-separated extra workloads timed in both precisions ('' for none)")
    ap.add_argument("--extra-steps", type=int, default=50)
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
        res[name] = measure(pkg, w, prec, args, want_T=not args.only and not args.no_mixed, tile=args.tile)
    if args.only:
        return

    r = res["fp64"]
    fp = precision_block(w, r, False, "f64")
    out = {
        "metric": "bond-evals/s (FP64, 1 MI355X)",
        "value": r["value"],
        "unit": "bond-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic jittered plate (counter-based Philox, seed 0)",
        "config": {"workload": w.name, "particles": w.n, "bonds": w.nb, "horizon_dx": w.horizon,
                   "family_max": int(np.diff(w.row).max()), "notches": len(w.info.get("notches", [])),
                   "porosity": 0.01 if w.name == "plate_4096" else 0.0, "steps": args.steps},
        "particle_steps_per_s": fp["particle_steps_per_s"],
        "step_hbm_frac": fp["step_hbm_frac"],
        "copy_probe_gbs": copy_gbs,
        "kernel_ms": fp["kernel_ms"],
        "phase_hbm_frac": fp["phase_hbm_frac"],
        "tile": fp["tile"],
        "roofline": fp["roofline"],
        "compute_roofline": fp["compute_roofline"],
    }
    if "mixed" in res:
        m = res["mixed"]
        mb = precision_block(w, m, True, "f32/f64")
        free = np.repeat(w.code < 0, 2)
        err = {k: rel(m["state"][k], r["state"][k], free) for k in ("u", "v", "f")}
        alive = np.repeat((r["broken"] == 0) & (m["broken"] == 0), 2)
        err["T"] = rel(m["T"], r["T"], alive)
        del alive
        m["T"] = r["T"] = None
        mb.update({
            "speedup_vs_fp64": (m["nb_alive"] / m["t"]) / (r["nb_alive"] / r["t"]),
            "rel_err_vs_fp64": err,
            "rel_err_note": "norm-wise over the free particles (u, v, f) and over the bonds alive in both "
                            "runs (T, the force state of every directed bond); F and sigma are per-bond "
                            "registers of the bond kernel and are not stored",
            "broken_sets_equal": bool(np.array_equal(m["broken"], r["broken"])),
            "broken_sha1": {"fp64": mask_digest(r["broken"]), "mixed": mask_digest(m["broken"])},
            "broken": int(m["st"].n_broken),
            "guarded_families": int(m["st"].n_guarded_families),
        })
        out["mixed"] = mb
    res.clear()

    extras = {}
    for name in [x for x in args.extra.split(",") if x]:
        if name == w.name:
            continue
        t = time.perf_counter()
        we = pkg.plates.WORKLOADS[name]()
        print(f"# extra workload {we.name}: n={we.n} nb={we.nb} setup {time.perf_counter() - t:.1f}s",
              file=sys.stderr, flush=True)
        e = {}
        for pn, prec in (("fp64", pkg.FP64), ("mixed", pkg.MIXED)):
            if pn == "mixed" and args.no_mixed:
                continue
            q = measure(pkg, we, prec, args, steps=args.extra_steps, warmup=args.warmup)
            blk = precision_block(we, q, prec == pkg.MIXED, "f32/f64" if pn == "mixed" else "f64")
            e[pn] = {k: blk[k] for k in ("value", "ms_per_step", "particle_steps_per_s", "step_hbm_frac",
                                         "kernel_ms", "tile")}
            e[pn]["bond_hbm_frac"] = blk["roofline"]["frac"]
            e[pn]["broken_sha1"] = mask_digest(q["broken"])
        if "mixed" in e:
            e["mixed_vs_fp64"] = e["mixed"]["value"] / e["fp64"]["value"]
            e["broken_sets_equal"] = e["mixed"]["broken_sha1"] == e["fp64"]["broken_sha1"]
        e["particles"], e["bonds"], e["steps"] = we.n, we.nb, args.extra_steps
        extras[name] = e
        del we
    if extras:
        out["extra_workloads"] = extras

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, r["state"])
        out["vs_cpu_port"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
