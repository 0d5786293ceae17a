#!/usr/bin/env python3
"""Benchmark: FLIP steps/s of the melt/freeze workload at 256^3 (~90M particles).

One "step" is mpf_step: hash/sort, forces, level set + flags, P2G, extrapolation,
variable-density Jacobi-PCG projection to 1e-6 relative residual, G2P, advection
with solid collision, re-sort + level set, heat (Jacobi-10), melt/freeze and
marching cubes — on inputs already resident in HBM.  The mixed-precision form
(FP32 stencils/transfers, FP64 dots and refinement) is timed, then the all-FP64
PCG form on the same initial state; both solve to the same residual.

Prints ONE JSON line (rank 0, N=1).  Extra fields: fp64 steps/s and the ratio,
particle-steps/s, PCG iterations/s, whether every timed solve reached its
tolerance on the FP64 true residual (the run exits 1 otherwise), a roofline
record for the dominant kernel (average launch duration from HIP events on
the context stream over the timed steps), the same mixed/FP64 comparison on a
scene with air (the bubble column at 256^3: the solver's slow-row path) and on
the boiling scene at 256^3, and a CPU baseline (the oracle with OpenMP, FP32
solve, two steps of the same 256^3 workload on this host's cores).
"""
import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per launch, as (bytes per cell, bytes per unknown cell,
# bytes per particle): each array element the kernel must touch, read or
# written once (neighbour reads of a stencil come from cache).  Unknown cells
# are the cells of the pressure system (FLUID or AIR); the solver's vectors
# are read and written there only.  FP32 vectors of the MIXED form; FP64_VEC
# holds the FP64 form's figures.
KERNEL_BYTES = {
    # mask 1 per cell; r 4 + p_{k-1} 4 read, p_k 4 + q 4 written per unknown
    "pcg_spmv": (1, 16, 0),
    # mask 1 per cell; per unknown, averaged over the alternating launches: one
    # reads r, q and writes r (12), the next also reads x, p_k, p_{k-1} and
    # writes x (28)
    "pcg_update": (1, 20, 0),
    # mask 1 per cell; b 4, x64 8, x 4 read, r 4 written per unknown
    "pcg_resid64": (1, 20, 0),
    # one-kernel iteration (MIXED form without slow rows): mask 1 per cell;
    # per unknown r, q, p_k-1 read and r, q, p_k written (24), and x advanced
    # every other launch (read, written, p_k-2 read: 12), fully before each
    # true residual (x 8) and folded into x64 after it (x64 16 more): over a
    # cycle of 20 launches 9 x 24 + 9 x 36 + 32 + 48 = 620, 31 per launch
    "pcg_cg_iter": (1, 31, 0),
    # the same iteration forming q_{k-1} = A p_{k-1} again (cfg.pcg_recompute_q):
    # mask 1 per cell; r, p_{k-1}, x read and r, p_k, x written per unknown
    # (q_k is written by one launch in refine_every, before the true
    # residual: not counted)
    "pcg_cg_iterq": (1, 24, 0),
    # start/end 8, mask 1, ls fluid+air 8 written; pos 12 + flag 1 per particle
    "genlevelset": (17, 0, 13),
    # one launch, 3 axes: start/end 8, mask 1, 3 faces 12 written; pos 12, vel 12, mass 4, flag 1
    "p2g_liquid": (21, 0, 29),
    # one launch per axis: start/end 8, mask 1, face 4 written; particles as above
    "p2g_air": (13, 0, 29),
    # start/end 8, mask 1, T 4 written; pos 12, T 4, flag 1
    "heat_p2g": (13, 0, 17),
    # 45-byte record read and written + perm 4
    "reorder": (0, 0, 94),
    # position and flag read, position written (neighbours from cache)
    "correct_pos": (0, 0, 25),
    # level-set sweep: ls 4 read, mark 1, ls 4 written
    "sweep": (9, 0, 0),
    # extrapolation pass per face: u 4, out 4, ls 2x4, mark 2
    "sweep_u": (18, 0, 0),
}
FP64_VEC = {"pcg_spmv": (1, 32, 0), "pcg_update": (1, 40, 0)}

PROFILES = os.path.join(ROOT, "profiles")


def measured_traffic(kernel, form):
    """HBM bytes per launch of `kernel` in `form` from the committed rocprofv3
    counter summary (FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel of
    known size), or None."""
    path = os.path.join(PROFILES, "hbm_traffic_melt_freeze256.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 7 and cells[0] == "`%s`" % kernel and cells[1] == form:
            try:
                return float(cells[6])
            except ValueError:
                return None
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


LOG_STEPS = False  # --log-steps: one line per step (long profiler runs stay audible)


def step_done(frame, info):
    if LOG_STEPS:
        log("[bench]   step %d done, %d PCG iterations" % (frame, info["project"]["iterations"]))


def run_form(pkg, scenes, name, n, precision, steps, warmup, kernel_timing=True, **overrides):
    """W warm-up + K timed mpf_step calls of one form.  With kernel_timing the
    timed steps record a HIP-event pair around every launch on the context
    stream, so the per-kernel table and the roofline come from those steps."""
    s = scenes.make(name, n=n, precision=precision)
    cfg = s["cfg"]
    for k, v in overrides.items():
        setattr(cfg, k, v)
    sim = pkg.Simulator(cfg)
    sim.set_particles(s["n_particles"], **s["particles"])
    if len(s["seeds"]):
        sim.set_seed_cells(s["seeds"])
    del s["particles"]
    log("[bench] %s n=%d precision=%d particles=%d: %d warmup + %d timed steps" % (
        name, n, precision, s["n_particles"], warmup, steps))
    frame = 0
    for _ in range(warmup):
        step_done(frame, sim.step(frame, allow_no_convergence=True))
        frame += 1
    sim.synchronize()
    iters = 0
    t0 = time.perf_counter()
    infos = []
    for _ in range(steps):
        info = sim.step(frame, allow_no_convergence=True)
        infos.append(info)
        iters += info["project"]["iterations"]
        step_done(frame, info)
        frame += 1
    sim.synchronize()
    dt = time.perf_counter() - t0
    solves = [i["project"] for i in infos]
    log("[bench]   %.3f s for %d steps, PCG iterations %s, true rel. residuals %s" % (
        dt, steps, [p["iterations"] for p in solves], ["%.2e" % p["rel_residual"] for p in solves]))
    ktimes = None
    kt_iters = None
    if kernel_timing:
        # per-launch HIP events add ~7% to a step, so they run in a second
        # window of the same length right after the timed one
        sim.reset_kernel_times()
        sim.set_kernel_timing(True)
        kt_iters = []
        for _ in range(steps):
            info = sim.step(frame, allow_no_convergence=True)
            kt_iters.append(info["project"]["iterations"])
            solves.append(info["project"])
            frame += 1
        ktimes = sim.kernel_times()  # waits for the recorded events
        sim.set_kernel_timing(False)
        log("[bench]   kernel-timing window: PCG iterations %s" % kt_iters)
    mark = sim.get_field("MARK")
    unknowns = int(((mark == pkg.TYPEFLUID) | (mark == pkg.TYPEAIR)).sum())
    res = {"seconds": dt, "steps": steps, "pcg_iters": iters, "infos": infos, "ktimes": ktimes, "kt_iters": kt_iters,
           "particles": sim.particle_count(), "G": cfg.nx * cfg.ny * cfg.nz, "unknowns": unknowns,
           "air_cells": int((mark == pkg.TYPEAIR).sum()),
           "converged": all(p["converged"] == 1 for p in solves),
           "max_rel_residual": max(p["rel_residual"] for p in solves), "tol": cfg.tol}
    sim.close()
    return res


def roofline(res, precision, form):
    """Roofline record of the kernel with the largest total time over the
    timed steps (HIP events on the context stream)."""
    kt = res["ktimes"] or {}
    if not kt:
        return None, None
    K = res["steps"]
    dom = max(kt.items(), key=lambda kv: kv[1][1])
    name, (cnt, ms) = dom
    table = dict(KERNEL_BYTES)
    if precision == 1:
        table.update(FP64_VEC)
    per_cell, per_unk, per_par = table.get(name, (0, 0, 0))
    G, P, U = res["G"], res["particles"], res["unknowns"] or 0
    bytes_per_launch = per_cell * G + per_unk * U + per_par * P
    avg_ms = ms / max(cnt, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if bytes_per_launch else None
    rf = {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
          "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": measured_traffic(name, form),
          "avg_launch_us": avg_ms * 1e3, "launches_per_step": cnt / K, "bytes_per_launch": bytes_per_launch,
          "cells": G, "unknown_cells": U, "particles": P,
          "measured_over": "HIP events on the context stream over the %d steps right after the timed ones "
                           "(PCG iterations per step %.1f there, %.1f in the timed steps)" % (
                               K, sum(res["kt_iters"]) / K, res["pcg_iters"] / K)}
    ranked = sorted(((k, v[0] / K, v[1] / K) for k, v in kt.items()), key=lambda kv: -kv[2])
    log("[bench]   kernel times per step of the kernel-timing window (name, launches, ms): " +
        ", ".join("%s %.1f %.3f" % r for r in ranked))
    top = [(k, ms_) for k, _, ms_ in ranked[:8]]
    return rf, top


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scenes, name, n, steps, warm):
    """The oracle (the reference's kernel bodies as OpenMP loops, FP32 fields
    and an FP32 pressure solve as the reference's, dots summed in double) on
    the same workload at full size, timed on this host."""
    import oracle
    s = scenes.make(name, n=n)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = oracle.set_threads(threads)
    o = oracle.Oracle(s["cfg"])
    o.set_solve_float(True)
    o.set_particles(s["n_particles"], **s["particles"])
    if len(s["seeds"]):
        o.set_seed_cells(s["seeds"])
    for f in range(warm):
        o.step(f, allow_no_convergence=True)
    t0 = time.perf_counter()
    its = []
    for f in range(warm, warm + steps):
        its.append(o.step(f, allow_no_convergence=True)["project"]["iterations"])
    dt = time.perf_counter() - t0
    o.close()
    return {"value": steps / dt, "unit": "steps/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": "%s %d^3 (%d particles, the measured workload itself), %d step(s) after %d untimed, "
                      "%.1f s on %d OpenMP threads, FP32 fields and FP32 PCG (PCG iterations %s)" % (
                          name, n, s["n_particles"], steps, warm, dt, threads, its)}


def other_scene(pkg, scenes, name, n, steps, warmup, tol, max_iters):
    """Mixed over FP64 step rate on another scene (with air cells the
    operator has slow rows: the single-kernel iteration with its slow-row
    list in the mixed form, the two-kernel iteration in FP64)."""
    mixed = run_form(pkg, scenes, name, n, pkg.MIXED, steps, warmup, kernel_timing=False, tol=tol,
                     max_iters=max_iters)
    fp64 = run_form(pkg, scenes, name, n, pkg.FP64, steps, warmup, kernel_timing=False, tol=tol,
                    max_iters=max_iters)
    m = mixed["steps"] / mixed["seconds"]
    f = fp64["steps"] / fp64["seconds"]
    return {"workload": "%s %d^3, %d particles, PCG to %.0e rel. residual in both forms" % (
                name, n, mixed["particles"], tol),
            "air_cells_after": mixed["air_cells"],
            "steps": steps, "warmup": warmup, "mixed_steps_per_s": m, "fp64_steps_per_s": f,
            "mixed_over_fp64": m / f, "mixed_pcg_iters_per_step": mixed["pcg_iters"] / mixed["steps"],
            "fp64_pcg_iters_per_step": fp64["pcg_iters"] / fp64["steps"],
            "converged": mixed["converged"] and fp64["converged"],
            "max_rel_residual": max(mixed["max_rel_residual"], fp64["max_rel_residual"])}, [mixed, fp64]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scene", default="melt_freeze")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-air", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="time the steps without the per-launch HIP events (no kernel table, no roofline)")
    ap.add_argument("--air-scene", default="bubble_column")
    ap.add_argument("--air-tol", type=float, default=1e-6)
    ap.add_argument("--air-max-iters", type=int, default=30000)
    ap.add_argument("--air-n", type=int, default=256)
    ap.add_argument("--air-steps", type=int, default=5)
    ap.add_argument("--air-warmup", type=int, default=2)
    ap.add_argument("--no-boiling", action="store_true")
    ap.add_argument("--boiling-n", type=int, default=256)
    ap.add_argument("--boiling-steps", type=int, default=5)
    ap.add_argument("--boiling-warmup", type=int, default=2)
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-warmup", type=int, default=0)
    ap.add_argument("--log-steps", action="store_true", help="log every step's end (for profiler runs)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a kernel of known HBM "
                         "traffic, for scaling profiler byte counters)")
    args = ap.parse_args()
    global LOG_STEPS
    LOG_STEPS = args.log_steps
    if args.gpus != 1:
        print("bench.py measures one GPU only: the z-slab multi-GPU path is not built; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    pkg = importlib.import_module("2017multiphase_amd")
    scenes = importlib.import_module("2017multiphase_amd.scenes")
    if args.copy_calibration_mib > 0:
        hm = importlib.import_module("2017multiphase_amd.hipmem")
        n = args.copy_calibration_mib << 20
        src = hm.DeviceArray(n, "u1")
        dst = hm.DeviceArray(n, "u1")
        src.fill_bytes(1)
        hm.synchronize()
        dst.copy_from(src, n)
        hm.synchronize()
        log("[bench] copy calibration: %d bytes read, %d bytes written" % (n, n))
        del src, dst
    kt = not args.no_kernel_timing
    mixed = run_form(pkg, scenes, args.scene, args.n, pkg.MIXED, args.steps, args.warmup, kernel_timing=kt)
    fp64 = None
    if not args.no_fp64:
        fp64 = run_form(pkg, scenes, args.scene, args.n, pkg.FP64, args.steps, args.warmup, kernel_timing=kt)
    sps = mixed["steps"] / mixed["seconds"]
    rf, top = roofline(mixed, pkg.MIXED, "mixed")
    out = {
        "metric": "FLIP steps/s at 256^3 (mixed vs FP64 PCG, same residual); % HBM peak",
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mixed["seconds"] / mixed["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (PCG dots, norms and refinement residual f64)",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": "%s %d^3, %d particles, variable-density Jacobi-PCG to 1e-6 rel. residual, "
                               "mixed precision (refine every 20)" % (args.scene, args.n, mixed["particles"]),
                   "grid": args.n, "particles": mixed["particles"], "parallelism": "single GPU"},
        "particle_steps_per_s": sps * mixed["particles"],
        "pcg_iters_per_s": mixed["pcg_iters"] / mixed["seconds"],
        "pcg_iters_per_step": mixed["pcg_iters"] / mixed["steps"],
        "converged": mixed["converged"],
        "max_true_rel_residual": mixed["max_rel_residual"],

        "roofline": rf,
        "top_kernels_ms_per_step": top,
    }
    forms = [mixed]
    if fp64:
        f_sps = fp64["steps"] / fp64["seconds"]
        out["fp64_steps_per_s"] = f_sps
        out["mixed_over_fp64"] = sps / f_sps
        out["fp64_pcg_iters_per_step"] = fp64["pcg_iters"] / fp64["steps"]
        out["fp64_converged"] = fp64["converged"]
        out["fp64_max_true_rel_residual"] = fp64["max_rel_residual"]
        rf64, _ = roofline(fp64, pkg.FP64, "fp64")
        out["fp64_roofline"] = rf64
        forms.append(fp64)
    if not args.no_air:
        out["air_scene"], more = other_scene(pkg, scenes, args.air_scene, args.air_n, args.air_steps,
                                             args.air_warmup, args.air_tol, args.air_max_iters)
        forms += more
    if not args.no_boiling:
        out["boiling_scene"], more = other_scene(pkg, scenes, "boiling", args.boiling_n, args.boiling_steps,
                                                 args.boiling_warmup, 1e-6, args.air_max_iters)
        forms += more
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scenes, args.scene, args.cpu_n, args.cpu_steps, args.cpu_warmup)
    print(json.dumps(out), flush=True)
    bad = [f for f in forms if not f["converged"] or f["max_rel_residual"] > f["tol"] * (1 + 1e-9)]
    if bad:
        log("[bench] a timed solve did not reach its tolerance: %s" % [(f["max_rel_residual"], f["tol"]) for f in bad])
        sys.exit(1)


if __name__ == "__main__":
    main()
