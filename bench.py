#!/usr/bin/env python3
"""Benchmark of the EDGE flow-map time step on one MI355X.

Workload: 256x128x128 cells, four seeded vortex rings, reinit every 12 steps,
CFL 0.5, BFECC clamp, projection to ||r|| <= 1e-6 ||b|| (FP64 norms).  A step
is one iteration of the reference's loop (run_edge.py:588-711): semi-Lagrangian
half step + projection, EDGE march, face march, BFECC pull-back, projection.

Prints ONE JSON line: ms/step with the chosen projection solver (value; by
default "mixed": BF16 multigrid levels, FP64 dots and dense FP64 coarsest
solve, each projection started from the same projection's pressure of the
previous step), the plain FP32 reference solver from zero beside it, Poisson
ms and iterations of plain, mixed, dense and the learned warm start on the
same right-hand sides (from zero), per-phase times, the roofline of the
dominant kernel and of the other main kernels (HIP events on the launch
stream), the step's error against the oracle and the CPU baseline (the
oracle restatement running one step of the same workload on this host's
cores).

Timing: `value` is K steps run back to back (as the step loop runs them)
between two synchronisations, divided by K; `step_ms`, `phases_ms` and the
per-launch kernel times come from K further steps, each synchronised and
carrying HIP events, which add a little host time per step.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--res NX NY NZ]
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "edge-flow-maps_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# algorithmic work per unit.  Flops: the FP32 arithmetic of the implemented
# separable formulation (FMA = 2), i.e. 64 lanes x (2 FMA + ADD + MUL) wave
# instructions per launch from SQ_INSTS_VALU_{FMA,ADD,MUL}_F32, divided by the
# cells (256x128x128, four rings); the reference's formulation (64 weight
# products per Hermite evaluation, 27 per B-spline) is 33,600 flop/cell for
# EDGE.  Bytes: compulsory FP32 reads + writes.
MARCH_FLOP_PER_CELL = 22250.0      # fused EDGE march: 5 points x (RK4 with Jacobian + Hermite compose)
# ED8: 8 points x (RK4 of the point + full-Hermite value); ED4: 4 points x (RK4 + FD-unrolled Hermite value);
# GE: 1 point (RK4 with Jacobian + Hermite gradient)
SCHEME_FLOP_PER_CELL = {"edge": MARCH_FLOP_PER_CELL, "ed8": 16550.0, "ed4": 8050.0, "ge": 4430.0}
# the face march (RK4 of a face point and its column): per face and velocity in EFM's backtrack
FACE_MARCH_FLOP_PER_FACE = 2860.0
# the reference formulation's count (64 weight products per Hermite evaluation, 27 per
# B-spline evaluation), the algorithmic work the roofline record is priced with
SCHEME_FLOP_PER_CELL_ALG = {"edge": 33600.0, "ed8": 21400.0, "ed4": 14500.0, "ge": 6800.0}
SCHEME_BYTES_PER_CELL = {"edge": 228.0, "ed8": 204.0, "ed4": 108.0, "ge": 324.0}
MG_BYTES_PER_CELL = {"plain": 145.0, "mixed": 98.0, "dense": 145.0}  # one CG iteration
MARCH_BYTES_PER_CELL = 228.0       # u 12 + old map 84 + new map 84 + its table's {psi, T} planes 48 (FP32)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def march_workload(res, case, scheme="edge"):
    """The map march's workload (what the march traffic counters depend on)."""
    what = {"cfg3": "delta-wing plate, oblique inflow", "cfg3plate": "four rings + plate span 0.3, inflow 1 at 20 deg",
            "cfg5": "one ring + passive dye particles", "cfg5laden": "one ring + laden particles"}.get(case, "four rings")
    return f"{scheme.upper()} step {res[0]}x{res[1]}x{res[2]} {what}"


def workload_name(res, case, scheme="edge", solver="dense"):
    solid = "solid " if case in ("cfg3", "cfg3plate") else ""
    return f"{march_workload(res, case, scheme)}, {solid}{solver} projection"


SOLVER_NOTE = {"plain": "plain (FP32 multigrid, the reference V-cycle with 20 bottom pairs, FP64 dots)",
               "mixed": "mixed (BF16 multigrid, FP64 dots, dense FP64 coarsest)",
               "dense": "dense (FP32 multigrid, FP64 dots, dense FP64 coarsest)"}


def read_pmc(d, counter):
    """{kernel short name: [per-dispatch value]} from a rocprofv3 --pmc csv directory."""
    import csv
    import glob
    out = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            if r.get("Counter_Name") != counter:
                continue
            name = r["Kernel_Name"].replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0]
            out.setdefault(name.split("<")[0], []).append(float(r["Counter_Value"]))
    return out


# kernels whose HBM traffic the counter runs record: short name -> algorithmic bytes per
# launch at 256x128x128 (4,194,304 cells, 12,664,832 faces) and what they are
TRAFFIC_KERNELS = {
    "k_march_edge": (228.0 * 4194304, "cell: u 12 + old map 84 + new map 84 + the new table's {psi, T} planes 48 B"),
    "k_map_u": (32.0 * 12664832, "face: u0 12 (gathered) + phi 12 + F 4 + out 4 B"),
    "k_advect_sl": (8.0 * 12664832, "face: u in 4 + out 4 B"),
    "k_advance_faces": (52.0 * 12664832, "face: phi, F read and written 48 + u 4 B"),
    "k_map_smoke": (68.0 * 4194304, "cell: faces of phi 36 + smoke in 16 + out 16 B"),
    "k_split_psi": (48.0 * 4194304 + 24.0 * 12664832,
                    "cell: the garm table's {psi, T} planes 48 B read once (face centres weigh no dT, d2T); "
                    "face: phi, F out 24 B"),
    "k_smooth8": (6.0 * 4194304, "cell: r 2 + z read 2 + z written 2 B (one half-sweep of the mixed projection's "
                                 "BF16 level 0)"),
}


def summarize_pmc(fetch_dir, write_dir, probe_mib):
    """HBM bytes per launch of the step's main kernels from FETCH_SIZE / WRITE_SIZE
    (KiB per dispatch), each scaled by what the counter reads for the copy probe of
    known size; the algorithmic bytes per launch beside them."""
    import numpy as np
    f, w = read_pmc(fetch_dir, "FETCH_SIZE"), read_pmc(write_dir, "WRITE_SIZE")
    known_kib = probe_mib * 1024.0
    # the probe is the last k_scale dispatch (the step's own k_scale launches come first)
    f_probe, w_probe = f["k_scale"][-1], w["k_scale"][-1]
    cf, cw = known_kib / f_probe, known_kib / w_probe
    rec = {}
    for name, (alg, what) in TRAFFIC_KERNELS.items():
        if name not in f or name not in w:
            continue
        fm, wm = float(np.median(f[name])), float(np.median(w[name]))
        hbm = (fm * cf + wm * cw) * 1024.0
        rec[name] = {
            "fetch_kib_per_launch": round(fm, 1), "write_kib_per_launch": round(wm, 1),
            "copy_probe_mib": probe_mib, "copy_fetch_kib": round(f_probe, 1), "copy_write_kib": round(w_probe, 1),
            "scale_fetch": round(cf, 4), "scale_write": round(cw, 4),
            "hbm_bytes_per_launch": round(hbm), "algorithmic_bytes_per_launch": round(alg),
            "traffic_over_algorithmic": round(hbm / alg, 2), "algorithmic": what,
            "launches": len(f[name]),
        }
    return rec

def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--res", type=int, nargs=3, default=[256, 128, 128])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-path comparison")
    ap.add_argument("--poisson-reps", type=int, default=3, help="0 skips the same-RHS solver comparison")
    ap.add_argument("--no-io", action="store_true", help="skip timing one frame output and one checkpoint")
    ap.add_argument("--solver", default=None, choices=["plain", "mixed", "dense"],
                    help="projection solver of the timed steps (default: mixed; dense for the 512x256x256 cases, "
                         "where the BF16 levels take 5 + 8 iterations against 4 + 5)")
    ap.add_argument("--start", default="extrapolated", choices=["zero", "previous", "extrapolated"],
                    help="where each projection's solve starts: zero (the reference), the same projection's "
                         "pressure of the previous step, or (default) the linear extrapolation in time of that "
                         "projection's last two pressures")
    ap.add_argument("--zero-start", action="store_true", help="the same as --start zero")
    ap.add_argument("--case", default="cfg2", choices=["cfg2", "cfg3", "cfg3plate", "cfg5", "cfg5laden"],
                    help="cfg2: the headline EDGE step; cfg3: the reference's delta-wing plate with the solid "
                         "projection; cfg3plate: the four rings plus a plate of span 0.3 in an inflow of speed 1 at "
                         "20 degrees (solid projection); cfg5: 512x256x256 with 16M passive dye particles; "
                         "cfg5laden: 512x256x256 with 16M laden particles")
    ap.add_argument("--particles", type=int, default=None, help="particle count for --case cfg5 / cfg5laden")
    ap.add_argument("--scheme", default="edge", choices=["edge", "ed8", "ed4", "ge", "efm"],
                    help="backward-map march: EDGE (run_edge.py), ED8 (run_ed8.py), ED4 (run_ed4.py), GE "
                         "(run_ge.py) or EFM (run_efm.py)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the run, one plain copy kernel (k_scale, alpha 1) over MIB MiB: the known-size "
                         "launch that scales rocprofv3's FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="turn two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of this script with "
                         "--copy-probe into profiles/traffic.json and exit")
    a = ap.parse_args()
    if a.solver is None:
        a.solver = "dense" if a.case in ("cfg5", "cfg5laden") else "mixed"
    return a


def phase_times(sim):
    """Per-phase ms of one step from its recorded marks."""
    marks = sim.timing.pop("_marks")
    out = {}
    for (n0, e0), (n1, e1) in zip(marks[:-1], marks[1:]):
        out[n0] = out.get(n0, 0.0) + e0.elapsed_time(e1)
    return out


def timed_steps(sim, k, torch):
    """k steps back to back between two synchronisations, nothing recorded in
    between (as the step loop runs them); returns ms per step."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        sim.substep()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / k


def run_steps(sim, k, torch):
    """k instrumented steps, each synchronised on both sides, with HIP events at
    the phase boundaries; returns wall ms per step (list) and phase ms (list of dicts)."""
    walls, phases = [], []
    for _ in range(k):
        sim.timing = {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.substep()
        torch.cuda.synchronize()
        walls.append((time.perf_counter() - t0) * 1e3)
        phases.append(phase_times(sim))
        sim.timing = None
    return walls, phases


def make_solver(sim, mode, warm_start=None):
    """A fresh projection solver for the simulation's grid (and its solids)."""
    import numpy as np
    from sim.mgpcg import MGPCG_3
    from sim.mgpcg_solid import MGPCG_Solid_3
    bt = np.full((3, 2), 2, np.int32)
    N = [sim.res_x, sim.res_y, sim.res_z]
    base = {"learned": "mixed", "learned_dense": "dense"}.get(mode, mode)
    kw = dict(base_level=3, mode=base, stop="relative", rel_tol=1e-6, warm_start=warm_start)
    if sim.solid:
        return MGPCG_Solid_3(bt, sim.boundary_mask, sim.boundary_vel, N, **kw)
    return MGPCG_3(bt, N, **kw)


def poisson_compare(sim, reps, torch):
    """Plain vs mixed vs dense (vs learned warm start + mixed, when the trained
    network is present) projection of the same velocities (the inputs of one
    step's two projections), to ||r|| <= 1e-6 ||b||."""
    import math
    sim.rhs_snapshots = []
    sim.substep()
    rhs = sim.rhs_snapshots
    sim.rhs_snapshots = None
    res = {}
    from sim.learned import load_default
    net = load_default()
    modes = ("plain", "mixed", "dense") + (("learned", "learned_dense") if net is not None else ())
    for mode in modes:
        solver = make_solver(sim, mode, warm_start=net if mode.startswith("learned") else None)
        per = []
        for which, base in enumerate(rhs):
            times, iters, conv = [], [], True
            for r in range(reps + 1):
                for f, b in zip((sim.tmp_u_x, sim.tmp_u_y, sim.tmp_u_z), base):
                    f.data.copy_(b)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                solver.Poisson(sim.tmp_u_x, sim.tmp_u_y, sim.tmp_u_z, max_iters=sim.poisson_max_iters)
                e1.record()
                torch.cuda.synchronize()
                if r > 0:
                    times.append(e0.elapsed_time(e1))
                    iters.append(solver.stats["iters"])
                    conv = conv and bool(solver.stats["converged"])
            st = solver.stats
            per.append(dict(ms=round(statistics.median(times), 3), iters=int(statistics.median(iters)),
                            converged=conv, rel_res=float("%.3g" % math.sqrt(max(st["rTr"], 0.0) / max(st["b2"], 1e-300))),
                            **({"x0_used": bool(solver.stats["x0_used"]),
                                "guess_rel_res": float("%.3g" % math.sqrt(solver.stats["r02"] /
                                                                          max(solver.stats["b2"], 1e-300)))}
                               if mode.startswith("learned") else {})))
        res[mode] = dict(per_solve=per, ms=round(sum(p["ms"] for p in per), 3),
                         iters=sum(p["iters"] for p in per))
        del solver
    res["speedup_mixed_vs_plain"] = round(res["plain"]["ms"] / res["mixed"]["ms"], 3)
    res["speedup_dense_vs_plain"] = round(res["plain"]["ms"] / res["dense"]["ms"], 3)
    if "learned" in res:
        res["speedup_learned_vs_plain"] = round(res["plain"]["ms"] / res["learned"]["ms"], 3)
        res["speedup_learned_dense_vs_plain"] = round(res["plain"]["ms"] / res["learned_dense"]["ms"], 3)
        res["learned_note"] = ("learned: the mixed solver started from the U-Net's p0, learned_dense: the dense "
                               "solver from it (inference included in ms)")
    return res


def io_timing(sim, torch):
    """One frame output (vorticity on the GPU, .vti) and one checkpoint save and load
    (velocity and smoke .npy, map state .npz) into a temporary directory, after the timed
    steps.  Host-side file work: reported beside the step, never part of it."""
    import shutil
    import tempfile
    from edge_hip import _lib
    from edge_hip.field import stream_handle
    d = tempfile.mkdtemp(prefix="edge_io_")
    try:
        wn, sm, _, _ = sim._frame_buffers()
        args = (sim.u_x.ptr, sim.u_y.ptr, sim.u_z.ptr, sim.smoke.ptr, 4, wn.data_ptr(), sm.data_ptr(), sim.res_x,
                sim.res_y, sim.res_z, float(sim.dx), stream_handle())
        _lib.call("edge_vtk_frame", *args)   # first launch (code object load) untimed
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()   # the frame kernel alone (|curl u| and smoke in VTK order), on torch's stream
        _lib.call("edge_vtk_frame", *args)
        e1.record()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.write_frame(d, frame=0)
        t1 = time.perf_counter()
        files = sim.save_checkpoint(d, frame=0)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        restored = sim.load_checkpoint(d, 0)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        fb = os.path.getsize(os.path.join(d, "field_0000.vti"))
        cb = sum(os.path.getsize(f) for f in files)
        return dict(frame_kernel_ms=round(e0.elapsed_time(e1), 3), frame_ms=round((t1 - t0) * 1e3, 1),
                    frame_bytes=fb, checkpoint_save_ms=round((t2 - t1) * 1e3, 1), checkpoint_bytes=cb,
                    checkpoint_load_ms=round((t3 - t2) * 1e3, 1), state_restored=bool(restored),
                    where="a temporary directory on the host's local disk")
    finally:
        shutil.rmtree(d, ignore_errors=True)


def _rel(a, b):
    """max |a - b| / max |b| over a list of arrays (FP64)."""
    import numpy as np
    num = max(float(np.max(np.abs(np.asarray(x, np.float64) - np.asarray(y, np.float64)))) for x, y in zip(a, b))
    den = max(float(np.max(np.abs(np.asarray(y, np.float64)))) for y in b)
    return num / max(den, 1e-300)


def cpu_baseline(sim):
    """The oracle (the CPU restatement, OpenMP) running one full step of the
    measured workload itself, from the simulation's current state: about 20 s of
    CPU work on 16 threads, no scaling.  The GPU takes the same step from the
    same state (a reinitialisation step, as the oracle's first one is), and the
    step's velocity and pressure are compared with the FP32 oracle and with the
    oracle in FP64 (the FP32 oracle's own distance from FP64 is the yardstick)."""
    import numpy as np
    import torch
    from oracle.step import OracleStep
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    u = sim.velocity_numpy()
    smoke = np.moveaxis(sim.smoke.to_numpy(), -1, 0)
    # the GPU step: make it the first step of a reinitialisation interval
    sim.i = (sim.i // sim.reinit_every + 1) * sim.reinit_every - 1
    dt = sim.substep()
    torch.cuda.synchronize()
    gu = sim.velocity_numpy()
    gp = sim.solver.p.to_numpy().astype(np.float64)

    def oracle(dtype):
        return OracleStep(u, smoke, sim.dx, sim.reinit_every, sim.CFL, sim.ed_eps, levels=sim.solver.n_mg_levels,
                          BFECC_clamp=sim.BFECC_clamp, stop_mode=1, rel_tol=1e-6, dtype=dtype)
    st = oracle(np.float32)
    t0 = time.perf_counter()
    st.substep(dt=dt)
    elapsed = time.perf_counter() - t0
    st64 = oracle(np.float64)
    st64.substep(dt=dt)
    mf = lambda x: np.asarray(x, np.float64) - np.mean(x)
    n = f"{sim.res_x}x{sim.res_y}x{sim.res_z}"
    base = dict(value=round(elapsed * 1e3, 1), unit="ms/step", cores=threads, kind="port",
                sample=f"one full oracle step (reinitialisation, advection, EDGE march, pull-backs, two "
                       f"projections to 1e-6; FP32, OpenMP) of the measured {n} workload from the timed "
                       f"run's state, {elapsed:.1f} s, unscaled")
    err = dict(step="one GPU step from the timed run's state against the oracle's step from the same state "
                    "(same dt; velocity: max |diff| / max |ref| over the three face grids; pressure of the "
                    "second projection, mean removed)",
               velocity_vs_fp32_oracle=float("%.3g" % _rel(gu, st.u)),
               velocity_vs_fp64_oracle=float("%.3g" % _rel(gu, st64.u)),
               fp32_oracle_vs_fp64_oracle=float("%.3g" % _rel(st.u, st64.u)),
               pressure_vs_fp64_oracle=float("%.3g" % _rel([mf(gp)], [mf(st64.p)])),
               fp32_oracle_pressure_vs_fp64=float("%.3g" % _rel([mf(st.p)], [mf(st64.p)])),
               poisson_iters_gpu=[s["iters"] for s in sim.poisson_stats[-2:]],
               poisson_iters_fp32_oracle=list(st.iters), poisson_iters_fp64_oracle=list(st64.iters))
    return base, err


def poisson_in_step(ph, plain):
    """The step's two projections (median phase times of the instrumented steps)
    with the timed solver and start, against the plain solver from zero on the
    same steps' own right-hand sides (the reference's solver and start)."""
    if plain is None or "poisson1" not in ph:
        return None
    ours = ph["poisson1"] + ph["poisson2"]
    ref = plain["phases_ms"]["poisson1"] + plain["phases_ms"]["poisson2"]
    return {"ms": round(ours, 3), "plain_zero_start_ms": round(ref, 3),
            "speedup_vs_plain": round(ref / ours, 3) if ours > 0 else None}


START_NOTE = {
    "zero": "zero (the reference)",
    "previous": "the same projection's pressure of the previous step",
    "extrapolated": "the linear extrapolation in time of the same projection's pressures of the previous two "
                    "steps (the previous pressure alone after a dropped guess)",
}


def main():
    a = parse()
    if a.summarize_pmc:
        rec = summarize_pmc(*a.summarize_pmc, probe_mib=a.copy_probe or 1024)
        for r in rec.values():
            r["workload"] = march_workload(a.res, a.case)
        with open(TRAFFIC_FILE, "w") as fh:
            json.dump(rec, fh, indent=1)
        print(json.dumps(rec))
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N > 1 is not supported "
                                   "(multi-GPU slab decomposition is a design note only)"}))
        sys.exit(2)
    import torch
    from run_edge import Simulation, freeze_setup_objects

    big = a.case in ("cfg5", "cfg5laden")
    res = tuple(a.res) if (not big or a.res != [256, 128, 128]) else (512, 256, 256)
    sim = Simulation(a.case, solver_mode=a.solver, stop="relative", rel_tol=1e-6, res=res, check_every=0,
                     pressure_start="zero" if a.zero_start else a.start,
                     num_particles=a.particles, scheme=a.scheme)
    sim.initialize()
    freeze_setup_objects()   # as Simulation.run does before its loop
    for _ in range(a.warmup):
        sim.substep()
    # the timed steps: back to back, as Simulation.run steps (the projection's
    # convergence checks and the CFL readback are the only host waits)
    ms_step = timed_steps(sim, a.steps, torch)
    pstats = sim.poisson_stats[-2 * a.steps:]
    step_iters = [s["iters"] for s in pstats]
    # then as many instrumented steps: per-step walls, phases and the HIP-event
    # times of the roofline kernels' launches (events and synchronisations add
    # a little host time that the timed steps do not carry)
    from edge_hip import timing
    timing.reset()
    timing.enable(True)
    i_first = sim.i + 1
    walls, phases = run_steps(sim, a.steps, torch)
    timing.enable(False)
    march_kernel = {"edge": "k_march_edge", "ed8": "k_march_ed8", "ed4": "k_march_ed4", "ge": "k_march_ge",
                    "efm": "k_backtrack_faces"}[a.scheme]
    march_launch_ms = timing.durations(march_kernel)
    other_launch_ms = {k: timing.durations(k) for k in ("k_map_u", "k_advect_sl", "k_advance_faces")}
    ph = {k: statistics.median(p.get(k, 0.0) for p in phases) for k in phases[0]}

    plain = None
    if not a.no_plain:
        # the reference's solver from the reference's zero start
        step_solver, reuse = sim.solver, sim.reuse_pressure
        sim.solver, sim.reuse_pressure = make_solver(sim, "plain"), False
        run_steps(sim, 1, torch)  # warm the plain solver
        pms = timed_steps(sim, a.steps, torch)
        plain_iters = [s["iters"] for s in sim.poisson_stats[-2 * a.steps:]]
        pw, pph = run_steps(sim, a.steps, torch)
        plain = dict(ms_per_step=round(pms, 3), step_ms=[round(w, 3) for w in pw],
                     phases_ms={k: round(statistics.median(p.get(k, 0.0) for p in pph), 3) for k in pph[0]},
                     poisson_iters=plain_iters,
                     projection_start="zero")
        sim.solver, sim.reuse_pressure = step_solver, reuse

    pcmp = poisson_compare(sim, a.poisson_reps, torch) if a.poisson_reps > 0 else None

    ncell = sim.res_x * sim.res_y * sim.res_z
    wl = workload_name(res, a.case, a.scheme, a.solver)
    # the dominant kernel: the fused EDGE march, one launch per step, timed per launch with
    # HIP events on its stream; FP32 vector-ALU bound (no matrix op on this path)
    march_ms = statistics.mean(march_launch_ms)
    if a.scheme == "efm":
        # step i re-marches every face through (i mod reinit_every) + 1 velocities in one launch:
        # the work per launch is the mean over the timed steps
        nface = 3 * ncell + sim.res_y * sim.res_z + sim.res_x * sim.res_z + sim.res_x * sim.res_y
        marches = statistics.mean((i % sim.reinit_every) + 1 for i in range(i_first, i_first + a.steps))
        flop_launch = FACE_MARCH_FLOP_PER_FACE * nface * marches
        # compulsory: the face map written once (24 B per face) and one velocity read per march
        bytes_launch = (24.0 + 4.0 * marches) * nface
    else:
        flop_launch = SCHEME_FLOP_PER_CELL.get(a.scheme, 0.0) * ncell   # 0: not counted for this kernel
        bytes_launch = SCHEME_BYTES_PER_CELL[a.scheme] * ncell
    march_tflops = flop_launch / (march_ms * 1e-3) / 1e12
    # the same launches priced with the reference formulation's count (64 weight products per
    # Hermite evaluation, 27 per B-spline evaluation): the algorithmic figure per cell
    flop_launch_alg = SCHEME_FLOP_PER_CELL_ALG.get(a.scheme, 0.0) * ncell if a.scheme != "efm" else flop_launch
    march_tflops_alg = flop_launch_alg / (march_ms * 1e-3) / 1e12
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and a.scheme == "edge":
        rec = json.load(open(TRAFFIC_FILE)).get("k_march_edge", {})
        if rec.get("workload") == march_workload(res, a.case, a.scheme):
            traffic = rec.get("hbm_bytes_per_launch")
    roof_march = dict(bound="valu", pipe="FP32 vector ALU (no matrix op on this path); 157.3 TFLOP/s is "
                                         "MI355X's FP32 vector peak",
                      achieved=round(march_tflops_alg, 2), peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                      frac=round(march_tflops_alg / FP32_PEAK_TFLOPS, 4), traffic=traffic, kernel=march_kernel,
                      flop_source="algorithmic: the reference formulation's flops per cell",
                      flop_per_cell=SCHEME_FLOP_PER_CELL_ALG.get(a.scheme), flop_per_launch=flop_launch_alg,
                      counter_flop_per_cell=SCHEME_FLOP_PER_CELL.get(a.scheme),
                      counter_achieved=round(march_tflops, 2) if flop_launch else None,
                      counter_frac=round(march_tflops / FP32_PEAK_TFLOPS, 4) if flop_launch else None,
                      algorithmic_bytes_per_launch=bytes_launch,
                      avg_launch_ms=round(march_ms, 4), launches=len(march_launch_ms))
    # the step's other main kernels, each against its own bound (algorithmic work per unit
    # from the reference formulation; face counts of the MAC grids)
    nface = 3 * ncell + sim.res_y * sim.res_z + sim.res_x * sim.res_z + sim.res_x * sim.res_y
    tr_all = json.load(open(TRAFFIC_FILE)) if os.path.exists(TRAFFIC_FILE) else {}
    kernels = {}
    for name, bound, per_unit, unit_note in (
            ("k_map_u", "hbm", 32.0, "32 B per face (u0 gathered 12, phi 12, F 4, out 4)"),
            ("k_advect_sl", "valu", 1980.0, "1,980 flop per face (RK4 + foot lookup, 5 B-spline MAC evaluations)"),
            ("k_advance_faces", "valu", 3400.0, "3,400 flop per face (RK4 of the face point and its column)")):
        ms = other_launch_ms.get(name) or []
        if not ms:
            continue
        t = statistics.mean(ms)
        work = per_unit * nface
        if bound == "hbm":
            ach, peak, unit = work / (t * 1e-3) / 1e9, HBM_PEAK_GBS, "GB/s"
        else:
            ach, peak, unit = work / (t * 1e-3) / 1e12, FP32_PEAK_TFLOPS, "TFLOP/s"
        tr = tr_all.get(name, {})
        kernels[name] = dict(bound=bound, achieved=round(ach, 2), peak=peak, unit=unit, frac=round(ach / peak, 4),
                             avg_launch_ms=round(t, 4), launches=len(ms), per_face=unit_note,
                             traffic=tr.get("hbm_bytes_per_launch") if tr.get("workload") == march_workload(
                                 res, a.case, a.scheme) else None)
    # the projection's CG iteration (many kernels), HBM-bound: bytes per iteration / ms per iteration
    if pcmp is not None:
        it_ms = pcmp[a.solver]["ms"] / max(pcmp[a.solver]["iters"], 1)   # includes per-solve setup
    else:
        it_ms = (ph.get("poisson1", 0.0) + ph.get("poisson2", 0.0)) / max(sum(step_iters[-2:]), 1)
    mg_gbs = MG_BYTES_PER_CELL[a.solver] * ncell / (it_ms * 1e-3) / 1e9
    roof_mg = dict(bound="hbm", achieved=round(mg_gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                   frac=round(mg_gbs / HBM_PEAK_GBS, 4), traffic=None, kernel=f"{a.solver} CG iteration (20 launches: 8 level-0 V-cycle passes, 6 coarse-level legs, the coarsest solve, 2 CG updates, 3 finishes)",
                   time_ms=round(it_ms, 4))

    if a.copy_probe:
        # one plain copy of known size for the counter runs (outside every timed region)
        from edge_hip import _lib
        from edge_hip.field import stream_handle
        nfl = a.copy_probe * 1024 * 1024 // 4
        src = torch.ones(nfl, device="cuda")
        dst = torch.empty_like(src)
        _lib.call("edge_scale", src.data_ptr(), 1.0, dst.data_ptr(), nfl, stream_handle())
        torch.cuda.synchronize()
        del src, dst

    io_rec = None if a.no_io else io_timing(sim, torch)
    cpu, oracle_err = (None, None) if (a.no_cpu or a.case != "cfg2" or a.scheme != "edge") else cpu_baseline(sim)

    line = {
        "metric": "ms/step @256x128x128; Poisson ms+iters to 1e-6 rel, learned/mixed vs plain",
        "value": round(ms_step, 3),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": {"cfg2": "synthetic: four seeded vortex rings (Biot-Savart init on the GPU)",
                 "cfg3": "synthetic: uniform oblique inflow past a voxelised triangular plate",
                 "cfg3plate": "synthetic: four seeded vortex rings in a uniform inflow (speed 1, 20 degrees) past a "
                              "voxelised triangular plate of span 0.3",
                 "cfg5": "synthetic: one vortex ring, passive dye particles uniform in its torus (numpy "
                         "default_rng(0))",
                 "cfg5laden": "synthetic: one vortex ring, laden particles uniform in its torus (numpy "
                              "default_rng(0))"}[a.case],
        "config": {"workload": wl, "cells": ncell, "reinit_every": sim.reinit_every,
                   "poisson_stop": "||r|| <= 1e-6 ||b||",
                   "solver": SOLVER_NOTE[a.solver],
                   "projection_start": START_NOTE[sim.pressure_start],
                   **({"particles": sim.laden_particle_num} if sim.lp is not None else {}),
                   **({"dye_particles": sim.dye_num} if sim.dye is not None else {})},
        "timing": "value: the K timed steps back to back between two synchronisations; step_ms, phases_ms and "
                  "the kernels' launch times: K further steps, each synchronised, with HIP events",
        "phases_ms": {k: round(v, 3) for k, v in ph.items()},
        "step_ms": [round(w, 3) for w in walls],
        "poisson_iters": step_iters,
        "plain": plain,
        "poisson_in_step": poisson_in_step(ph, plain),
        "poisson_same_rhs": pcmp,
        "roofline": roof_march,
        "roofline_cg_iteration": roof_mg,
        "roofline_other_kernels": kernels,
        "io": io_rec,
        "cpu_baseline": cpu,
        "error_vs_oracle": oracle_err,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
