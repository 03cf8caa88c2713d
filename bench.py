"""Benchmark of the particle step (one PICCloud::evolve() per step) on one MI355X.

Workload (default): two-species thermalisation, 32^3 all-cyclic hex box, 64 electrons +
64 singly charged ions (m_i = 1836 m_e) per cell = 4,194,304 parcels, Maxwellian 10 eV /
1 eV at n = 1e18 m^-3, Boris push, Nanbu binary Coulomb collisions with Takizuka-Abe
pairing (lnL = 10), no field solve, FP64 throughout.  Synthetic initial state generated on
the device.  Prints one JSON line.

Roofline: every kernel of the step is HBM-bound.  Algorithmic bytes per parcel:
  move kernel  : read + write of the parcel state (3 f64 barycentric + 3 U + i32 tet)
                 plus the i32 id that travels with it = 2 x 56 B = 112 B (FP64),
                 2 x 44 B = 88 B (mixed: U in f32)
  Coulomb      : read + write of U plus the id read = 24 + 24 + 4 = 52 B (FP64)
achieved = bytes per launch / average launch time (HIP events on the context stream).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s + %HBM peak; Poisson ms/step learned-guess vs plain PCG"


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hot_kernels_thermalisation32_fp64_traffic.md")


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of this workload (FP64, 32^3, 64+64 ppc), or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic" and f[1] == kernel:
                return int(f[2])
    except OSError:
        return None
    return None


def state_bytes(precision):
    u = 4 if precision == "mixed" else 8
    return 3 * 8 + 3 * u + 4 + 4


def cpu_baseline(ncell_sample=8, ppc=64, target_s=10.0):
    """The oracle (FP64 C restatement) on a bounded sample of the same workload:
    ncell_sample^3 cells of the same size, same ppc, dt and models.  Timed twice: on every
    host thread this process may use (OpenMP over the move's parcels and the Coulomb
    cells; `value`) and on one thread (`value_1core`).  Its results do not depend on the
    thread count."""
    from oracle import picref
    from picfoam_amd import cases
    mesh, species, models, ex = cases.thermalisation(ncell=ncell_sample, ppc=ppc)
    rm = picref.RefMesh(mesh)
    nc = rm.n_cells
    tc, _, _ = rm.tets()
    starts = np.searchsorted(tc, np.arange(nc))
    dt = ex["dt"]
    sp = [(species[s]["mass"], species[s]["charge"], species[s]["nEquivalentParticles"]) for s in range(2)]
    cl = np.zeros((8, 8)); cl[:2, :2] = 10.0
    zeros = np.zeros((nc, 3))
    Q0 = abs(sp[0][1] * sp[0][2])
    n_total = 2 * nc * ppc

    def run(threads):
        picref.lib().ref_set_threads(threads)
        rng = np.random.default_rng(1)
        parts = []
        for s, T in ((0, ex["Te_K"]), (1, ex["Ti_K"])):
            n = nc * ppc
            tet = (starts.repeat(ppc) + rng.integers(0, 12, n)).astype(np.int32)
            e = -np.log(1 - rng.random((n, 4)))
            y = np.ascontiguousarray((e / e.sum(1, keepdims=True))[:, 1:])
            U = rng.normal(size=(n, 3)) * np.sqrt(cases.KB * T / species[s]["mass"])
            parts.append([y, tet, U, np.arange(n, dtype=np.uint32)])
        steps = 0
        t0 = time.perf_counter()
        while True:
            for s in range(2):
                y, tet, U, ids = parts[s]
                y, U, tet, _, _ = rm.move("Boris", "CellAverage", y, U, tet, sp[s][1], sp[s][0], dt, zeros, None, zeros)
                order, cs = rm.sort_by_cell(tet)
                parts[s] = [y[order], tet[order], U[order], ids[order], cs]
            Us, _ = picref.coulomb(rm, sp, [parts[0][2], parts[1][2]], [parts[0][3], parts[1][3]],
                                   [parts[0][4], parts[1][4]], dt, seed=1, step=steps, coulomb_log=cl)
            for s in range(2):
                parts[s][2] = Us[s]
            rm.deposit("CellAverage", [(parts[s][0], parts[s][1], sp[s][1] * sp[s][2]) for s in range(2)], Q0)
            for s in range(2):
                parts[s] = parts[s][:4]
            steps += 1
            el = time.perf_counter() - t0
            if el >= target_s:
                return steps, el

    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS") or avail), avail))
    steps_n, el_n = run(threads)
    steps_1, el_1 = run(1)
    return {"value": n_total * steps_n / el_n, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "value_1core": n_total * steps_1 / el_1,
            "sample": f"FP64 oracle, {ncell_sample}^3 cells x {ppc}+{ppc} ppc ({n_total} parcels), same dt/models "
                      f"as the GPU workload: {steps_n} steps in {el_n:.1f} s on {threads} threads (OpenMP over "
                      f"move parcels and Coulomb cells), {steps_1} steps in {el_1:.1f} s on 1 thread"}


def copy_calibration(mib):
    """hipMemcpyDeviceToDevice of mib MiB, three times: the blit kernel reads and writes
    exactly mib MiB per dispatch.  Uses the HIP runtime libpicmi is linked against."""
    import ctypes
    from picfoam_amd import picmi
    picmi.load_library()
    hip = ctypes.CDLL("libamdhip64.so.7")
    n = ctypes.c_size_t(mib * (1 << 20))
    x, y = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (x, y):
        if hip.hipMalloc(ctypes.byref(p), n) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(x, 1, n) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(y, x, n, 3) != 0:
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(x)
    hip.hipFree(y)
    print(f"bench: copy calibration {mib} MiB x3 done", file=sys.stderr, flush=True)


def run_steps(args, precision, steps, warmup, profile, sync_steps=0):
    """warmup steps, then `steps` steps through pic_run (one host synchronisation at the end),
    timed with the host clock around the call.  With profile: first a pass of `steps` steps
    with HIP events around every stage (the stage breakdown; two event records per stage and
    step cost about 25 us per step in all), then the timed pass with events around the
    dominant stage only (move or Coulomb, the larger in the breakdown), whose launch time the
    roofline uses.  sync_steps > 0 adds that many pic_step calls (synchronous, one round trip
    each) for comparison.  Returns (parcels, seconds, breakdown, dominant stage, its
    (ms, launches) in the timed pass, dt, synchronous ms per step)."""
    from picfoam_amd import cases, picmi
    mesh, species, models, ex = cases.thermalisation(ncell=args.ncell, ppc=args.ppc, precision=precision)
    models["sortInterval"] = args.sort_interval or 1
    species[1]["sortInterval"] = args.ion_sort_interval
    cloud = picmi.Cloud(mesh, species, models)
    cloud.init_maxwellian(0, args.ppc, ex["Te_K"], id0=0, seed=11)
    cloud.init_maxwellian(1, args.ppc, ex["Ti_K"], id0=0, seed=12)
    dt = ex["dt"]
    n_total = cloud.count(0) + cloud.count(1)
    if warmup:
        cloud.run(dt, warmup)
    cloud.sync()
    stages, dom, dom_times = None, None, None
    if profile:
        cloud.profile(True)
        cloud.run(dt, steps)
        cloud.sync()
        stages = cloud.stage_times()
        dom = max(("move", "coulomb"), key=lambda k: stages[k][0])
        cloud.profile(True, stages=[dom])
    t0 = time.perf_counter()
    st = cloud.run(dt, steps)
    cloud.sync()
    t1 = time.perf_counter()
    if profile:
        dom_times = cloud.stage_times()[dom]
        cloud.profile(False)
    sync_ms = None
    if sync_steps:
        cloud.sync()
        t2 = time.perf_counter()
        for _ in range(sync_steps):
            cloud.step(dt)
        sync_ms = 1e3 * (time.perf_counter() - t2) / sync_steps
    if st.coulomb_collisions <= 0:
        raise RuntimeError("no Coulomb collisions happened in the timed steps")
    cloud.close()
    return n_total, t1 - t0, stages, dom, dom_times, dt, sync_ms


def poisson_benchmark(ncell, ppc, snaps, train_steps, evals, warm, loss_kind=0, guesses=("previous", "extrapolate")):
    """Poisson ms per step (right-hand side, initial guess, PCG to ||r||/||b|| <= 1e-8, gradient)
    on the field-solve workload, for four initial guesses: the previous phi (the reference's
    warm start), the learned guess on top of it, the linear extrapolation 2 phi(n-1) -
    phi(n-2), and the learned guess on top of that.  The network (3 message-passing layers,
    width 64, BF16 MFMA) is trained on the snapshots the same run records."""
    from picfoam_amd import cases, picmi
    out = {}
    for guess in guesses:
        mesh, species, models, ex = cases.poisson_plasma(ncell=ncell, ppc=ppc, guess=guess)
        c = picmi.Cloud(mesh, species, models, ex["bc"])
        c.set_field(picmi.FIELD_SPACE_CHARGE, np.full(ex["n_cells"], ex["space_charge"]))
        c.init_maxwellian(0, ppc, ex["T_K"], seed=21)
        dt = ex["dt"]
        for _ in range(warm):
            c.step(dt)
        c.gnn_init(layers=3, width=64, seed=5)
        c.gnn_loss(loss_kind)
        c.gnn_record(snaps)
        base = [c.step(dt).poisson for _ in range(snaps)]
        t0 = time.perf_counter()
        first, last = c.gnn_train(train_steps, lr=2e-3, seed=2)
        t_train = time.perf_counter() - t0
        c.gnn_record(0)
        c.gnn_use(True)
        learned = [c.step(dt).poisson for _ in range(evals)]
        c.gnn_use(False)
        after = [c.step(dt).poisson for _ in range(evals)]
        c.close()
        ms_b = float(np.mean([p.ms for p in base] + [p.ms for p in after]))
        it_b = float(np.mean([p.iterations for p in base] + [p.iterations for p in after]))
        ms_l = float(np.mean([p.ms for p in learned]))
        it_l = float(np.mean([p.iterations for p in learned]))
        out[guess] = {"ms": round(ms_b, 3), "iterations": round(it_b, 2)}
        out[guess + "+learned"] = {"ms": round(ms_l, 3), "iterations": round(it_l, 2),
                                   "train_loss_first_last": [round(first, 4), round(last, 4)],
                                   "train_s": round(t_train, 2)}
    if "previous" in out:
        out["learned_vs_previous_ms_ratio"] = round(out["previous+learned"]["ms"] / out["previous"]["ms"], 3)
    out["training_loss"] = ("A-norm error", "residual norm", "A-norm + residual")[loss_kind]
    out["config"] = (f"{ncell}^3 hex, {ncell ** 3 * ppc} electrons + neutralising background, phi = 0 on x walls, "
                     f"cyclic y/z, PCG+DIC FP64 to 1e-8; {snaps} snapshots, {train_steps} Adam steps, "
                     f"{evals} evaluated steps")
    return out


def timed(cloud, dt, steps, warmup):
    """warmup, then `steps` timed pic_run steps with events around every stage (about 25 us
    per step, under 1 % of these steps); (ms/step, stage ms/step, stats, species-0 parcels
    when the timed steps start).  The timed window stays right after the warm-up: the
    argon case's step grows as the field heats the electrons (move 4.7 -> 8.4 ms over 22
    steps: more tet crossings per step)."""
    cloud.run(dt, warmup)
    cloud.sync()
    n0 = cloud.count(0)
    cloud.profile(True)
    t0 = time.perf_counter()
    st = cloud.run(dt, steps)
    cloud.sync()
    el = time.perf_counter() - t0
    stages = cloud.stage_times()
    cloud.profile(False)
    return 1e3 * el / steps, {k: round(v[0] / steps, 4) for k, v in stages.items() if v[1]}, st, n0


def extra_workloads(names, steps, warmup, sort_interval=None):
    """The other full-size configurations, measured the same way (pic_run, device-resident
    inputs): the magnetised electron push on the jittered 50^3x6 tet lattice (50M electrons,
    Higuera-Cary, FP32 push / FP64 track) and the argon-like MCC on a 100^3 hex box (100M
    electrons, Nanbu e-n with ionization, FP32 push).  The parcels move into cell order every
    8 steps on the tet lattice (no collision model needs the order there; it only keeps the
    move's tet loads local) and every 2 steps on the argon case (index sort in between),
    unless sort_interval is given."""
    from picfoam_amd import cases, picmi
    out = {}
    if "c3" in names:
        print("bench: magnetised tets, 50^3 x 6 tet lattice, 50M electrons", file=sys.stderr, flush=True)
        mesh, species, models, ex = cases.magnetised_tets(n_lattice=50, n_total=50_000_000, precision="mixed")
        models["sortInterval"] = sort_interval or 8
        c = picmi.Cloud(mesh, species, models)
        c.set_field(picmi.FIELD_E, ex["E"])
        c.set_field(picmi.FIELD_B, ex["B"])
        c.init_maxwellian(0, ex["ppc"], ex["T_K"], seed=5)
        ms, stages, _, n = timed(c, ex["dt"], steps, warmup)
        c.close()
        out["magnetised_tets"] = {"particle_steps_per_s": n / (ms * 1e-3), "ms_per_step": round(ms, 3),
                                  "stage_ms_per_step": stages, "parcels": n, "cells": mesh["n_cells"],
                                  "precision": "FP32 push (Higuera-Cary), FP64 track, int64 deposit",
                                  "sort_interval": models["sortInterval"]}
    if "c4" in names:
        print("bench: argon MCC, 100^3 hex, 100M electrons", file=sys.stderr, flush=True)
        mesh, species, models, ex = cases.argon_mcc(ncell=100, ppc=100, precision="mixed")
        models["sortInterval"] = sort_interval or 2
        c = picmi.Cloud(mesh, species, models)
        e, el, exs, iz = ex["xs"]
        c.set_cross_section(0, e, el, 0.0)
        c.set_cross_section(1, e, exs, 11.5)
        c.set_cross_section(2, e, iz, 15.76)
        c.mcc_initialize(1000.0, 0.01)
        c.set_field(picmi.FIELD_E, ex["E"])
        c.init_maxwellian(0, ex["ppc"], ex["T_K"], seed=3)
        ms, stages, st, n = timed(c, ex["dt"], steps, warmup)
        n_end = c.count(0)
        c.close()
        out["argon_mcc"] = {"particle_steps_per_s": n / (ms * 1e-3), "ms_per_step": round(ms, 3),
                            "stage_ms_per_step": stages, "electrons_timed_start_end": [n, n_end],
                            "mcc_candidates_last_run": st.mcc_candidates, "ionizations_last_run": st.mcc_ionization,
                            "precision": "FP32 push (Boris), FP64 track, FP64 MCC",
                            "sort_interval": models["sortInterval"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ncell", type=int, default=32)
    ap.add_argument("--ppc", type=int, default=64)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--sort-interval", type=int, default=None,
                    help="move the parcels into cell order every k steps (index sort in between); default 1 "
                         "for the thermalisation electrons, 8 for the tet lattice, 2 for the argon case")
    ap.add_argument("--ion-sort-interval", type=int, default=8,
                    help="the same for the ions of the thermalisation workload alone (about 0.5 %% of "
                         "them change cell per step; the electrons, 30 %%, stay at --sort-interval)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision comparison run")
    ap.add_argument("--no-poisson", action="store_true", help="skip the Poisson initial-guess comparison")
    ap.add_argument("--poisson-ncell", type=int, default=100)
    ap.add_argument("--poisson-ppc", type=int, default=8)
    ap.add_argument("--poisson-loss", type=int, default=0, choices=[0, 1, 2],
                    help="training loss of the learned guess: 0 A-norm error, 1 residual norm, 2 both")
    ap.add_argument("--workloads", default="", metavar="LIST",
                    help="also measure other full-size configurations: c3 (magnetised tets, 50M electrons), "
                         "c4 (argon MCC, 100M electrons), comma-separated")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device three times (a plain copy of known size whose "
                         "FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic figures)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU domain decomposition is a design, not code)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
    print(f"bench: {args.ncell}^3 mesh, {2 * args.ncell ** 3 * args.ppc} parcels, {args.precision}", file=sys.stderr,
          flush=True)
    n_total, elapsed, stages, dom, dom_times, dt, sync_ms = run_steps(args, args.precision, args.steps,
                                                                      args.warmup, True, sync_steps=10)
    ms_per_step = 1e3 * elapsed / args.steps
    value = n_total * args.steps / elapsed
    # dominant kernel and its roofline
    sb = state_bytes(args.precision)
    move_bytes = n_total * 2 * sb                # every species in one launch
    coul_bytes = n_total * (2 * 3 * (4 if args.precision == "mixed" else 8) + 4)
    b = {"move": move_bytes, "coulomb": coul_bytes}[dom]
    t_ms = dom_times[0] / max(dom_times[1], 1)   # the dominant stage's launches in the timed pass
    achieved = b / (t_ms * 1e-3) / 1e9
    # with a fixed Coulomb logarithm, cells of up to 420 parcels per species on average run in
    # the grouped kernel (four cells per wave at 64 + 64), denser ones one per wave
    coul_kernel = "k_coulomb_group_nanbu" if args.ppc <= 420 else "k_coulomb_nanbu"
    kname = {"move": "k_move", "coulomb": coul_kernel}[dom]
    default_workload = args.ncell == 32 and args.ppc == 64 and args.precision == "fp64"
    traffic = profiled_traffic(kname) if default_workload else None
    roof = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_launch": b, "avg_launch_ms": round(t_ms, 4)}
    if dom == "coulomb":
        roof["limiter"] = ("VALU: FP64 scattering math, the Philox streams per collision and shuffle key, and the "
                           "shuffle ranks (SQ counters in profiles/); the HBM fraction is low by construction")
    stage_ms = {k: round(v[0] / args.steps, 4) for k, v in stages.items() if v[1]}
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32 push / f64 track",
        "data": "synthetic",
        "config": {"workload": f"two-species thermalisation, {args.ncell}^3 cyclic hex, {args.ppc}+{args.ppc} ppc "
                               f"({n_total} parcels), Boris + Nanbu/TakizukaAbe Coulomb (lnL 10), no field solve",
                   "parcels": n_total, "dt": dt, "precision": args.precision,
                   "sort_interval": {"electrons": args.sort_interval or 1,
                                     "ions": args.ion_sort_interval}},
        "roofline": roof,
        "stage_ms_per_step": stage_ms,
        "stage_timing": "stage_ms_per_step from a pass of the same length before the timed one, with events "
                        "around every stage; the timed pass has events around the dominant stage only",
        "ms_per_step_synchronous_api": sync_ms,
    }
    if not args.no_mixed and args.precision == "fp64":
        print("bench: mixed-precision run of the same workload", file=sys.stderr, flush=True)
        n2, el2, _, _, _, _, _ = run_steps(args, "mixed", args.steps, args.warmup, False)
        out["mixed_precision"] = {"value": n2 * args.steps / el2, "ms_per_step": 1e3 * el2 / args.steps,
                                  "speedup_vs_fp64": (n2 * args.steps / el2) / value}
    if not args.no_poisson:
        print("bench: Poisson initial-guess comparison", file=sys.stderr, flush=True)
        out["poisson"] = poisson_benchmark(args.poisson_ncell, args.poisson_ppc, snaps=48, train_steps=400, evals=24,
                                           warm=4, loss_kind=args.poisson_loss)
    if args.workloads:
        out["workloads"] = extra_workloads(args.workloads.split(","), steps=10, warmup=2,
                                           sort_interval=args.sort_interval)
    if not args.no_cpu_baseline:
        print("bench: timing the CPU oracle on a bounded sample", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(target_s=args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
