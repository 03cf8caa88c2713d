"""Benchmark of the particle step (one PICCloud::evolve() per step) on one MI355X.

Headline workload: electron-impact MCC in an argon-like gas, the largest single-GPU
configuration: all-cyclic 100^3 hex box of 1 cm (1M cells), 100 electrons per cell
(100M), Maxwellian at 2 eV, driven by a uniform E = 5e4 V/m (no field solve), a uniform
background of 1e21 m^-3 at 300 K, synthetic tabulated cross sections (elastic with a
Ramsauer dip, excitation from 11.5 eV, ionization from 15.76 eV), null-collision MCC with
Nanbu scattering and ionization, Boris push.  Mixed precision: FP32 push and velocities,
FP64 barycentric tracking and positions, FP64 MCC arithmetic, int64 deposit.  The
synthetic initial state is generated on the device.  The same workload in FP64 runs
beside it for the mixed / FP64 ratio.  Also measured, as sub-objects: the two-species
thermalisation (32^3, 64 + 64 parcels per cell, Nanbu/Takizuka-Abe Coulomb, FP64 and
mixed), the magnetised electron push on the jittered 50^3 x 6 tet lattice (50M electrons,
Higuera-Cary) and the Poisson solve at 100^3 (plain warm-started PCG, the learned initial
guess, FP32-inner iterative refinement).  Prints one JSON line.

Roofline: every particle stage is HBM-bound.  The unit of work is a particle-step; its
algorithmic bytes are the read and write of the persistent parcel state once: 2 x 44 B =
88 B in mixed precision (3 f64 barycentric, 3 f32 velocity, i32 tet, u32 id), 2 x 56 B =
112 B in FP64.  The move kernel (k_move: push, interpolation, tracking, cell count) reads
40 B of that state per parcel (the id is not read) and writes 40 B plus the 8 B
(cell, slot) pair the sort scatters from: 88 B per parcel in mixed precision, 120 B in
FP64.  achieved = those bytes per launch / the kernel's average launch time, measured with
HIP events on the context stream around the move stage of the timed steps.
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
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_move_argon_mcc_traffic.md")


def profiled_traffic(kernel, precision):
    """HBM bytes per launch of `kernel` on the argon MCC workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a plain copy), or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            f = line.split()
            if len(f) == 4 and f[0] == "traffic" and f[1] == kernel and f[2] == precision:
                return int(f[3])
    except OSError:
        return None
    return None


def move_bytes_per_parcel(precision):
    """k_move: state read (y 24, U 12 or 24, tet 4) + state write (the same) + place 8"""
    u = 12 if precision == "mixed" else 24
    return 2 * (24 + u + 4) + 8


def state_bytes(precision):
    """persistent parcel state: y 3 x f64, U 3 x f32/f64, tet i32, id u32"""
    return 24 + (12 if precision == "mixed" else 24) + 4 + 4


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


def argon_cloud(precision, ncell, ppc, sort_interval):
    from picfoam_amd import cases, picmi
    mesh, species, models, ex = cases.argon_mcc(ncell=ncell, ppc=ppc, precision=precision)
    models["sortInterval"] = sort_interval
    c = picmi.Cloud(mesh, species, models)
    e, el, exs, iz = ex["xs"]
    c.set_cross_section(0, e, el, 0.0)
    c.set_cross_section(1, e, exs, 11.5)
    c.set_cross_section(2, e, iz, 15.76)
    c.mcc_initialize(1000.0, 0.01)
    c.set_field(picmi.FIELD_E, ex["E"])
    c.init_maxwellian(0, ex["ppc"], ex["T_K"], seed=3)
    return c, ex


def run_argon(precision, steps, warmup, ncell, ppc, sort_interval, breakdown=True):
    """warmup steps, then `steps` steps through pic_run (one host synchronisation at the
    end) timed by the host clock, with HIP events around the move stage only (its average
    launch time for the roofline); then, with breakdown, the same window once more on a
    fresh cloud from the same seeds (the run is deterministic) with events around every
    stage, for the stage split.  The field heats the electrons, so later windows cross more
    tets per step than earlier ones: the split is taken over the same steps as the rate."""
    c, ex = argon_cloud(precision, ncell, ppc, sort_interval)
    dt = ex["dt"]
    if warmup:
        c.run(dt, warmup)
    c.sync()
    n0 = c.count(0)
    c.profile(True, stages=["move"])
    t0 = time.perf_counter()
    st = c.run(dt, steps)
    c.sync()
    el = time.perf_counter() - t0
    move_ms, move_n = c.stage_times()["move"]
    n1 = c.count(0)
    out = {"ms_per_step": 1e3 * el / steps, "electrons_start_end": [n0, n1],
           "particle_steps_per_s": 0.5 * (n0 + n1) * steps / el, "dt": dt,
           "move_launch_ms": move_ms / max(move_n, 1), "mcc_candidates": st.mcc_candidates,
           "ionizations": st.mcc_ionization}
    c.profile(False)
    c.close()
    if breakdown:
        c, _ = argon_cloud(precision, ncell, ppc, sort_interval)
        if warmup:
            c.run(dt, warmup)
        c.profile(True)
        t0 = time.perf_counter()
        c.run(dt, steps)
        c.sync()
        out["stage_window_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / steps
        out["stage_ms_per_step"] = {k: round(v[0] / steps, 4) for k, v in c.stage_times().items() if v[1]}
        c.profile(False)
        c.close()
    return out


def run_thermalisation(precision, steps, warmup, ncell=32, ppc=64, ion_sort_interval=8):
    """The two-species thermalisation: electrons into cell order every step, ions every 8
    (index sort between; about 0.5 % of the ions change cell per step)."""
    from picfoam_amd import cases, picmi
    mesh, species, models, ex = cases.thermalisation(ncell=ncell, ppc=ppc, precision=precision)
    models["sortInterval"] = 1
    species[1]["sortInterval"] = ion_sort_interval
    c = picmi.Cloud(mesh, species, models)
    c.init_maxwellian(0, ppc, ex["Te_K"], id0=0, seed=11)
    c.init_maxwellian(1, ppc, ex["Ti_K"], id0=0, seed=12)
    dt = ex["dt"]
    n = c.count(0) + c.count(1)
    c.run(dt, warmup)
    c.sync()
    t0 = time.perf_counter()
    st = c.run(dt, steps)
    c.sync()
    el = time.perf_counter() - t0
    if st.coulomb_collisions <= 0:
        raise RuntimeError("no Coulomb collisions happened in the timed steps")
    c.profile(True)
    c.run(dt, steps)
    c.sync()
    stages = {k: round(v[0] / steps, 4) for k, v in c.stage_times().items() if v[1]}
    c.profile(False)
    c.close()
    return {"particle_steps_per_s": n * steps / el, "ms_per_step": 1e3 * el / steps, "parcels": n,
            "stage_ms_per_step_next_window": stages, "precision": precision,
            "sort_interval": {"electrons": 1, "ions": ion_sort_interval}}


def run_magnetised_tets(steps, warmup, sort_interval=8):
    """Magnetised electron push on the jittered 50^3 x 6 tet lattice (750k cells), 50M
    electrons, Higuera-Cary FP32 push, FP64 tracking; cell order every 8 steps (no
    collision model needs it; it keeps the move's tet loads local)."""
    from picfoam_amd import cases, picmi
    mesh, species, models, ex = cases.magnetised_tets(n_lattice=50, n_total=50_000_000, precision="mixed")
    models["sortInterval"] = sort_interval
    c = picmi.Cloud(mesh, species, models)
    c.set_field(picmi.FIELD_E, ex["E"])
    c.set_field(picmi.FIELD_B, ex["B"])
    c.init_maxwellian(0, ex["ppc"], ex["T_K"], seed=5)
    dt = ex["dt"]
    c.run(dt, warmup)
    c.sync()
    n = c.count(0)
    c.profile(True)
    t0 = time.perf_counter()
    c.run(dt, steps)
    c.sync()
    el = time.perf_counter() - t0
    stages = {k: round(v[0] / steps, 4) for k, v in c.stage_times().items() if v[1]}
    c.profile(False)
    c.close()
    return {"particle_steps_per_s": n * steps / el, "ms_per_step": round(1e3 * el / steps, 3),
            "stage_ms_per_step": stages, "parcels": n, "cells": mesh["n_cells"],
            "precision": "FP32 push (Higuera-Cary), FP64 track, int64 deposit", "sort_interval": sort_interval}


def run_vw_field_solve(steps, warmup, ncell=100, ppc=8):
    """The field-solve plasma of the Poisson block (100^3 hex, 8M electrons over a
    neutralising background, phi = 0 on the x walls) with VolumeWeighting field weighting and
    charge distribution: the move kernel then deposits each parcel's barycentric shares on
    its cell and three points (LDS-privatised int64 sums), and the solve is followed by the
    cell-to-point interpolation of E."""
    from picfoam_amd import cases, picmi
    mesh, species, models, ex = cases.poisson_plasma(ncell=ncell, ppc=ppc)
    models["FieldWeigthing"] = "VolumeWeighting"
    models["ChargeDistribution"] = "VolumeWeighting"
    c = picmi.Cloud(mesh, species, models, ex["bc"])
    c.set_field(picmi.FIELD_SPACE_CHARGE, np.full(ex["n_cells"], ex["space_charge"]))
    c.init_maxwellian(0, ppc, ex["T_K"], seed=21)
    dt = ex["dt"]
    for _ in range(warmup):
        c.step(dt)
    n = c.count(0)
    c.profile(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        c.step(dt)
    el = time.perf_counter() - t0
    stages = {k: round(v[0] / steps, 4) for k, v in c.stage_times().items() if v[1]}
    c.profile(False)
    c.close()
    return {"ms_per_step": round(1e3 * el / steps, 3), "stage_ms_per_step": stages, "parcels": n,
            "particle_steps_per_s": n * steps / el,
            "note": "synchronous pic_step per step (the Poisson solve checks convergence on the host); move includes "
                    "the VolumeWeighting deposit"}


def run_walls(steps, warmup, ncell=100, ppc=8):
    """The field-solve plasma (100^3 hex, 8M electrons, phi = 0 on the x faces) twice: with
    the x faces open (parcels reaching them are deleted) and as MaxwellianThermal walls at
    300 K (reflected, tracked on within the step by the wall pass).  ms per step, the move
    stage (the move kernel, and the wall pass when there are walls), and the wall hits per
    step from the step statistics."""
    from picfoam_amd import cases, picmi
    out = {}
    for kind in ("open", "walls"):
        if kind == "open":
            mesh, species, models, ex = cases.poisson_plasma(ncell=ncell, ppc=ppc)
        else:
            mesh, species, models, ex = cases.wall_plasma(ncell=ncell, ppc=ppc)
        c = picmi.Cloud(mesh, species, models, ex["bc"])
        c.set_field(picmi.FIELD_SPACE_CHARGE, np.full(ex["n_cells"], ex["space_charge"]))
        c.init_maxwellian(0, ppc, ex["T_K"], seed=21)
        dt = ex["dt"]
        for _ in range(warmup):
            c.step(dt)
        n = c.count(0)
        c.profile(True)
        hits = deleted = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            st = c.step(dt)
            hits += st.n_wall_hits
            deleted += st.n_deleted
        el = time.perf_counter() - t0
        stages = {k: round(v[0] / steps, 4) for k, v in c.stage_times().items() if v[1]}
        c.profile(False)
        out[kind] = {"ms_per_step": round(1e3 * el / steps, 3), "stage_ms_per_step": stages, "parcels": n,
                     "parcels_end": c.count(0), "wall_hits_per_step": hits / steps,
                     "deleted_per_step": deleted / steps}
        c.close()
    out["note"] = ("synchronous pic_step per step; the move stage of the wall run holds the move kernel and the "
                   "wall pass")
    return out


def run_binary(steps, warmup, ncell=100):
    """Neutral DSMC (BinaryCollisionModel, HardSphere, sub-cells, Nanbu weight correction) on
    the neon-argon gas of the reference's thermalization_binaryCollision case: 100^3 cells of
    1.875 um, 26 Ne + 13 Ar parcels per cell (39M parcels), dt 1e-10 s.  The step is the
    neutral move (no push), the sort and the pair pass (timed as the collision stage)."""
    from picfoam_amd import cases, picmi
    mesh, species, models, ex = cases.binary_mixture(ncell=ncell)
    c = picmi.Cloud(mesh, species, models)
    for s in range(2):
        c.init_maxwellian(s, ex["ppc"][s], ex["T"][s], id0=0, seed=21 + s)
    c.binary_initialize(list(ex["T"]), list(ex["n"]))
    dt = ex["dt"]
    n = c.count(0) + c.count(1)
    c.run(dt, warmup)
    c.sync()
    t0 = time.perf_counter()
    st = c.run(dt, steps)
    c.sync()
    el = time.perf_counter() - t0
    if st.binary_collisions <= 0:
        raise RuntimeError("no binary collisions happened in the timed steps")
    c.profile(True)
    st2 = c.run(dt, steps)
    c.sync()
    stages = {k: round(v[0] / steps, 4) for k, v in c.stage_times().items() if v[1]}
    c.profile(False)
    c.close()
    return {"particle_steps_per_s": n * steps / el, "ms_per_step": 1e3 * el / steps, "parcels": n,
            "cells": ncell ** 3, "candidates_per_step": st.binary_candidates / steps,
            "collisions_per_step": st.binary_collisions / steps,
            "stage_ms_per_step_next_window": stages,
            "note": "the collision stage (\"mcc\") is the binary pair pass; no field solve"}


def poisson_benchmark(ncell, ppc, snaps, train_steps, evals, warm, loss_kind=0, guesses=("previous", "extrapolate")):
    """Poisson ms per step (right-hand side, initial guess, PCG to ||r||/||b|| <= 1e-8, gradient)
    on the field-solve workload, for four initial guesses: the previous phi (the reference's
    warm start), the learned guess on top of it, the linear extrapolation 2 phi(n-1) -
    phi(n-2), and the learned guess on top of that.  The network (3 message-passing layers,
    width 64, BF16 MFMA) is trained on the snapshots the same run records.  Then the same
    plasma with the FP32-inner, FP64-refined solver (previous phi)."""
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
    # FP32-inner CG with FP64 iterative refinement, same plasma, previous phi as the start
    mesh, species, models, ex = cases.poisson_plasma(ncell=ncell, ppc=ppc)
    models["poissonMode"] = "mixed"
    c = picmi.Cloud(mesh, species, models, ex["bc"])
    c.set_field(picmi.FIELD_SPACE_CHARGE, np.full(ex["n_cells"], ex["space_charge"]))
    c.init_maxwellian(0, ppc, ex["T_K"], seed=21)
    for _ in range(warm):
        c.step(ex["dt"])
    ir = [c.step(ex["dt"]).poisson for _ in range(evals)]
    c.close()
    out["previous_fp32_inner_fp64_refinement"] = {
        "ms": round(float(np.mean([p.ms for p in ir])), 3),
        "inner_iterations": round(float(np.mean([p.iterations for p in ir])), 2),
        "refinements": round(float(np.mean([p.refinements for p in ir])), 2),
        "rel_residual_max": float(max(p.rel_residual for p in ir))}
    if "previous" in out:
        out["fp32_ir_vs_fp64_ms_ratio"] = round(out["previous_fp32_inner_fp64_refinement"]["ms"] /
                                                out["previous"]["ms"], 3)
    out["config"] = (f"{ncell}^3 hex, {ncell ** 3 * ppc} electrons + neutralising background, phi = 0 on x walls, "
                     f"cyclic y/z, PCG+DIC to 1e-8; {snaps} snapshots, {train_steps} Adam steps, "
                     f"{evals} evaluated steps")
    return out


def cpu_baseline(ncell_sample=10, ppc=100, target_s=10.0):
    """The oracle (FP64 C restatement) on a bounded sample of the headline workload:
    ncell_sample^3 cells of the same size (1e-4 m), the same electrons per cell, dt, field,
    background and cross sections.  A step is the oracle's move (Boris, CellAverage field),
    sort by cell, MCC and CellAverage deposit, as in the GPU step.  Timed on every host
    thread this process may use (OpenMP over the move's parcels; the MCC's cell loop and the
    sort are serial) and on one thread; the results do not depend on the thread count."""
    from oracle import picref
    from picfoam_amd import cases
    L = 0.01 * ncell_sample / 100
    mesh, species, models, ex = cases.argon_mcc(ncell=ncell_sample, ppc=ppc, precision="fp64", L=L, growth=3.0)
    rm = picref.RefMesh(mesh)
    nc = rm.n_cells
    tc, _, _ = rm.tets()
    starts = np.searchsorted(tc, np.arange(nc))
    dt = ex["dt"]
    e, el, exs, iz = ex["xs"]
    bg = models["uniformBackground"]
    P = {"xs": [(e, el), (e, exs), (e, iz)], "massE": cases.ME, "massN": bg["mass"], "nN": bg["numberDensity"],
         "TN": bg["temperature"], "wE": 1.0, "wN": 1.0, "th_ex": 11.5, "th_iz": 15.76, "seed": 3, "step": 0}
    E = ex["E"]
    zeros = np.zeros((nc, 3))
    smax0 = float(np.max(np.sqrt(2 * e * cases.QE / cases.ME) * (el + exs + iz)))

    def run(threads):
        picref.lib().ref_set_threads(threads)
        rng = np.random.default_rng(1)
        n = nc * ppc
        tet = (starts.repeat(ppc) + rng.integers(0, 12, n)).astype(np.int32)
        ee = -np.log(1 - rng.random((n, 4)))
        y = np.ascontiguousarray((ee / ee.sum(1, keepdims=True))[:, 1:])
        U = rng.normal(size=(n, 3)) * np.sqrt(cases.KB * ex["T_K"] / cases.ME)
        ids = np.arange(n, dtype=np.uint32)
        yI, UI, tI, iI = np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int32), np.zeros(0, np.uint32)
        smax = np.full(nc, smax0)
        rem = np.zeros(nc)
        steps, psteps = 0, 0
        t0 = time.perf_counter()
        while True:
            y, U, tet, _, _ = rm.move("Boris", "CellAverage", y, U, tet, -cases.QE, cases.ME, dt, E, None, zeros)
            order, cs = rm.sort_by_cell(tet)
            psteps += len(tet)
            y, U, tet, ids = y[order], U[order], tet[order], ids[order]
            P["step"] = steps
            cap = int(len(tet) * 1.1) + 64
            out = picref.mcc(rm, P, dt, y, U, tet, ids, cs, cap, yI, UI, tI, iI, len(tI) + cap, smax, rem)
            y, U, tet, ids = out["yE"], out["UE"], out["tetE"], out["idE"]
            yI, UI, tI, iI = out["yI"], out["UI"], out["tetI"], out["idI"]
            smax, rem = out["smax"], out["rem"]
            rm.deposit("CellAverage", [(y, tet, -cases.QE), (yI, tI, cases.QE)], cases.QE)
            steps += 1
            el_ = time.perf_counter() - t0
            if el_ >= target_s:
                return steps, psteps, el_

    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS") or avail), avail))
    steps_n, ps_n, el_n = run(threads)
    steps_1, ps_1, el_1 = run(1)
    return {"value": ps_n / el_n, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "value_1core": ps_1 / el_1,
            "sample": f"FP64 oracle on {ncell_sample}^3 cells of the headline workload ({nc * ppc} electrons, same "
                      f"cell size, dt, field, background and cross sections): {steps_n} steps in {el_n:.1f} s on "
                      f"{threads} threads (OpenMP over the move's parcels; sort, MCC and deposit serial), "
                      f"{steps_1} steps in {el_1:.1f} s on 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ncell", type=int, default=100, help="cells per side of the argon MCC box (default 100)")
    ap.add_argument("--ppc", type=int, default=100, help="electrons per cell of the argon MCC workload")
    ap.add_argument("--sort-interval", type=int, default=2,
                    help="the argon electrons move into cell order every k steps (index sort in between)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 run of the headline workload")
    ap.add_argument("--no-thermalisation", action="store_true")
    ap.add_argument("--no-tets", action="store_true", help="skip the magnetised tet-lattice workload")
    ap.add_argument("--no-poisson", action="store_true", help="skip the Poisson initial-guess comparison")
    ap.add_argument("--no-walls", action="store_true", help="skip the open-faces / walls comparison")
    ap.add_argument("--no-binary", action="store_true", help="skip the neutral DSMC workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only-headline", action="store_true", help="the mixed headline run alone (for profiling)")
    ap.add_argument("--only-vw-field-solve", action="store_true",
                    help="the VolumeWeighting field-solve workload alone (for profiling the deposit)")
    ap.add_argument("--only", choices=["thermalisation", "magnetised-tets", "poisson", "walls", "binary"],
                    help="one sub-workload alone, printed as its JSON object (for profiling): the thermalisation "
                         "in --precision (default fp64), the magnetised tets, or the Poisson initial-guess comparison")
    ap.add_argument("--precision", default=None, choices=["mixed", "fp64"],
                    help="precision of the headline run (default mixed; the other one runs beside it unless "
                         "--no-fp64)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--poisson-ncell", type=int, default=100)
    ap.add_argument("--poisson-ppc", type=int, default=8)
    ap.add_argument("--poisson-loss", type=int, default=0, choices=[0, 1, 2],
                    help="training loss of the learned guess: 0 A-norm error, 1 residual norm, 2 both")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device three times (a plain copy of known size whose "
                         "FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic figures)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU domain decomposition is a design, not code)", file=sys.stderr)
        sys.exit(2)
    if args.only_headline:
        args.no_fp64 = args.no_thermalisation = args.no_tets = args.no_poisson = args.no_cpu_baseline = True
        args.no_walls = args.no_binary = True
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
    if args.only:
        if args.only == "thermalisation":
            r = run_thermalisation(args.precision or "fp64", args.steps, args.warmup)
        elif args.only == "magnetised-tets":
            r = run_magnetised_tets(steps=args.steps, warmup=args.warmup)
        elif args.only == "walls":
            r = run_walls(args.steps, args.warmup)
        elif args.only == "binary":
            r = run_binary(args.steps, args.warmup)
        else:
            r = poisson_benchmark(args.poisson_ncell, args.poisson_ppc, snaps=48, train_steps=400, evals=24, warm=4,
                                  loss_kind=args.poisson_loss)
        print(json.dumps({args.only: r}), flush=True)
        return
    if args.only_vw_field_solve:
        print(json.dumps({"field_solve_volume_weighting": run_vw_field_solve(steps=args.steps, warmup=args.warmup)}),
              flush=True)
        return

    prec = args.precision or "mixed"
    N = args.ncell ** 3 * args.ppc
    print(f"bench: argon MCC {args.ncell}^3 cells, {N} electrons, {prec}", file=sys.stderr, flush=True)
    head = run_argon(prec, args.steps, args.warmup, args.ncell, args.ppc, args.sort_interval,
                     breakdown=not args.only_headline)
    value = head["particle_steps_per_s"]
    bpp = move_bytes_per_parcel(prec)
    n_move = 0.5 * sum(head["electrons_start_end"])
    b = n_move * bpp
    t_ms = head["move_launch_ms"]
    achieved = b / (t_ms * 1e-3) / 1e9
    traffic = profiled_traffic("k_move", prec) if (args.ncell, args.ppc) == (100, 100) else None
    roof = {"bound": "hbm", "kernel": "k_move", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "bytes_per_launch": int(b),
            "bytes_per_parcel": bpp, "avg_launch_ms": round(t_ms, 4)}
    step_bytes = 2 * state_bytes(prec)
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 push / f64 track" if prec == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"argon-like electron-impact MCC, {args.ncell}^3 cyclic hex (1 cm box), "
                               f"{N} electrons Maxwellian 2 eV, E = 5e4 V/m, background 1e21 m^-3 at 300 K, "
                               f"Boris + NanbuElectronNeutral with ionization, no field solve",
                   "electrons": N, "dt": head["dt"], "precision": prec, "sort_interval": args.sort_interval},
        "roofline": roof,
        "step_hbm": {"bytes_per_particle_step": step_bytes,
                     "achieved_GBs": round(value * step_bytes / 1e9, 1),
                     "frac": round(value * step_bytes / 1e9 / HBM_PEAK_GBS, 4)},
        "electrons_start_end": head["electrons_start_end"],
        "stage_ms_per_step": head.get("stage_ms_per_step"),
        "stage_timing": "the timed window has HIP events around the move stage only; the stage split repeats the "
                        "same steps on a fresh cloud from the same seeds with events around every stage"
                        + (f" ({head['stage_window_ms_per_step']:.3f} ms/step with those events)"
                           if "stage_window_ms_per_step" in head else ""),
    }
    if not args.no_fp64:
        other = "fp64" if prec == "mixed" else "mixed"
        print(f"bench: the same workload in {other}", file=sys.stderr, flush=True)
        o = run_argon(other, args.steps, args.warmup, args.ncell, args.ppc, args.sort_interval, breakdown=False)
        out[other] = {"value": o["particle_steps_per_s"], "ms_per_step": o["ms_per_step"],
                      "move_launch_ms": round(o["move_launch_ms"], 4)}
        out["mixed_vs_fp64"] = (value / o["particle_steps_per_s"]) if prec == "mixed" else \
            (o["particle_steps_per_s"] / value)
    wl = {}
    if not args.no_thermalisation:
        print("bench: thermalisation 32^3, 64 + 64 ppc, fp64 and mixed", file=sys.stderr, flush=True)
        t64 = run_thermalisation("fp64", args.steps, args.warmup)
        tmx = run_thermalisation("mixed", args.steps, args.warmup)
        t64["mixed"] = {"particle_steps_per_s": tmx["particle_steps_per_s"], "ms_per_step": tmx["ms_per_step"],
                        "speedup_vs_fp64": tmx["particle_steps_per_s"] / t64["particle_steps_per_s"]}
        wl["thermalisation"] = t64
    if not args.no_tets:
        print("bench: magnetised tets, 50^3 x 6 tet lattice, 50M electrons", file=sys.stderr, flush=True)
        wl["magnetised_tets"] = run_magnetised_tets(steps=10, warmup=2)
    if not args.no_poisson:
        print("bench: field solve with VolumeWeighting, 100^3, 8M electrons", file=sys.stderr, flush=True)
        wl["field_solve_volume_weighting"] = run_vw_field_solve(steps=10, warmup=2)
    if not args.no_walls:
        print("bench: field solve with open x faces and with MaxwellianThermal walls, 100^3, 8M electrons",
              file=sys.stderr, flush=True)
        wl["walls"] = run_walls(steps=10, warmup=2)
    if not args.no_binary:
        print("bench: neutral DSMC, Ne/Ar hard spheres, 100^3 cells, 39M parcels", file=sys.stderr, flush=True)
        wl["binary_dsmc"] = run_binary(steps=10, warmup=2)
    if wl:
        out["workloads"] = wl
    if not args.no_poisson:
        print("bench: Poisson initial-guess comparison", file=sys.stderr, flush=True)
        out["poisson"] = poisson_benchmark(args.poisson_ncell, args.poisson_ppc, snaps=48, train_steps=400, evals=24,
                                           warm=4, loss_kind=args.poisson_loss)
    if not args.no_cpu_baseline:
        print("bench: timing the CPU oracle on a bounded sample", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(target_s=args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
