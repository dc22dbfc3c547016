#!/usr/bin/env python3
"""Benchmark of the SimPLIC alpha transport on one MI355X.

Metric: interface-cell updates/s = sum over steps of the mixed
cell count N_ic / sum over steps of (t_reconstruct + t_advect), both times
measured with HIP events on the library's stream.  A step is one
reconstruct() + advect() of the deformation test (applications/test/
plicVofAdvectionFoam) on the workload below; the unscaled velocity and flux
are resident in HBM, and each step rescales them on the GPU by the harness's
time factor before the timed hot path runs.

Default workload: 128^3 hex mesh of the unit cube with interior vertices
displaced by U(-0.2h, 0.2h) per axis (seed 1), sphere r=0.15 at
(0.35, 0.35, 0.35) initialised by adaptive exact-cut volumes, deformation
field with period 6 run to t = 3 at Courant 0.5, FP64, the reference test
case's alpha dictionary (LS normals, nAlphaBounds 3, snapTol 0, clip off).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
  --steps 0 (default) times the full run to t = T.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "interface-cell advect updates/s; plain vs learned normal at equal L1 error"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "hex128p": dict(kind="hex", N=128, perturb=0.2, seed=1, shapes=[((0.35, 0.35, 0.35), 0.15)],
                    flow="deformation", end_time=3.0,
                    desc="128^3 hex, interior vertices +-20% h (seed 1), sphere r=0.15 at 0.35, "
                         "deformation T=3, Co 0.5"),
    "hex64": dict(kind="hex", N=64, perturb=0.0, seed=0, shapes=[((0.35, 0.35, 0.35), 0.15)],
                  flow="deformation", end_time=3.0,
                  desc="64^3 uniform hex, sphere r=0.15 at 0.35, deformation T=3, Co 0.5"),
    # tetrahedral workload: 96^3 cubes x 6 Kuhn tets, vertices +-10% h, 8 seeded
    # spheres, one revolution of solid-body rotation about a seeded axis
    "tet96": dict(kind="tet", N=96, perturb=0.1, seed=2, shapes="spheres:7", flow="rotation:3",
                  end_time=1.0, desc="96^3 x 6 tets (5.3M cells), vertices +-10% h (seed 2), 8 seeded "
                                     "spheres r 0.05-0.12, one revolution of solid-body rotation, Co 0.5"),
}

ALPHA_DICT = dict(orientationMethod="LS", nAlphaBounds=3, snapTol=0.0, clip=False, mixedCellTol=1e-8)

# learned-normal comparison: 128^3 uniform hex, 3 seeded ellipsoids, same
# deformation; the network is trained on the GPU at the start of the run
LEARNED_WORKLOAD = dict(kind="hex", N=128, perturb=0.0, seed=0, shapes_seed=2024, flow="deformation",
                        end_time=3.0, desc="128^3 uniform hex, 3 seeded ellipsoids (semi-axes 0.08-0.15), "
                                           "deformation T=3, Co 0.5")
# finer plain meshes of the learned-normal case, to find the plain path's
# cost at the learned path's error
EQUAL_ERROR_MESHES = (144, 160)
LEARNED_TRAINING = dict(samples=4 << 20, depth=4, epochs=20, batch=4096, lr0=3e-3, lr1=3e-5, seed=11)


def algorithmic_bytes(stage, n_cells, n_faces, n_ic, cell_kind):
    """Algorithmic HBM bytes of one launch of a stage (FP64 values, int32 labels).

    detect      : alpha read per cell, mixed label written per interface cell
    normal      : per interface cell the 27 stencil alphas and centres, normal out
    fsd         : per interface cell its points, face->point labels, alpha, n in; d, C, S out
    upwind      : phi, owner, neighbour read and dVf written per face; alpha read per cell
    iflux       : per interface cell points, face labels, point velocities, n, d, C, and
                  phi in / dVf out on its downwind faces
    update      : alpha_old, V read and alpha written per cell, dVf read per face
    """
    npts, nfc, fpts = (8, 6, 4) if cell_kind == "hex" else (4, 4, 3)
    per_ic_fsd = npts * 24 + nfc * fpts * 4 + 8 + 24 + 8 + 24 + 24
    per_ic_iflux = npts * 24 + nfc * fpts * 4 + npts * 24 + 24 + 8 + 24 + (nfc // 2) * 16
    table = {
        "bounding": None,   # per step from the pass statistics: bounding_bytes()
        "detect": 8 * n_cells + 4 * n_ic,
        "normal": n_ic * (27 * 32 + 24),
        "findSignedDistance": n_ic * per_ic_fsd,
        "upwind": 24 * n_faces + 8 * n_cells,
        "interfaceFlux": n_ic * per_ic_iflux,
        "update": 24 * n_cells + 8 * n_faces,
    }
    return table.get(stage)


def bounding_bytes(n_cells, totals, cell_kind):
    """Algorithmic bytes of the bounding passes summed in `totals` (running
    sums: candidates, active, rounds, corrected faces, passes): per pass the
    candidate bitmap (N_c/8); per candidate its label, alpha, alpha_old, V
    and, per face, the face label, owner, phi and dVf; per corrected face the
    correction written and dVf updated, and alpha and V of the cell it
    corrects."""
    cand, _act, _rounds, corr, passes = totals
    nfc = 6 if cell_kind == "hex" else 4
    per_cand = 4 + 24 + nfc * (4 + 4 + 8 + 8)
    return passes * (n_cells // 8) + per_cand * cand + 48 * corr


def step_bytes(n_cells, n_faces, n_ic, cell_kind):
    """Whole-step algorithmic bytes 24 N_c + 24 N_f + B_ic N_ic (B_ic 512 hex / 272 tet)."""
    return 24 * n_cells + 24 * n_faces + (512 if cell_kind == "hex" else 272) * n_ic


# kernels of each pipeline stage, for the HBM traffic read from profiler counters
STAGE_KERNELS = {
    "detect": ("k_detect_list",),
    "normal": ("k_ls_normal", "k_grad_normal", "k_rdf_"),
    "findSignedDistance": ("k_fsd_group", "k_find_signed_distance"),
    "upwind": ("k_upwind",),
    "interfaceFlux": ("k_interface_flux", "k_un0", "k_flux_list", "k_flux_fan"),
    "update": ("k_update",),
    "bounding": ("k_bnd_",),
    "snapClip": ("k_snap_clip",),
}
COPY_PROBE_BYTES = 256 << 20


def copy_probe():
    """One device-to-device copy of COPY_PROBE_BYTES: the reference kernel
    whose FETCH_SIZE / WRITE_SIZE readings calibrate the counters (through
    the HIP runtime directly, so the probe does not depend on torch)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), COPY_PROBE_BYTES) == 0 and hip.hipMalloc(C.byref(b), COPY_PROBE_BYTES) == 0
    assert hip.hipMemset(a, 1, COPY_PROBE_BYTES) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, COPY_PROBE_BYTES, 3) == 0    # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"copied {COPY_PROBE_BYTES} bytes", flush=True)


def _counter_sums(run_dir):
    """kernel name -> [dispatches, counter sum (KiB)] of a rocprofv3 --pmc run."""
    import csv
    import glob
    sums = {}
    for path in glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for row in csv.DictReader(fh):
                name = row["Kernel_Name"].split("(")[0].replace("void ", "")
                e = sums.setdefault(name, [0, 0.0])
                e[0] += 1
                e[1] += float(row["Counter_Value"])
    return sums


def pmc_summary(copy_fetch_dir, copy_write_dir, run_fetch_dir, run_write_dir):
    """HBM bytes per launch of every stage from FETCH_SIZE and WRITE_SIZE runs,
    each scaled by what the same counter reads for the copy probe (a copy of
    COPY_PROBE_BYTES reads and writes exactly that many bytes)."""
    def probe_kib(d):
        sums = _counter_sums(d)
        name, (n, v) = max(sums.items(), key=lambda kv: kv[1][1] / max(kv[1][0], 1))
        return v / n
    f_fac = COPY_PROBE_BYTES / (1024.0 * probe_kib(copy_fetch_dir))
    w_fac = COPY_PROBE_BYTES / (1024.0 * probe_kib(copy_write_dir))
    fetch, write = _counter_sums(run_fetch_dir), _counter_sums(run_write_dir)
    # one update launch per step (k_update, or k_update<MF> by cell size)
    steps = max(v[0] for k, v in fetch.items() if re.search(r"k_update(<\d+>)?$", k))
    out = {"fetch_factor": f_fac, "write_factor": w_fac, "steps": steps, "stages": {}, "kernels": {}}
    for k in sorted(set(fetch) | set(write)):
        fb = 1024.0 * f_fac * fetch.get(k, [0, 0.0])[1] / steps
        wb = 1024.0 * w_fac * write.get(k, [0, 0.0])[1] / steps
        out["kernels"][k] = {"fetch_bytes_per_step": fb, "write_bytes_per_step": wb,
                             "dispatches_per_step": fetch.get(k, [0, 0.0])[0] / steps}
        for stage, pats in STAGE_KERNELS.items():
            if any(p in k for p in pats):
                e = out["stages"].setdefault(stage, {"fetch_bytes": 0.0, "write_bytes": 0.0})
                e["fetch_bytes"] += fb
                e["write_bytes"] += wb
    for e in out["stages"].values():
        e["bytes"] = e["fetch_bytes"] + e["write_bytes"]
    return out


def build_case(wl, alpha_dict=ALPHA_DICT):
    import geometricvofext_amd as g
    if wl["kind"] == "hex":
        mesh = g.hex_box(wl["N"], perturb=wl["perturb"], seed=wl["seed"])
    else:
        mesh = g.tet_box(wl["N"], perturb=wl["perturb"], seed=wl["seed"])
    from geometricvofext_amd import cases
    solver = g.SolveVofEqu(mesh, dict(alpha_dict))
    shapes = wl["shapes"]
    if isinstance(shapes, str):
        shapes = cases.random_spheres(int(shapes.split(":")[1]))
    solver.init_alpha(shapes, max_depth=5 if wl["kind"] == "hex" else 4)
    alpha0 = solver.alpha()
    if wl["flow"].startswith("rotation"):
        omega = cases.rotation_omega(int(wl["flow"].split(":")[1]), wl["end_time"])
        U0, phi0 = solver.rotation_field(omega, (0.5, 0.5, 0.5))
    else:
        U0, phi0 = solver.deformation_field()
    solver.upload_base_flow(U0, phi0)
    return mesh, solver, alpha0, U0, phi0


def traffic_of(workload, stage):
    """HBM bytes per launch of `stage` measured with profiler counters on this
    workload (profiles/hbm_traffic_<workload>.json, written by --pmc-summary),
    or None when no such measurement is committed."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    try:
        with open(path) as fh:
            return json.load(fh)["stages"][stage]["bytes"]
    except (OSError, KeyError, ValueError):
        return None


def flow_period(wl):
    """Period of the harness time factor: 2T for the reversing deformation, none for rotation."""
    return 0.0 if wl["flow"].startswith("rotation") else 2.0 * wl["end_time"]


def learned_comparison(say):
    """Plain least-squares normal against the learned normal on the
    ellipsoid workload: L1 shape error at t=T and interface-cell updates/s of
    the same timed loop, plus the network's training cost."""
    import geometricvofext_amd as g
    from geometricvofext_amd import cases, harness
    tr = LEARNED_TRAINING
    model = g.NormalModel(seed=tr["seed"])
    t0 = time.time()
    model.generate(tr["samples"], seed=tr["seed"] + 12, depth=tr["depth"])
    gen_s = time.time() - t0
    losses, train_s = model.train(epochs=tr["epochs"], batch=tr["batch"], lr0=tr["lr0"], lr1=tr["lr1"],
                                  seed=tr["seed"] + 1)
    say(f"learned normal: {tr['samples']} samples in {gen_s:.1f}s, {tr['epochs']} epochs in {train_s:.1f}s")
    wl = LEARNED_WORKLOAD
    mesh = g.hex_box(wl["N"])
    shapes = cases.random_ellipsoids(wl["shapes_seed"])
    out = {"workload": wl["desc"], "training": {"samples": tr["samples"], "epochs": tr["epochs"],
                                                "generate_s": gen_s, "train_device_s": train_s,
                                                "final_loss": float(losses[-1])}}
    seq = None
    # plain normals; learned normals with the default residual fallback; and
    # learned normals with the fallback test off (tolerance >= 1)
    variants = (("plain", "LS", None), ("learned", "learned", None), ("learned_no_fallback", "learned", 1.0))
    for name, method, tol in variants:
        d = dict(ALPHA_DICT, orientationMethod=method)
        if tol is not None:
            d["learnedResidualTol"] = tol
        s = g.SolveVofEqu(mesh, d)
        if method == "learned":
            s.set_normal_model(model)
        s.init_alpha(shapes, max_depth=5)
        a0 = s.alpha()
        U0, phi0 = s.deformation_field()
        s.upload_base_flow(U0, phi0)
        if seq is None:
            seq = time_sequence(s, a0, wl["end_time"])
        s.set_alpha(a0)
        device_sync(s)
        r0, ad0, n0 = s.timers()
        run_steps(s, seq)
        device_sync(s)
        r1, ad1, n1 = s.timers()
        hot = (r1 - r0) + (ad1 - ad0)
        V = s.geometry()["V"]
        a1 = s.alpha()
        row = {"updates_per_s": (n1 - n0) / hot, "ms_per_step": 1e3 * hot / len(seq),
               "E_s": harness.shape_error(a1, a0, V),
               "E_v": harness.volume_error(a1, V, float(np.sum(a0 * V))), "steps": len(seq)}
        if method == "learned":
            st = s.learned_stats()
            row["last_step_cells"] = {"interface": st[0], "learned": st[1], "residual_fallback": st[2],
                                      "least_squares_only": st[3]}
        if tol is not None:
            row["learnedResidualTol"] = tol
        out[name] = row
        say(f"{name}: E_s {row['E_s']:.5f}, {row['updates_per_s'] / 1e6:.2f}M updates/s")
        s.close()
    model.close()
    for name in ("learned", "learned_no_fallback"):
        out[name]["speedup_over_plain"] = out[name]["updates_per_s"] / out["plain"]["updates_per_s"]
        out[name]["E_s_ratio_to_plain"] = out[name]["E_s"] / out["plain"]["E_s"]
    out["equal_error"] = equal_error_comparison(out, shapes, wl, say)
    return out


BATCH_WORKLOAD = dict(n_cases=256, seed=2025, nmin=16, nmax=48, tmin=0.5, tmax=1.0,
                      desc="256 independent cases: uniform hex N^3, N ~ U{16..48} (seeded), 1-4 ellipsoids "
                           "(semi-axes 0.05-0.2, random orientation), rotation or deformation to a seeded "
                           "T in [0.5, 1], per-case Courant U(0.3, 0.5)")


def batched_cases_run(say, oracle_cases=4):
    """The batched workload run to every case's end time as one batch (one
    launch per stage for all cases, launch shapes from the tuner): aggregate
    interface-cell updates/s over the batch steps' GPU time, the same cases
    stepped one handle at a time over the first steps for comparison, per-case
    shape error at the end time, and the largest per-cell difference to the
    oracle of a few cases run in lockstep."""
    import geometricvofext_amd as g
    from geometricvofext_amd import cases
    from oracle import oracle as O
    wl = BATCH_WORKLOAD
    t0 = time.time()
    specs = cases.batch_cases(wl["n_cases"], seed=wl["seed"], nmin=wl["nmin"], nmax=wl["nmax"], tmin=wl["tmin"],
                              tmax=wl["tmax"])
    from concurrent.futures import ThreadPoolExecutor
    from geometricvofext_amd import harness
    with ThreadPoolExecutor(8) as pool:
        made = list(pool.map(lambda sp: cases.make_batch_case(sp, ALPHA_DICT), specs))
    n_cells = int(sum(sp["N"] ** 3 for sp in specs))
    say(f"batch: {len(specs)} cases, {n_cells} cells, set up in {time.time() - t0:.1f}s")
    # the same cases one handle at a time over their first steps, then back
    # to their initial state
    seq_steps = 20
    t_alone, nic_alone = 0.0, 0
    for c, sp in zip(made, specs):
        s = c["solver"]
        r0, a0_, n0 = s.timers()
        rt = harness.RunTime(end_time=sp["end_time"], max_co=sp["co"], max_alpha_co=sp["co"])
        for _ in range(seq_steps):
            co, aco = s.courant()
            rt.set_delta_t(0.5 * co * rt.delta_t, 0.5 * aco * rt.delta_t)
            rt.increment()
            s.scale_flow(harness.u_factor(rt.value, rt.delta_t, c["period"]))
            s.reconstruct()
            s.advect(rt.delta_t)
        r1, a1_, n1 = s.timers()
        t_alone += (r1 - r0) + (a1_ - a0_)
        nic_alone += n1 - n0
        s.set_alpha(c["alpha0"])
        s.scale_flow(0.0)
    say(f"batch: the cases one handle at a time, {seq_steps} steps each: {nic_alone / t_alone / 1e6:.1f}M updates/s")
    # oracle cases in lockstep: the smallest few
    small = sorted(range(len(specs)), key=lambda k: specs[k]["N"])[:oracle_cases]
    ors = {}
    for k in small:
        om = O.OracleMesh(made[k]["mesh"])
        o = O.OracleSolver(om, nAlphaBounds=ALPHA_DICT["nAlphaBounds"], snapTol=ALPHA_DICT["snapTol"],
                           clip=ALPHA_DICT["clip"], threads=16)
        o.set_alpha(made[k]["alpha0"])
        ors[k] = o
    worst = {k: 0.0 for k in small}
    b = g.CaseBatch([c["solver"] for c in made])
    shapes = None
    steps = 0
    sec0 = nic0 = None
    while any(c["rt"].run() for c in made):
        co, aco = b.courant()
        dt, fac = cases.batch_time_step(made, co, aco)
        b.scale_flow(fac)
        if shapes is None:
            shapes = b.tune(dt)
            sec0, nic0 = b.timers()
            say(f"batch: tuned launch shapes {shapes}")
        b.step(dt)
        steps += 1
        for k, o in ors.items():
            if dt[k] > 0.0:
                o.set_flow(made[k]["phi0"] * fac[k], made[k]["U0"] * fac[k])
                o.reconstruct()
                o.advect(dt[k])
                worst[k] = max(worst[k], float(np.max(np.abs(made[k]["solver"].alpha() - o.alpha()))))
        if steps % 250 == 0:
            say(f"batch: step {steps}, {sum(c['rt'].run() for c in made)} cases running")
    sec1, nic1 = b.timers()
    hot = sec1 - sec0
    b.close()
    es = []
    for c in made:
        s = c["solver"]
        V = s.geometry()["V"]
        es.append(harness.shape_error(s.alpha(), c["alpha0"], V))
    out = {"workload": wl["desc"], "cells": n_cells, "batch_steps": steps,
           "updates_per_s": (nic1 - nic0) / hot, "ms_per_batch_step": 1e3 * hot / steps,
           "n_ic_per_batch_step": (nic1 - nic0) / steps, "launch_shapes": shapes,
           "one_handle_at_a_time": {"updates_per_s": nic_alone / t_alone, "steps_per_case": seq_steps},
           "E_s_at_T": {"median": float(np.median(es)), "max": float(np.max(es))},
           "oracle_lockstep": {"cases": [int(specs[k]["N"]) for k in small],
                               "max_abs_alpha_diff": max(worst.values())}}
    out["speedup_over_one_at_a_time"] = out["updates_per_s"] / out["one_handle_at_a_time"]["updates_per_s"]
    say(f"batch: {steps} steps, {out['updates_per_s'] / 1e6:.1f}M updates/s, oracle lockstep max |diff| "
        f"{out['oracle_lockstep']['max_abs_alpha_diff']:.1e}")
    for c in made:
        c["solver"].close()
    return out


def equal_error_comparison(out, shapes, wl, say):
    """The plain path on finer meshes of the same case: the time it needs to
    reach the learned path's L1 shape error at t=T (hot-path seconds of the
    whole run, linear in E_s between the bracketing resolutions), against
    the learned path's time on the base mesh."""
    import geometricvofext_amd as g
    from geometricvofext_amd import harness
    pts = [(wl["N"], out["plain"]["E_s"], out["plain"]["ms_per_step"] * out["plain"]["steps"] / 1e3,
            out["plain"]["updates_per_s"])]
    for N in EQUAL_ERROR_MESHES:
        s = g.SolveVofEqu(g.hex_box(N), dict(ALPHA_DICT))
        s.init_alpha(shapes, max_depth=5)
        a0 = s.alpha()
        U0, phi0 = s.deformation_field()
        s.upload_base_flow(U0, phi0)
        seq = time_sequence(s, a0, wl["end_time"])
        s.set_alpha(a0)
        device_sync(s)
        r0, ad0, n0 = s.timers()
        run_steps(s, seq)
        device_sync(s)
        r1, ad1, n1 = s.timers()
        hot = (r1 - r0) + (ad1 - ad0)
        e = harness.shape_error(s.alpha(), a0, s.geometry()["V"])
        pts.append((N, e, hot, (n1 - n0) / hot))
        say(f"plain {N}^3: E_s {e:.5f} in {hot:.3f}s")
        s.close()
    res = {"plain_runs": [{"N": n, "E_s": e, "hot_s": t, "updates_per_s": u} for n, e, t, u in pts]}
    for name in ("learned", "learned_no_fallback"):
        target = out[name]["E_s"]
        t_learn = out[name]["ms_per_step"] * out[name]["steps"] / 1e3
        t_plain = None
        for (n0_, e0, t0, _), (n1_, e1, t1, _) in zip(pts, pts[1:]):
            if e0 >= target >= e1:
                t_plain = t0 + (t1 - t0) * (e0 - target) / (e0 - e1)
                break
        res[name] = {"E_s": target, "hot_s": t_learn, "plain_hot_s_at_equal_error": t_plain,
                     "time_to_solution_speedup": (t_plain / t_learn) if t_plain else None}
    return res


def time_sequence(solver, alpha0, end_time, max_steps=None, period=6.0):
    """(t, dt, uFactor) of every step to end_time.  With maxCo == maxAlphaCo
    the alpha-restricted Courant number never exceeds the full one, so the
    sequence does not depend on alpha and is computed from the flow alone."""
    from geometricvofext_amd import harness
    rt = harness.RunTime(end_time=end_time)
    assert rt.max_co == rt.max_alpha_co
    solver.set_alpha(alpha0)
    solver.scale_flow(0.0)
    seq = []
    while rt.run():
        co, aco = solver.courant()
        rt.set_delta_t(0.5 * co * rt.delta_t, 0.5 * aco * rt.delta_t)
        rt.increment()
        uf = harness.u_factor(rt.value, rt.delta_t, period)
        seq.append((rt.value, rt.delta_t, uf))
        solver.scale_flow(uf)
        if max_steps and len(seq) >= max_steps:
            break
    return seq


def host_io_rate(solver, alpha0, seq, U0, phi0):
    """Steps as a host-side solver drives them through SVP_HOST buffers: alpha,
    phi and U copied in, reconstruct + advect, alpha and alphaPhi copied out,
    every step.  Wall time with the copies included (PCIe-inclusive); not the
    benchmark's value, which has the fields resident in HBM."""
    solver.set_alpha(alpha0)
    a = np.array(alpha0, dtype=np.float64)
    n_ic = 0
    wall = 0.0
    device_sync(solver)
    for (_, dt, uf) in seq:
        phi, U = phi0 * uf, U0 * uf          # the host solver's own fields (not timed)
        t0 = time.perf_counter()
        solver.set_alpha(a)
        solver.set_flow(phi, U)
        solver.reconstruct()
        n_ic += solver.n_mixed()
        solver.advect(dt)
        a = solver.alpha()
        solver.alpha_phi()
        wall += time.perf_counter() - t0
    return {"steps": len(seq), "ms_per_step": 1e3 * wall / len(seq), "updates_per_s": n_ic / wall,
            "note": "host buffers: alpha, phi, U in and alpha, alphaPhi out over PCIe each step "
                    "(pageable numpy arrays); the host's own flow scaling is not timed"}


def run_steps(solver, seq):
    for (_, dt, uf) in seq:
        solver.scale_flow(uf)
        solver.reconstruct()
        solver.advect(dt)


def device_sync(solver):
    solver.synchronize()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def cpu_samples(solver, alpha0, seq, n_samples):
    """Start states for the CPU baseline: the GPU field at n_samples steps
    spread evenly over the whole run, each with that step's dt and uFactor, so
    the sample sees the interface sizes of the whole run and not only of its
    first steps."""
    idx = sorted({int(round(x)) for x in np.linspace(0, len(seq) - 1, n_samples)})
    solver.set_alpha(alpha0)
    out, k = [], 0
    for j in idx:
        run_steps(solver, seq[k:j])
        k = j
        out.append((j, solver.alpha(), seq[j][1], seq[j][2]))
    return out


def cpu_threads():
    """OpenMP threads of the CPU baseline: the CPU share the host gives this
    job (OMP_NUM_THREADS when set, else the cores this process may run on),
    with the machine's core count and model for the record."""
    try:
        avail = len(os.sched_getaffinity(0))
    except Exception:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else avail
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return min(threads, avail), {"nproc": os.cpu_count(), "affinity": avail, "model": model}


def cpu_baseline(mesh, samples, U0, phi0, threads, alpha_dict=ALPHA_DICT, om=None):
    """The oracle (CPU restatement of the reference path) on the same mesh:
    one reconstruct + advect from each sampled GPU state with that step's
    flow and dt, timed by the oracle's own reconstruct/advect timers."""
    from oracle import oracle as O
    om = om or O.OracleMesh(mesh)
    s = O.OracleSolver(om, mixedCellTol=alpha_dict["mixedCellTol"], nAlphaBounds=alpha_dict["nAlphaBounds"],
                       snapTol=alpha_dict["snapTol"], clip=alpha_dict["clip"], threads=threads,
                       orientationMethod=alpha_dict["orientationMethod"])
    n_ic = 0
    for k, (j, a, dt, uf) in enumerate(samples):
        s.set_alpha(a)
        s.set_flow(phi0 * uf, U0 * uf)
        s.reconstruct()
        n_ic += len(s.mixed()["cells"])
        s.advect(dt)
        print(f"[bench] cpu baseline ({threads} threads) sample {k + 1}/{len(samples)} (step {j})",
              file=sys.stderr, flush=True)
    rec, adv = s.timers()
    return n_ic, rec + adv, om


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (0 = run to t=T)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="hex128p", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-steps", type=int, default=8,
                    help="sampled steps of the CPU baseline, spread over the run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--stage", default="auto", help="stage timed for the roofline (auto = slowest)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-normal comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-cases run")
    ap.add_argument("--orientation", default="LS", choices=["LS", "alphaGrad", "RDF"],
                    help="interface normal method (default LS, the reference test case's)")
    ap.add_argument("--host-io", type=int, default=0, metavar="STEPS",
                    help="also time STEPS steps with host buffers (alpha, phi and U in, alpha and "
                         "alphaPhi out over PCIe every step, as a host-side solver would)")
    ap.add_argument("--copy-probe", action="store_true", help="run only the counter-calibration copy")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("COPY_FETCH", "COPY_WRITE", "RUN_FETCH", "RUN_WRITE"),
                    help="summarise rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of the copy probe and "
                         "of this benchmark into HBM bytes per stage launch (JSON)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.pmc_summary:
        print(json.dumps(pmc_summary(*args.pmc_summary), indent=1))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU; the alpha transport shards by spatial domain "
              "decomposition (halo exchange of alpha, fluxes and bounding corrections) but no multi-GPU "
              "path is built, so --gpus must be 1.",
              file=sys.stderr)
        sys.exit(2)

    import geometricvofext_amd as g
    wl = WORKLOADS[args.workload]
    t_setup = time.time()
    alpha_dict = dict(ALPHA_DICT, orientationMethod=args.orientation)
    mesh, solver, alpha0, U0, phi0 = build_case(wl, alpha_dict)
    seq = time_sequence(solver, alpha0, wl["end_time"], period=flow_period(wl))
    n_total = len(seq)
    K = args.steps if args.steps > 0 else n_total
    K = min(K, n_total)
    W = min(args.warmup, n_total)
    setup_s = time.time() - t_setup
    print(f"[bench] {args.workload}: {mesh.n_cells} cells, {mesh.n_faces} faces, {n_total} steps to "
          f"t={wl['end_time']}, setup {setup_s:.1f}s", file=sys.stderr, flush=True)

    def say(msg):
        print(f"[bench] {msg} ({time.time() - t_setup:.1f}s)", file=sys.stderr, flush=True)

    # warm-up: W steps from t=0, then rewind the field
    solver.set_alpha(alpha0)
    run_steps(solver, seq[:W])
    device_sync(solver)
    say("warm-up done")

    # stage breakdown (untimed pass over the first steps) to pick the roofline stage
    solver.set_alpha(alpha0)
    solver.profile_set(0xFF)
    run_steps(solver, seq[:min(60, K)])
    prof_all = solver.profile_get(reset=True)
    solver.profile_set(0)
    stages = {k: v for k, v in prof_all.items() if v[1] > 0}
    stage = args.stage
    if stage == "auto":
        # the slowest stage; bounding counts its bytes from the pass statistics
        stage = max(stages, key=lambda k: stages[k][0])

    say(f"stage breakdown done, roofline stage {stage}")
    # timed region
    solver.set_alpha(alpha0)
    stage_bit = 1 << g.SolveVofEqu.STAGES.index(stage)
    solver.profile_set(stage_bit)
    device_sync(solver)
    rec0, adv0, nic0 = solver.timers()
    bt0 = solver.bound_totals()
    t0 = time.perf_counter()
    run_steps(solver, seq[:K])
    device_sync(solver)
    wall = time.perf_counter() - t0
    rec1, adv1, nic1 = solver.timers()
    bt1 = solver.bound_totals()
    prof = solver.profile_get(reset=True)
    solver.profile_set(0)
    solver.synchronize()

    say(f"timed region done: {K} steps in {wall:.2f}s wall")
    n_ic = nic1 - nic0
    hot_s = (rec1 - rec0) + (adv1 - adv0)
    value = n_ic / hot_s
    ms_per_step = 1e3 * hot_s / K
    n_ic_mean = n_ic / K
    st_sec, st_cnt = prof[stage]
    st_avg = st_sec / max(st_cnt, 1)
    if stage == "bounding":
        dtot = [b - a for a, b in zip(bt0, bt1)]
        st_bytes = bounding_bytes(mesh.n_cells, dtot, wl["kind"]) / max(st_cnt, 1)
    else:
        st_bytes = algorithmic_bytes(stage, mesh.n_cells, mesh.n_faces, n_ic_mean, wl["kind"])
    achieved = st_bytes / st_avg / 1e9
    step_gbs = step_bytes(mesh.n_cells, mesh.n_faces, n_ic_mean, wl["kind"]) / (hot_s / K) / 1e9

    alpha_end = solver.alpha()
    V = solver.geometry()["V"]
    from geometricvofext_amd import harness
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "interface-cell updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl["desc"], "cells": mesh.n_cells, "faces": mesh.n_faces,
                   "alpha_dict": alpha_dict, "path": f"plain {args.orientation}"},
        "roofline": {"bound": "hbm", "stage": stage, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_of(args.workload, stage),
                     "algorithmic_bytes_per_launch": st_bytes, "avg_launch_us": 1e6 * st_avg},
        "step_roofline": {"achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_gbs / HBM_PEAK_GBS},
        "n_ic_mean": n_ic_mean,
        "wall_s_timed_region": wall,
        "stage_ms_per_step": {k: 1e3 * v[0] / max(v[1], 1) for k, v in stages.items()},
        "bounding_per_pass": dict(zip(("candidates", "active", "rounds", "corrected_faces"),
                                      [(b - a) / max(bt1[4] - bt0[4], 1) for a, b in zip(bt0[:4], bt1[:4])])),
        "t_end": seq[K - 1][0],
    }
    if K == n_total:
        res["accuracy"] = {"E_s_vs_t0": harness.shape_error(alpha_end, alpha0, V),
                           "E_v": harness.volume_error(alpha_end, V, float(np.sum(alpha0 * V)))}

    if args.host_io > 0:
        res["host_io"] = host_io_rate(solver, alpha0, seq[:args.host_io], U0, phi0)
        say("host-buffer steps done")

    if not args.no_batch:
        res["batched"] = batched_cases_run(say)

    if not args.no_learned:
        res["learned_vs_plain"] = learned_comparison(say)

    if not args.no_cpu_baseline:
        threads, host = cpu_threads()
        t_c = time.time()
        samples = cpu_samples(solver, alpha0, seq, args.cpu_steps)
        say(f"cpu baseline: {len(samples)} start states from the GPU run")
        nic_c, sec_c, om = cpu_baseline(mesh, samples, U0, phi0, threads, alpha_dict)
        steps_txt = ", ".join(str(j) for j, *_ in samples)
        res["cpu_baseline"] = {"value": nic_c / sec_c, "unit": "interface-cell updates/s", "cores": threads,
                               "kind": "port", "host": host,
                               "sample": f"oracle on the same mesh and flow, one reconstruct + advect from each "
                                         f"of {len(samples)} GPU states spread over the run (steps {steps_txt} "
                                         f"of {n_total}), {threads} OpenMP threads = the job's CPU share"}
        half = samples[::2]
        nic_1, sec_1, _ = cpu_baseline(mesh, half, U0, phi0, 1, alpha_dict, om)
        res["cpu_baseline_serial"] = {"value": nic_1 / sec_1, "unit": "interface-cell updates/s", "cores": 1,
                                      "kind": "port",
                                      "sample": f"the same, every second state ({len(half)} steps), 1 thread"}
        print(f"[bench] cpu baseline {time.time() - t_c:.1f}s", file=sys.stderr, flush=True)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
