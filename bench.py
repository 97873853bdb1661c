"""SDPD step benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--case suspension_10m] [--skin 0.2]

A step is one Particle::UpdatePar (neighbour build, forces, integration) of
the whole suspension, replayed from a HIP graph; by default the engine keeps
Verlet lists with a skin of 0.2 (re-sorted and rebuilt on the device when
the particles have moved far enough; --skin 0 re-sorts every step as the
reference does).  Inputs are generated on
the host from a seed and resident in HBM before the timed region.  Prints
one JSON line: whole-job particle-steps/s, the dominant kernel's HBM
roofline (HIP events on the engine stream), the step-level roofline, and
the FP64 OpenMP oracle timed on a bounded sample on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle-step of the whole step (five passes:
# integrate+key 56, sort 18.3, gather+predict 124, density 16,
# force+corrector 72) and per membrane vertex (440).
STEP_BYTES_PER_PARTICLE = 286.0
STEP_BYTES_PER_VERTEX = 440.0
# Algorithmic bytes per particle of the passes (neighbour data from cache),
# the five-pass dataflow above: integrate+wrap+key reads the FP64 position
# 24 B, velocity 16 B and force 12 B and writes the key 4 B; the counting
# sort and gather+predictor move 18.3 + 124 B; density reads the FP32
# position 12 B and writes the density 4 B; the SDPD force pass (EOS, walls,
# body force and corrector fused) reads position 12 B + predicted velocity
# 12 B + density 4 B + id 4 B + half-step velocity 12 B and writes
# velocity/density 16 B + force 12 B.
KERNEL_BYTES_PER_PARTICLE = {"integrate+key": 56.0, "sort+gather+predict": 142.3,
                             "density+wall velocity": 16.0, "sdpd force": 72.0}
KERNEL_NAMES = {"integrate+key": "k2_integrate", "sort+gather+predict": "k2_sortgather",
                "density+wall velocity": "k2_density", "sdpd force": "k2_force"}
# Verlet skin mode: no per-step re-sort.  The advance pass reads the FP64
# position 24 B, velocity 16 B and force 12 B and writes the position 24 B,
# its FP32 copy 16 B, the predicted velocity 16 B and the half-step velocity
# 12 B (120 B); density 16 B and force 72 B as above.  The list rebuild
# (re-sort and list build, every few tens of steps) is its own phase.
SKIN_STEP_BYTES_PER_PARTICLE = 208.0
SKIN_KERNEL_BYTES_PER_PARTICLE = {"advance+predict": 120.0, "density+wall velocity": 16.0, "sdpd force": 72.0}
SKIN_KERNEL_NAMES = {"advance+predict": "k3_advance", "density+wall velocity": "k3_density",
                     "sdpd force": "k2_force"}
DEFAULT_SKIN = 0.2
# autotuner family of each pass (its chosen shape names the launch variant)
KERNEL_FAMILY = {"integrate+key": "integrate", "advance+predict": "integrate",
                 "density+wall velocity": "density", "sdpd force": "sdpd force"}
# the bound of the neighbour-loop passes (instruction issue, not HBM)
KERNEL_BOUND_NOTE = {
    "density+wall velocity": "VALU issue in the 27-cube scan (~95 candidates per particle, lane imbalance)",
    "sdpd force": "VALU issue: Philox-4x32-10 (20 quarter-rate 64-bit products per pair) + pair math"}
SKIN_BOUND_NOTE = {
    "density+wall velocity": "scattered 16 B position gathers over the Verlet lists (~20 entries per particle)",
    "sdpd force": "VALU issue: Philox-4x32-10 (20 quarter-rate 64-bit products per pair) + pair math; "
                  "scattered list-entry loads through the neighbour bits",
    "advance+predict": "hbm"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def duct(c):
    """The duct's interior size from the case's box, and its size with the
    three frozen wall layers on each walled face (unit lattice)."""
    b = [float(v) for v in c.bbox]
    L, H, W = b[1] - b[0], b[3] - b[2], b[5] - b[4]
    return f"{L:.0f}x{H:.0f}x{W:.0f} duct ({H + 6:.0f}x{W + 6:.0f} with 3 frozen wall layers)"


def make_case(name):
    from gpucell_amd import casegen
    if name == "suspension_1m":
        c = casegen.suspension_1m()
        return c, (f"suspension_1m_100rbc: {duct(c)}, {c.nCell} RBCs x 642 vertices at "
                   f"{100 * c.hematocrit:.0f}% hematocrit, Poiseuille body force, aggregation off")
    if name == "stenosis_2m":
        c = casegen.stenosis_2m()
        return c, (f"stenosis_2m_300rbc: {duct(c)} with a 50% cosine stenosis, {c.nCell} RBCs at "
                   f"{100 * c.hematocrit:.0f}% hematocrit (aligned lattice slots), Morse De=0.3kT, "
                   "r0=0.3, beta=1.5")
    if name == "suspension_10m":
        c = casegen.suspension_10m()
        return c, (f"suspension_10m_2000rbc: {duct(c)}, {c.nCell} RBCs x 642 vertices at "
                   f"{100 * c.hematocrit:.0f}% hematocrit, Poiseuille body force, aggregation off")
    if name == "box32":
        return casegen.periodic_box(), "periodic 32^3 box, one RBC, reverse Poiseuille"
    if name == "ensemble64":
        return casegen.ensemble64(), ("ensemble64: 64 independent sheared suspensions (plane Couette "
                                      "cells, x and z periodic, walls moving at -U/+U), 27k-406k "
                                      "particles (log-uniform), 1-40 RBCs each, per-replica shear rate "
                                      "(0.002-0.01), Morse De and seed, one batched engine")
    raise SystemExit(f"unknown case {name}")


def cpu_model():
    """The host CPU's model string (/proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, seconds_target=15.0):
    """FP64 OpenMP oracle on the same workload: as many whole steps as fit
    in ~seconds_target (at least one), from the same initial state."""
    from oracle import pyoracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    o = orc.Oracle(case, threads=threads)
    st = o.initial_state()
    t0 = time.perf_counter()
    o.step(st, 1)
    t1 = time.perf_counter()
    per = t1 - t0
    n = max(0, int(seconds_target / max(per, 1e-6)) - 1)
    n = min(n, 50)
    t2 = time.perf_counter()
    for _ in range(n):
        o.step(st, 1)
    t3 = time.perf_counter()
    steps = n if n > 0 else 1
    dt = (t3 - t2) if n > 0 else per
    return {"value": case.N * steps / dt, "unit": "particle-steps/s", "cores": threads,
            "cpu_model": cpu_model(), "kind": "port",
            "sample": f"{steps} full steps of the same {case.N}-particle case (FP64 oracle, "
                      f"{threads} OpenMP threads, after 1 untimed step)"}


SURROGATE_THRESHOLD = 2e-5   # residual of the Verlet relation, in rest edge lengths
SURROGATE_EPOCHS = 32


def surrogate_leg(case, case_name, skin=0.0):
    """Learned membrane substepper applied to this suspension: ~1e6
    training rows from plain 8-substep trajectories of the 32^3 one-RBC box
    and of the 1M-particle 100-RBC channel, training with the library's
    BF16-MFMA kernels, then surrogate vs plain substepping from the same
    state (wall time per step at equal dt, deformation index and cell volume
    against plain)."""
    from gpucell_amd import casegen, substepper
    import torch
    t0 = time.perf_counter()
    box = casegen.periodic_box()
    src = case if case_name == "suspension_1m" else casegen.suspension_1m()
    # rows from relaxed suspensions, as the comparison below starts from
    # one (substepper.compare, relax)
    X1, Y1 = substepper.collect([box], steps=150, every=1, warmup=150)
    X2, Y2 = substepper.collect([src], steps=56, every=4, warmup=150,
                                max_rows=14 * src.nCellPar)
    del src
    X, Y = torch.cat([X1, X2]), torch.cat([Y1, Y2])
    t_data = time.perf_counter() - t0
    log(f"[bench] surrogate: {X.shape[0]} training rows in {t_data:.1f}s")
    # 32 epochs: at 8 the surrogate drifted the cell volumes by 0.57 % over
    # the 40 compared steps of the 10M suspension (plain substepping 0.29 %),
    # at 24 and 64 by 0.29 % and 0.28 %
    net, norm, loss, t_train = substepper.train(X, Y, epochs=SURROGATE_EPOCHS, batch=4096, lr=2e-3)
    log(f"[bench] surrogate: trained in {t_train:.1f}s, loss {loss[0]:.4g} -> {loss[-1]:.4g}")
    r = substepper.compare(case, net, norm, SURROGATE_THRESHOLD, steps=40, skin=skin)
    log("[bench] surrogate vs plain: " + json.dumps(r))
    r.update({"metric": "simulated time per wall-second, surrogate / plain substepping "
                        "(k=8, equal dt)",
              "case": case_name, "threshold": SURROGATE_THRESHOLD,
              "train_rows": int(X.shape[0]), "train_epochs": SURROGATE_EPOCHS,
              "train_seconds": t_train, "data_seconds": t_data,
              "loss_first_epoch": float(loss[0]), "loss_last_epoch": float(loss[-1])})
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--case", default="suspension_10m",
                    help="suspension_10m (default: 10M particles + 2000 RBCs, the largest single-GPU "
                         "workload), suspension_1m, stenosis_2m, ensemble64 or box32")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--eager", action="store_true", help="launch kernels without a graph")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device copy three times (the reference kernel "
                         "for scaling FETCH_SIZE / WRITE_SIZE counters to bytes) and exit")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the learned-substepper leg (training + surrogate vs plain)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes (no autotuning before the timed run)")
    ap.add_argument("--skin", type=float, default=DEFAULT_SKIN,
                    help="Verlet skin of the engine's neighbour lists (default %(default)s; 0 = the "
                         "reference's per-step re-sort and cube scan)")
    ap.add_argument("--surrogate-only", action="store_true",
                    help="only the learned-substepper leg on --case (prints its JSON and exits; "
                         "for kernel traces of the two substep modes)")
    a = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path "
              "(a suspension runs on one GPU; across GPUs it would be sharded by slabs of cubes)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    if a.calibrate_copy:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copied {src.numel() * 4} bytes x 3")
        return
    from gpucell_amd.engine import Engine, Ensemble

    log(f"[bench] generating case {a.case}")
    t = time.perf_counter()
    case, workload = make_case(a.case)
    if a.surrogate_only:
        print(json.dumps(surrogate_leg(case, a.case, a.skin)), flush=True)
        return
    batch = isinstance(case, list)
    cases = case if batch else [case]
    N, Nm = sum(c.N for c in cases), sum(c.nCellPar for c in cases)
    log(f"[bench] case ready: {len(cases)} suspension(s), N={N} Nm={Nm} "
        f"({time.perf_counter() - t:.1f}s)")
    skin = a.skin
    eng = Ensemble(cases, skin=skin) if batch else Engine(case, skin=skin)
    tune_rows = None
    if not a.no_autotune:
        log("[bench] autotuning launch shapes")
        tune_rows = eng.autotune(steps=5)
        for r in tune_rows:
            log(f"[bench]   {r['family']:15s} {r['shape']:22s} {r['us']:9.2f} us{'  <-' if r['chosen'] else ''}")
    log("[bench] engine created; warmup")
    graph = not a.eager
    eng.step(a.warmup, graph=graph)
    eng.info()  # synchronises the engine stream
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.step(a.steps, graph=graph)
    eng.info()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / a.steps
    finite = True
    for r in range(len(cases)):
        S = eng.state(r) if batch else eng.state()
        finite = finite and bool(np.isfinite(S["pos"]).all() and np.isfinite(S["force"]).all())
    log(f"[bench] {a.steps} steps: {ms:.4f} ms/step, finite={finite}")
    nrb = eng.info().nrebuild
    prof = eng.profile(20)
    log("[bench] per-phase us: " + json.dumps(prof))
    kbytes = SKIN_KERNEL_BYTES_PER_PARTICLE if skin > 0 else KERNEL_BYTES_PER_PARTICLE
    knames = SKIN_KERNEL_NAMES if skin > 0 else KERNEL_NAMES
    knotes = SKIN_BOUND_NOTE if skin > 0 else KERNEL_BOUND_NOTE
    # dominant single kernel of the step (HIP events on the engine stream)
    dom = max((k for k in kbytes if k in prof), key=lambda k: prof.get(k, 0.0))
    dom_us = prof[dom]
    achieved = kbytes[dom] * N / (dom_us * 1e-6) / 1e9
    # measured HBM bytes per launch from the committed counter profile of this
    # case, used only when it was taken on the launch variant this run uses
    traffic, traffic_variant = None, None
    mode = "_skin" if skin > 0 else ""
    pmc = os.path.join(ROOT, "profiles", f"traffic_{knames[dom]}_{a.case}{mode}.json")
    used = None if tune_rows is None else {r["family"]: r["shape"] for r in tune_rows if r["chosen"]}
    if os.path.exists(pmc):
        tj = json.load(open(pmc))
        fam = KERNEL_FAMILY.get(dom)
        # only a profile of the same mode (skin) and launch variant
        if abs(float(tj.get("skin", 0.0)) - skin) < 1e-6 and (
                used is None or fam is None or used.get(fam) == tj.get("shape")):
            traffic = tj.get("bytes_per_launch")
            traffic_variant = tj.get("variant")
    # instruction-issue roofline of the dominant kernel: VALU instructions per
    # wave from the committed counter profile of this case (rocprofv3 SQ
    # counters), 4 cycles per wave64 instruction on each of 1024 SIMDs
    valu = None
    vp = os.path.join(ROOT, "profiles", f"valu_{a.case}{mode}.json")
    vdoc = json.load(open(vp)) if os.path.exists(vp) else {}
    if vdoc and abs(float(vdoc.get("skin", 0.0)) - skin) < 1e-6:
        vj = vdoc.get("kernels", {}).get(knames[dom])
        if vj:
            # waves the kernel launches on this case (the force pass runs one
            # lane per moving particle), as recorded with the counters
            waves = vj.get("waves", (N + 63) // 64)
            issue_us = vj["valu_per_wave"] * waves * 4 / (1024 * 2.4e9) * 1e6
            valu = {"kernel": knames[dom], "valu_per_wave": vj["valu_per_wave"],
                    "issue_bound_us": issue_us, "avg_us": dom_us, "frac": issue_us / dom_us,
                    "peak": "1024 SIMDs x 16 lanes x 2.4 GHz (wave64 VALU instruction = 4 cycles)",
                    "note": "quarter-rate instructions counted once: a lower bound"}
    step_bytes = (SKIN_STEP_BYTES_PER_PARTICLE if skin > 0 else STEP_BYTES_PER_PARTICLE) * N \
        + STEP_BYTES_PER_VERTEX * Nm
    step_gbs = step_bytes / (ms * 1e-3) / 1e9
    # the same time against the per-step scheme's dataflow (re-sort every step)
    ref_bytes = STEP_BYTES_PER_PARTICLE * N + STEP_BYTES_PER_VERTEX * Nm
    out = {
        "metric": "SDPD particle-steps/s on 1 GPU (% HBM roofline); surrogate vs plain speedup",
        "value": N * a.steps / (t1 - t0),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded lattice fluid, walls and biconcave RBCs)",
        "config": {"workload": workload, "N": N, "Nm": Nm, "suspensions": len(cases),
                   "h": float(cases[0].consts.CutRadius), "dt": float(cases[0].consts.TimeStep),
                   "graph": graph, "finite": finite, "skin": skin,
                   "list_rebuilds_per_step": (nrb / (a.warmup + a.steps)) if skin > 0 else 1.0},
        "roofline": {"bound": "hbm", "kernel": knames[dom], "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_variant": traffic_variant, "avg_us": dom_us,
                     "algorithmic_bytes_per_particle": kbytes[dom],
                     "note": knotes.get(dom)},
        "valu_roofline": valu,
        "step_roofline": {"bytes_per_step": step_bytes, "achieved": step_gbs,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                          "dataflow": "skin mode: advance 120 + density 16 + force 72 B per particle, "
                                      "440 B per membrane vertex" if skin > 0 else
                                      "per-step re-sort: 286 B per particle, 440 B per membrane vertex",
                          "per_step_dataflow_frac": ref_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "phases_us": prof,
        "passes": {k: {"kernel": knames[k], "avg_us": prof.get(k),
                       "algorithmic_bytes_per_particle": b,
                       "achieved_gbs": (b * N / (prof[k] * 1e-6) / 1e9) if prof.get(k) else None,
                       "frac_hbm": (b * N / (prof[k] * 1e-6) / 1e9 / HBM_PEAK_GBS) if prof.get(k) else None,
                       "bound": knotes.get(k, "hbm")}
                   for k, b in kbytes.items()},
        "tuning": None if tune_rows is None else {
            r["family"]: r["shape"] for r in tune_rows if r["chosen"]},
    }
    eng.close()
    if not a.no_surrogate and not batch and case.nCell > 0:
        out["surrogate"] = surrogate_leg(case, a.case, skin)
    if not a.no_cpu:
        log("[bench] CPU baseline (FP64 oracle)")
        out["cpu_baseline"] = cpu_baseline(max(cases, key=lambda c: c.N))
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
