#!/usr/bin/env python3
"""Benchmark: SCF time-to-converge (|dE| < 1e-8 Eh) and Fock builds/s on one MI355X,
mixed precision + learned guess against the plain FP64 GPU path.

Workload (default): seeded 64-water cluster, 6-31G* (1216 Cartesian functions), RHF,
Pulay DIIS, converged to |dE| < 1e-8 Eh; synthetic geometry from the project's own
generator (oxygens uniform in a 12.4 A cube at 1 g/cm^3, O-O >= 2.5 A, Haar-random
orientations, seed 64001).  One step = one full SCF from the geometry: context build
(shell pairs, Schwarz bounds on the GPU), one-electron matrices, the guess and every
iteration.

  * timed steps (the metric): mixed precision (FP32 for quartets with Schwarz bound
    below 1e-6, incremental density-change builds, the last iterations in full FP64)
    with the learned density guess;
  * reference run (reported beside it): plain FP64, core-Hamiltonian guess, full builds.
    The final energies of the two must agree to 1e-8 Eh.

The learned guess loads unomol_amd/models/guess_631gs.bin; without it, a network is
trained first on small clusters (outside the timed region, reported).

Prints ONE compact JSON line (stdout) with the metric, the errors of the timed path
against the plain FP64 run (energy, orbital energies, density), an ablation (FP64 +
learned guess, mixed + core-H guess), a roofline object for the dominant ERI kernel of
the timed steps (the class/precision kernel with the largest summed time; per-launch
time from HIP events recorded around each launch, on the stream it runs on, inside the
timed region -- a build's launches run on three streams, so these spans include time
shared with other kernels; `one_stream` repeats the step outside the timed region with
every launch on one stream for the kernel's isolated duration; algorithmic FLOPs = unomol_amd.flopmodel's per-primitive-quartet and
per-shell-quartet operation counts of the kernel's own arithmetic times the quartets the
kernel evaluated, counted by the kernel; HBM traffic per launch from the committed
FETCH_SIZE/WRITE_SIZE passes when they cover the same kernel), the 16-water cluster's
plain FP64 GPU time beside the reference program's measured time on it, and a
cpu_baseline object (the reference's own code compiled from its sources into
oracle/_ref, timed on samples of the workload on one host core, extrapolated to the
SCF).  The per-class kernel table goes to stderr.
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PAIR_L = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]
PAIR_NAME = ["ss", "ps", "pp", "ds", "dp", "dd"]
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (and matrix) dense peak
FP32_PEAK_TFLOPS = 157.3
LANE_MAX_COMP = 162       # classes up to this many components run one quartet per lane
PROF_SLOTS = 432          # profiling slots: 72 kind + 36 precision + 6 X + Y
ROW_KINDS = {1: "true, true", 2: "false, true", 3: "false, false"}   # k_eri_rows <TL, FB>
GROUP_KINDS = {4: 2, 5: 4}                                            # k_eri_rows_mb <NBR>
MODEL = os.path.join(ROOT, "unomol_amd", "models", "guess_631gs.bin")


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def ncart(l):
    return (l + 1) * (l + 2) // 2


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(deck_path, fp64_iterations, budget_s=15.0):
    """The reference's CPU path, compiled from its sources into oracle/_ref/libref_harness.so,
    on one host core.  The 64-water workload cannot run through the reference whole (its
    integral store would hold 2.7e11 ERIs, 6.6 TB), so its SCF time is extrapolated from two
    measured samples:
      * its Rys ERI code on a random sample of the deck's canonical shell quartets (about
        budget_s seconds): ERI-pass time and the fraction of ERIs it would store;
      * its direct-store integral pass and Fock replay (formGmatrix) on a 4-water 6-31G*
        cluster: seconds per stored ERI per Fock build (the in-memory rate, no disk spill,
        which favours the reference).
    Estimate = ERI pass + (FP64 GPU iteration count) x replay of the estimated store."""
    lib_path = os.path.join(ROOT, "oracle", "_ref", "libref_harness.so")
    if not os.path.exists(lib_path):
        return None
    lib = C.CDLL(lib_path)
    secs, nf, tot, nk = C.c_double(), C.c_long(), C.c_double(), C.c_long()
    n = 2000
    lib.ref_eri_sample(deck_path.encode(), C.c_long(n), C.c_ulong(7), C.byref(secs), C.byref(nf), C.byref(tot),
                       C.byref(nk))
    if secs.value > 0:
        n = max(n, int(n * budget_s / max(secs.value, 1e-3)))
    lib.ref_eri_sample(deck_path.encode(), C.c_long(n), C.c_ulong(11), C.byref(secs), C.byref(nf), C.byref(tot),
                       C.byref(nk))
    rate = n / secs.value
    eri_pass = tot.value / rate
    stored = tot.value * nk.value / n
    import unomol_amd
    tmp = tempfile.mkdtemp()
    w4 = os.path.join(tmp, "w4.dat")
    unomol_amd.generate_deck(w4, "water_cluster", 4, 4001, "6-31G*", 0, 1e-8, 50)
    es, gs, ns_ = C.c_double(), C.c_double(), C.c_double()
    lib.ref_fock_sample(w4.encode(), os.path.join(tmp, "MINTS.DAT").encode(), 5, C.byref(es), C.byref(gs),
                        C.byref(ns_))
    per_eri = gs.value / ns_.value
    replay = stored * per_eri
    est = eri_pass + fp64_iterations * replay
    return {"value": round(est, 1), "unit": "s (reference SCF of this workload, extrapolated)", "cores": 1,
            "kind": "reference",
            "sample": "reference Rys ERI code on %d random canonical shell quartets of the deck (%.1f s): "
                      "%.3e shell quartets/s, ERI pass %.0f s, %.3e ERIs stored (|val| > 1e-14); reference "
                      "formGmatrix replay on a 4-water 6-31G* cluster (%.3e stored ERIs, %.1f ms per build): "
                      "%.2f ns per stored ERI, %.0f s per Fock build here; SCF = ERI pass + %d iterations "
                      "(the FP64 GPU run's count) of replay" %
                      (n, secs.value, rate, eri_pass, stored, ns_.value, 1e3 * gs.value, 1e9 * per_eri, replay,
                       fp64_iterations),
            "eri_pass_s": round(eri_pass, 1), "replay_s_per_build": round(replay, 2),
            "shell_quartets_per_s": round(rate, 1), "host_cores": host_cores()}


def reference_config_b():
    """The reference program itself on the 16-water 6-31G* cluster (three SCF iterations
    from the core-Hamiltonian guess, serial Unomol and 8-rank UnomolMPI), measured once by
    profiles/cpu_reference_config_b.md (its last line is JSON), or None."""
    path = os.path.join(ROOT, "profiles", "cpu_reference_config_b.md")
    if not os.path.exists(path):
        return None
    rec = None
    for line in open(path):
        if line.strip().startswith("{"):
            rec = json.loads(line)
    return rec


def run_config_b(unomol_amd):
    """Plain FP64 GPU path on the 16-water 6-31G* cluster (304 functions): context build,
    guess and three SCF iterations with the reference's damping, the same trajectory the
    reference runs in profiles/cpu_reference_config_b.md (energies equal to 1e-9 Eh in the
    parity tests).  Returns wall seconds of the best of three runs and the energy."""
    tmp = tempfile.mkdtemp()
    path = os.path.join(tmp, "b.dat")
    unomol_amd.generate_deck(path, "water_cluster", 16, 16001, "6-31G*", 0, 1e-10, 3)
    best, e = None, None
    for _ in range(3):
        t0 = time.perf_counter()
        b = unomol_amd.Basis(path)
        t = unomol_amd.TwoElectronInts(b)
        hf = unomol_amd.RestrictedHartreeFock(b, t)
        e = hf.findEnergy(maxits=3, eps=1e-10, cflag=2, accel=0, guess=0)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best, e


def read_profile(N, t):
    """Per-class kernel time (HIP events on the stream each launch runs on), launches and evaluated
    primitive / shell quartets accumulated since profiling was switched on."""
    ms, launches, prim, shell = (np.zeros(PROF_SLOTS) for _ in range(4))
    N.check(N.lib().uno_profile(t.ctx, -1, N.ptr(ms), N.ptr(launches), N.ptr(prim), N.ptr(shell)))
    N.check(N.lib().uno_profile(t.ctx, 0, None, None, None, None))
    return ms, launches, prim, shell


def profile_classes(N, t, P):
    """One profiled Fock build on density P."""
    N.check(N.lib().uno_profile(t.ctx, 1, None, None, None, None))
    G = np.zeros_like(P)
    t.formGmatrix(P, G)
    return class_table(*read_profile(N, t))


def class_table(ms, launches, prim, shell):
    """Per-class entries with the algorithmic FLOPs of unomol_amd.flopmodel (the kernel's
    own arithmetic per primitive quartet and per shell quartet, constant-folded, dead code
    removed) times the quartets the kernel counted."""
    from unomol_amd import flopmodel
    classes = []
    for kind in range(PROF_SLOTS // 72):
        for prec in range(2):
            for X in range(6):
                for Y in range(X + 1):
                    ci = 72 * kind + 36 * prec + 6 * X + Y
                    if launches[ci] == 0:
                        continue
                    la, lb = PAIR_L[X]
                    lc, ld = PAIR_L[Y]
                    fl = flopmodel.class_flops(la, lb, lc, ld, "f32" if prec else "f64", prim[ci], shell[ci])
                    cls = {"class": "(%s|%s)" % (PAIR_NAME[X], PAIR_NAME[Y]), "precision": "f32" if prec else "f64",
                           "kind": kind}
                    cls.update({"kernel": kernel_name(cls), "ms": ms[ci], "launches": launches[ci],
                                "prim_quartets": prim[ci], "shell_quartets": shell[ci], "gflop": fl / 1e9,
                                "tflops": fl / (ms[ci] * 1e-3) / 1e12 if ms[ci] > 0 else 0.0})
                    classes.append(cls)
    return classes


def kernel_name(cls):
    """Demangled name of the Fock-build kernel of a class entry (as rocprofv3 prints it):
    kind 0 the wave-task (lane classes) or per-quartet wave kernel, kinds 1-3 the
    row-merged kernel variants."""
    X, Y = PAIR_NAME.index(cls["class"][1:3]), PAIR_NAME.index(cls["class"][4:6])
    la, lb = PAIR_L[X]
    lc, ld = PAIR_L[Y]
    t = "float" if cls["precision"] == "f32" else "double"
    if cls.get("kind", 0) in GROUP_KINDS:
        return "void uno::k_eri_rows_mb<%d, %d, %d, %d, 0, %s, %d>(uno::EriArgs)" % (la, lb, lc, ld, t,
                                                                                     GROUP_KINDS[cls["kind"]])
    if cls.get("kind", 0):
        return "void uno::k_eri_rows<%d, %d, %d, %d, 0, %s, %s>(uno::EriArgs)" % (la, lb, lc, ld, t,
                                                                                  ROW_KINDS[cls["kind"]])
    lane = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld) <= LANE_MAX_COMP
    return "void uno::%s<%d, %d, %d, %d, 0, %s>(uno::EriArgs)" % ("k_eri_bra" if lane else "k_eri_wave", la, lb, lc,
                                                                  ld, t)


def traffic_of(kernel):
    """HBM bytes per launch of the kernel from the committed FETCH_SIZE / WRITE_SIZE passes
    (profiles/eri_traffic_w64_631gs.md, last line = JSON), or None if they cover another kernel."""
    path = os.path.join(ROOT, "profiles", "eri_traffic_w64_631gs.md")
    if not os.path.exists(path):
        return None
    rec = None
    for line in open(path):
        line = line.strip()
        if line.startswith("{"):
            rec = json.loads(line)
    if not rec or rec.get("kernel") != kernel:
        return None
    return rec.get("bytes_per_launch")


def roofline_of(classes):
    """Roofline entry for the dominant ERI kernel (largest summed time) of the classes."""
    dom = max(classes, key=lambda c: c["ms"])
    peak = FP32_PEAK_TFLOPS if dom["precision"] == "f32" else FP64_PEAK_TFLOPS
    tot_fl = sum(c["gflop"] for c in classes) * 1e9
    tot_ms = sum(c["ms"] for c in classes)
    name = kernel_name(dom)
    return {"bound": "valu", "pipe": "%s VALU (no dense contraction); MI355X's published dense %s peak, "
                                     "%.1f TFLOP/s, is the bound" % (dom["precision"].upper(), dom["precision"].upper(), peak),
            "kernel": name, "class": dom["class"], "precision": dom["precision"],
            "launches": int(dom["launches"]), "avg_launch_ms": round(dom["ms"] / max(dom["launches"], 1), 4),
            "achieved": round(dom["tflops"], 4), "peak": peak, "unit": "TFLOP/s",
            "frac": round(dom["tflops"] / peak, 5), "traffic": traffic_of(name),
            "all_eri_kernels": {"ms": round(tot_ms, 3), "tflops": round(tot_fl / (tot_ms * 1e-3) / 1e12, 4)}}


def get_model(unomol_amd):
    if os.path.exists(MODEL):
        return unomol_amd.MLP(path=MODEL), {"source": "unomol_amd/models/guess_631gs.bin"}
    from unomol_amd import guess_train as GT
    t0 = time.perf_counter()
    X, Y, M = GT.make_training_set(n_clusters=60, min_w=2, max_w=8, seed=1000, log=log)
    net, loss = GT.train_guess(X, Y, M, steps=2000)
    dt = time.perf_counter() - t0
    log("trained a guess network in %.1f s (loss %.2e)" % (dt, loss))
    return net, {"source": "trained in this run (60 clusters of 2-8 waters)", "train_s": round(dt, 2),
                 "loss": loss}


def run_batched(args):
    """Batched SCF: all molecules iterate in lockstep on the device (one ERI launch per class
    pair over every molecule, the whole Fock build captured once as a HIP graph and replayed
    with one host launch per build, grouped GEMMs, batched Jacobi eigensolver).  Timed:
    building the batch (every molecule's pairs and Schwarz bounds) and the SCF to |dE| < 1e-8.
    Beside it: the single-molecule engine run molecule after molecule on a sample."""
    import unomol_amd
    from unomol_amd.workloads import batched_pool
    tmp = tempfile.mkdtemp()
    paths = batched_pool(tmp, args.molecules, 5)
    opts = dict(maxits=200, eps=1e-8, cflag=3, accel=2)
    from unomol_amd.deck import read_deck
    sizes = [read_deck(p)["norb"] for p in paths]
    log("batched: %d molecules, %d-%d functions (%d RHF, %d UHF)" %
        (len(paths), min(sizes), max(sizes), sum(1 for i in range(len(paths)) if i % 2 == 0),
         sum(1 for i in range(len(paths)) if i % 2 == 1)))
    for i in range(args.warmup):
        t0 = time.perf_counter()
        unomol_amd.Batch(paths).run(**opts)
        log("warmup %d: %.2f s" % (i, time.perf_counter() - t0))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch = unomol_amd.Batch(paths)
        e, res = batch.run(**dict(opts, verbose=1 if args.verbose else 0))
    wall = (time.perf_counter() - t0) / args.steps
    graph = int(batch.get("graph"))
    fock_launch = {"graph": graph, "list_builds": int(batch.get("space_builds")),
                   "graph_builds": int(batch.get("graph_builds")), "graph_replays": int(batch.get("graph_replays")),
                   "kernels_per_build": int(batch.get("launches_per_build")),
                   "host_launches_per_build": 1 if graph else int(batch.get("launches_per_build"))}
    nconv = sum(1 for r in res if r.converged)
    its = sorted(r.iterations for r in res)
    iters = its[-1]
    log("batched: %.2f s (99%% of the molecules done after %.2f s; Fock %.2f s, eigen/GEMM %.2f s), %d/%d "
        "converged, %d lockstep iterations (per molecule: median %d, 90%% %d)" %
        (wall, res[0].wall99_s, res[0].fock_s, res[0].diag_s, nconv, len(paths), iters, its[len(its) // 2],
         its[(9 * len(its)) // 10]))
    stragglers = []
    for i, r in enumerate(res):
        if not r.converged:
            d = read_deck(paths[i])
            stragglers.append({"molecule": i, "functions": d["norb"], "electrons": d["nelec"],
                               "atoms": int(len(d["charge"])), "ediff": r.ediff, "pdiff": r.pdiff})
            log("not converged: molecule %d (%d functions, %d electrons, %d atoms), ediff %.2e pdiff %.2e" %
                (i, d["norb"], d["nelec"], len(d["charge"]), r.ediff, r.pdiff))
    # the same molecules one by one through the single-molecule engine (a sample)
    nsample = min(args.sample, len(paths))
    t1 = time.perf_counter()
    worst = 0.0
    for i in range(nsample):
        b = unomol_amd.Basis(paths[i])
        t = unomol_amd.TwoElectronInts(b)
        cls = unomol_amd.UnRestrictedHartreeFock if b.deck["nelec"] % 2 else unomol_amd.RestrictedHartreeFock
        e1 = cls(b, t).findEnergy(**opts)
        worst = max(worst, abs(e1 - e[i]))
    single = (time.perf_counter() - t1) / nsample
    log("single-molecule engine: %.3f s per molecule on %d molecules; max |E_batch - E_single| %.2e" %
        (single, nsample, worst))
    print(json.dumps({
        "metric": "batched SCF throughput, %d mixed RHF/UHF molecules (6-31G / 6-31G*), |dE|<1e-8 Eh" % len(paths),
        "value": round(len(paths) / wall, 2), "unit": "molecules/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * wall, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "%d molecules: distorted H2 HF H2O NH3 CH4 N2 CO CO2 C2H2 and 1-8-water clusters, "
                               "6-31G or 6-31G*, %d-%d functions, every other one a +1 cation (UHF), seed 5" %
                               (len(paths), min(sizes), max(sizes)), "global_batch": len(paths),
                   "parallelism": "single-gpu"},
        "converged": nconv, "lockstep_iterations": iters,
        "wall_until_99pct_s": round(res[0].wall99_s, 3), "fock_s": round(res[0].fock_s, 3),
        "diag_s": round(res[0].diag_s, 3), "fock_build_launches": fock_launch,
        "iterations_median": its[len(its) // 2], "iterations_p90": its[(9 * len(its)) // 10],
        "not_converged": stragglers,
        "single_molecule_s_per_molecule": round(single, 4), "single_sample": nsample,
        "speedup_vs_single_molecule": round(single * len(paths) / wall, 2),
        "max_energy_diff_vs_single": worst}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--waters", type=int, default=64)
    ap.add_argument("--seed", type=int, default=64001)
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="default row-merge configuration, no tuning")
    ap.add_argument("--fp64-only", action="store_true", help="time the plain FP64 path as the steps")
    ap.add_argument("--fock-only", action="store_true",
                    help="hardware-counter runs: context build, 2 FP64 SCF updates, one profiled Fock "
                         "build and a 1 GiB calibration copy; prints the per-class table")
    ap.add_argument("--batched", action="store_true",
                    help="batched SCF of many small molecules instead of the 64-water step")
    ap.add_argument("--molecules", type=int, default=512)
    ap.add_argument("--sample", type=int, default=32, help="molecules run one by one for comparison")
    ap.add_argument("--verbose", action="store_true", help="batched: one stderr line per lockstep iteration")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build measures one MI355X (--gpus 1); "
              "multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    import unomol_amd
    from unomol_amd import _native as N
    if args.batched:
        run_batched(args)
        return

    tmp = tempfile.mkdtemp()
    deck = os.path.join(tmp, "patin.dat")
    unomol_amd.generate_deck(deck, "water_cluster" if args.waters > 1 else "water", args.waters, args.seed,
                             "6-31G*", 0, args.eps, 200)
    basis = unomol_amd.Basis(deck)
    log("deck: %d functions, %d electrons" % (basis.number_of_orbitals(), basis.deck["nelec"]))
    fp64_opts = dict(maxits=200, eps=args.eps, cflag=3, accel=2, guess=0)

    if args.fock_only:
        t = unomol_amd.TwoElectronInts(basis)
        hf = unomol_amd.RestrictedHartreeFock(basis, t)
        hf.findEnergy(**dict(fp64_opts, maxits=2))
        classes = profile_classes(N, t, hf.P)
        ms = C.c_double()
        N.check(N.lib().uno_copy_probe(C.c_ulonglong(1 << 30), 1, C.byref(ms)))
        print(json.dumps({"fock_only": True, "copy_probe_bytes": 1 << 30, "copy_ms": ms.value,
                          "copy_GBps": 2 * (1 << 30) / (ms.value * 1e-3) / 1e9, "roofline": roofline_of(classes),
                          "eri_classes": classes}))
        return

    model, model_info = (None, None) if args.fp64_only else get_model(unomol_amd)
    fast_opts = dict(fp64_opts, precision=1, incremental=1, guess=2, fp64_final=2)

    ctx_s = [0.0]

    def one_scf(opts, profile=False, fp32_threshold=None, launch_streams=None):
        t0 = time.perf_counter()
        t = unomol_amd.TwoElectronInts(basis)
        ctx_s[0] = time.perf_counter() - t0
        if fp32_threshold is not None:
            t.set("fp32_threshold", fp32_threshold)
        if launch_streams is not None:
            t.set("launch_streams", launch_streams)
        if opts.get("guess") == 2:
            unomol_amd.set_guess_model(t, model)
        if profile:
            N.check(N.lib().uno_profile(t.ctx, 1, None, None, None, None))
        hf = unomol_amd.RestrictedHartreeFock(basis, t)
        e = hf.findEnergy(**opts)
        dt = time.perf_counter() - t0
        if not hf.result.converged:
            raise RuntimeError("SCF did not converge")
        return t, hf, e, dt

    step_opts = fp64_opts if args.fp64_only else fast_opts
    # autotuner (outside the timed region, once per basis layout): every quartet space's
    # row-merge configuration timed on a two-iteration density, for the mixed and the FP64
    # builds; later contexts of this deck pick the choices up
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tt = unomol_amd.TwoElectronInts(basis)
        hft = unomol_amd.RestrictedHartreeFock(basis, tt)
        hft.findEnergy(**dict(fp64_opts, maxits=2))
        ms_mixed = tt.autotune(hft.P, mixed=not args.fp64_only)
        ms_fp64 = tt.autotune(hft.P, mixed=False)
        tune = {"seconds": round(time.perf_counter() - t0, 3), "tuned_build_ms_mixed": round(ms_mixed, 3),
                "tuned_build_ms_fp64": round(ms_fp64, 3)}
        log("autotune: %.2f s (tuned build %.1f ms mixed, %.1f ms FP64)" % (tune["seconds"], ms_mixed, ms_fp64))
        del tt, hft
    # the plain FP64 reference run (core-H guess, full FP64 builds), the ablation runs and
    # the 16-water run, all before the timed steps so that the timed steps' launches are the
    # last of their kernels in the process
    ablation = {}
    det = None
    if not args.fp64_only:
        t64, hf64, e64, dt64 = one_scf(fp64_opts)
        ctx64_s = ctx_s[0]
        r64 = hf64.result
        ev64, P64 = hf64.evals.copy(), hf64.P.copy()
        log("fp64 reference: E=%.10f in %.2f s, %d iterations" % (e64, dt64, r64.iterations))
        for name, opts in (("fp64+learned", dict(fp64_opts, guess=2)),
                           ("mixed+coreH", dict(fp64_opts, precision=1, incremental=1, fp64_final=2))):
            _, hfa, ea, dta = one_scf(opts)
            ablation[name] = {"scf_s": round(dta, 4), "iterations": hfa.result.iterations,
                              "energy_diff_vs_fp64": ea - e64}
            log("ablation %s: E=%.10f in %.2f s, %d iterations" % (name, ea, dta, hfa.result.iterations))
        # the timed path with no FP32 quartets (fp32_threshold 0): what the FP32 split costs
        # or saves with these kernels; the configured workload keeps its 1e-6 threshold
        _, hfa, ea, dta = one_scf(fast_opts, fp32_threshold=0.0)
        ablation["mixed+learned, no FP32 quartets"] = {"scf_s": round(dta, 4), "iterations": hfa.result.iterations,
                                                       "energy_diff_vs_fp64": ea - e64}
        log("ablation no FP32 quartets: E=%.10f in %.2f s, %d iterations" % (ea, dta, hfa.result.iterations))
        # deterministic Fock builds (fixed-point J/K accumulation, wave tasks only) beside the
        # default FP64 build on the FP64 run's converged density: wall time per build (packed P
        # up, packed G down included), bitwise repeat, largest difference from the default
        def timed_build(ctx):
            G = np.zeros_like(P64)
            t0 = time.perf_counter()
            ctx.formGmatrix(P64, G)
            return G, 1e3 * (time.perf_counter() - t0)

        timed_build(t64)
        Gdef, ms_def = timed_build(t64)
        t64.set("deterministic", 1)
        timed_build(t64)
        Gdet, ms_det = timed_build(t64)
        Gdet2, _ = timed_build(t64)
        t64.set("deterministic", 0)
        det = {"build_ms_default": round(ms_def, 2), "build_ms_deterministic": round(ms_det, 2),
               "repeat_bitwise": bool(np.array_equal(Gdet, Gdet2)),
               "max_abs_diff_vs_default": float(np.abs(Gdet - Gdef).max())}
        log("deterministic build: %.1f ms (default %.1f ms), repeat bitwise %s, max |dG| %.1e" %
            (ms_det, ms_def, det["repeat_bitwise"], det["max_abs_diff_vs_default"]))
    # the 16-water cluster: plain FP64 GPU against the reference program's measured run
    log("config b: 16-water 6-31G*, three FP64 iterations")
    b_s, b_e = run_config_b(unomol_amd)
    refb = reference_config_b()
    cfgb = {"workload": "16-water 6-31G* RHF (304 functions), seed 16001, core-H guess, 3 SCF iterations with "
                        "the reference's damping", "gpu_fp64_s": round(b_s, 4), "energy": b_e}
    if refb:
        cfgb["reference"] = refb
        for k in ("serial_s", "mpi8_s"):
            if refb.get(k):
                cfgb["speedup_vs_reference_" + k[:-2]] = round(refb[k] / b_s, 1)
    log("config b: %.3f s on the GPU%s" % (b_s, "" if not refb else "; reference serial %.1f s" %
                                                                       refb.get("serial_s", float("nan"))))
    iso_classes, iso_step_s = None, None
    if not args.no_profile and not args.fp64_only:
        ti, _, _, iso_step_s = one_scf(step_opts, profile=True, launch_streams=1)
        iso_classes = class_table(*read_profile(N, ti))
        del ti
    for i in range(args.warmup):
        _, _, e, dt = one_scf(step_opts)
        log("warmup %d: E=%.10f in %.2f s" % (i, e, dt))
    # timed region: every step's Fock-build kernels carry per-class HIP events on the
    # engine's stream (and slot-spread evaluated-quartet counters)
    results = []
    acc = [np.zeros(PROF_SLOTS) for _ in range(4)]
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t, hf, e, dt = one_scf(step_opts, profile=not args.no_profile)
        step_ctx_s = ctx_s[0]
        if not args.no_profile:
            for k, v in enumerate(read_profile(N, t)):
                acc[k] += v
        log("step: E=%.10f in %.2f s, %d iterations" % (e, dt, hf.result.iterations))
        results.append((t, hf, e, dt))
    wall = time.perf_counter() - t_all
    t, hf, energy, dt = results[-1]
    r = hf.result
    classes, roofline = [], None
    if not args.no_profile:
        for v in acc:   # per step
            v /= args.steps
        classes = class_table(*acc)
        roofline = roofline_of(classes)
        log("dominant kernel %s: %d launches per step, %.3f ms per launch, %.3f TFLOP/s" %
            (roofline["kernel"], roofline["launches"], roofline["avg_launch_ms"], roofline["achieved"]))
        # the per-class table is long: stderr only, one JSON line
        print("[bench] eri_classes " + json.dumps(classes), file=sys.stderr, flush=True)

    # The timed steps spread each build's ERI launches over three streams, so a launch's
    # events span the time it shared the GPU with the other streams' kernels.  Beside the
    # timed figure: the same step with every launch on one stream (run before the timed
    # region, so the timed launches stay the last of their kernels), the dominant kernel's
    # duration and rate in isolation.
    if roofline is not None and iso_classes is not None:
        same = [c for c in iso_classes if kernel_name(c) == roofline["kernel"]]
        if same:
            ms1 = sum(c["ms"] for c in same)
            n1 = sum(c["launches"] for c in same)
            fl1 = sum(c["gflop"] for c in same) * 1e9
            peak = roofline["peak"]
            tot_fl = sum(c["gflop"] for c in iso_classes) * 1e9
            tot_ms = sum(c["ms"] for c in iso_classes)
            roofline["streams"] = 3
            roofline["one_stream"] = {
                "step_s": round(iso_step_s, 4), "launches": int(n1), "avg_launch_ms": round(ms1 / max(n1, 1), 4),
                "achieved": round(fl1 / (ms1 * 1e-3) / 1e12, 4), "frac": round(fl1 / (ms1 * 1e-3) / 1e12 / peak, 5),
                "all_eri_kernels": {"ms": round(tot_ms, 3), "tflops": round(tot_fl / (tot_ms * 1e-3) / 1e12, 4)}}
            log("one stream: step %.3f s, dominant kernel %.4f ms per launch, %.3f TFLOP/s" %
                (iso_step_s, roofline["one_stream"]["avg_launch_ms"], roofline["one_stream"]["achieved"]))

    if args.fp64_only:
        t64, hf64, e64, dt64 = t, hf, energy, dt
        ctx64_s = step_ctx_s
        r64 = hf64.result
        ev64, P64 = hf64.evals.copy(), hf64.P.copy()
    # errors of the timed path against the plain FP64 GPU run (same deck, same threshold)
    nocc = basis.deck["nelec"] // 2
    err = {"energy": energy - e64,
           "orbital_energy_max_abs": float(np.abs(hf.evals - ev64).max()),
           "occupied_orbital_energy_max_abs": float(np.abs(hf.evals[:nocc] - ev64[:nocc]).max()),
           "density_max_abs": float(np.abs(hf.P - P64).max())}
    log("errors vs FP64: dE %.2e, max |d eps| %.2e (occupied %.2e), max |dP| %.2e" %
        (err["energy"], err["orbital_energy_max_abs"], err["occupied_orbital_energy_max_abs"],
         err["density_max_abs"]))
    # streaming copy of 1 GiB: the HBM reference point for the traffic counters
    cms = C.c_double()
    N.check(N.lib().uno_copy_probe(C.c_ulonglong(1 << 30), 1, C.byref(cms)))
    copy = {"bytes_each_way": 1 << 30, "ms": round(cms.value, 4),
            "GBps": round(2 * (1 << 30) / (cms.value * 1e-3) / 1e9, 1)}
    log("cpu baseline")
    cpu = None if args.no_cpu_baseline else cpu_baseline(deck, r64.iterations)

    def summary(res, secs, ctx):
        return {"scf_s": round(secs, 4), "context_s": round(ctx, 4), "iterations": res.iterations,
                "fock_builds": res.nfock, "full_builds": res.nfull, "fock_s": round(res.fock_s, 4),
                "diag_s": round(res.diag_s, 4), "guess_s": round(res.guess_s, 4), "setup_s": round(res.setup_s, 4),
                "diis_s": round(res.diis_s, 4), "shell_quartets": res.quartets, "fp32_quartets": res.quartets32,
                "fock_builds_per_s": round(res.nfock / res.fock_s, 4) if res.fock_s > 0 else None,
                "shell_quartets_per_s": round(res.quartets / res.fock_s, 1) if res.fock_s > 0 else None}

    out = {
        "metric": "SCF time-to-converge (|dE|<1e-8 Eh) + Fock builds/s, mixed+learned vs FP64 GPU",
        "value": round(wall / args.steps, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.fp64_only else "f64+f32",
        "data": "synthetic",
        "config": {"workload": "%d-water 6-31G* RHF (%d functions), seed %d, DIIS, |dE|<%g Eh; steps: %s" %
                               (args.waters, t.norb, args.seed, args.eps,
                                "plain FP64, core-H guess" if args.fp64_only else
                                "mixed FP32/FP64 + incremental builds + learned guess"),
                   "global_batch": 1, "parallelism": "single-gpu"},
        "energy": energy,
        "energy_fp64": e64,
        "errors_vs_fp64": err,
        "fp64_s": round(dt64, 4),
        "speedup_vs_fp64": round(dt64 / (wall / args.steps), 3),
        "fock_builds_per_s": round(r.nfock / r.fock_s, 3) if r.fock_s > 0 else None,
        "steps_detail": summary(r, wall / args.steps, step_ctx_s),
        "fp64_detail": summary(r64, dt64, ctx64_s),
        "ablation": ablation,
        "guess_model": model_info,
        "autotune": tune,
        "deterministic_fock": det,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "config_b": cfgb,
        "hbm_copy_probe": copy,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
