#!/usr/bin/env python3
"""Benchmark: Metropolis moves/s and pair-evals/s on the 12,288-atom water box.

Workload (one GPU): 4,096 rigid TIP3P-like waters (12,288 atoms) in a cubic
box of 49.7 A, switched Coulomb+LJ (10/12 A), 2 A list skin, T = 300 K,
residue TRN (0.3 A) and ROT (15 deg) moves.  A "step" is a window of
`--moves` Metropolis moves run by the device loop (per-move dE, accept test
and commit on the GPU), with the Verlet list rebuilt on the GPU every 50
moves as the reference schedules it.  value = moves per second.

Batched replicas (same run): 256 water boxes of 300-3,000 molecules
(rho = 0.0334 / A^3, geometric temperature ladder 280-400 K), one persistent
kernel launch per sweep with on-device replica exchange every 100 sweeps,
replayed as a hipGraph, launch shape picked by the library's tuner; compared
with the same replicas run one at a time through the single-chain device
loop -> "batched".

Large mixed system (same run): the ~100k-atom water + chain box -- pair
kernels, list build, and the device loop with the 50-move list schedule
and with displacement-triggered rebuilds -> "large_mixed".

Force-bias (same run): residue force-bias moves (A = 0.01 A^2, drift with
F in kcal/mol/A; and the drift at the Metropolis test's own beta, A adapted)
against plain residue translations on the 12,288-atom box: moves/s,
acceptance and accepted mean-square displacement per second; the adjoint
kernel's time on that box and a 64-configuration batched adjoint call ->
"force_bias".

Also measured in the same run (HIP events on the library's own stream):
  * "roofline": the Metropolis kernel of the headline loop -- algorithmic
    FLOP of a move's dE (moved atoms x full-row candidates x 2 x (17 + 23
    f_in)) x moves per launch / mean launch time, every launch of the timed
    region bracketed by HIP events (MCX_KERNEL_EVENTS=1);
  * "secondary_rooflines": the O(N^2) pair kernel, the switched list
    energy kernel (the drop-in getE), the full list build, the in-run list
    update and the batched sweep kernel;
  * "getE_end_to_end": host coordinate upload + energy, as the drop-in
    FField_MCX::getE does it, beside the reference's getE on the CPU;
  * "cpu_baseline": the reference's own CPU loop (oracle/_ref/libmcref.so,
    compiled from the mc_cpp sources) on one core, the FP64 restatement on
    all cores and in its CPU dE mode, bounded samples on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# algorithmic FLOP per unit, FMA = 2 (the constants frozen in csrc/mcx_internal.h)
FLOP_FULL_PAIR = 32.0          # full O(N^2) pair
FLOP_LIST_ENTRY = 17.0         # switched list entry (distance, min image, r^2)
FLOP_LIST_INCUT = 23.0         # + per entry inside the cutoff
COPY_BYTES = 256 << 20         # calibration copy of the counter runs


def batched_replicas(args):
    """256-replica ladder: batched sweeps vs the same replicas looped one at a
    time on the GPU (one sweep each through the single-chain device loop)."""
    from mc_cpp_amd import mcx, systems
    reps = systems.water_replicas(seed=4, nrep=args.replicas)
    handles = [mcx.from_system(s) for s, _ in reps]
    b = mcx.Batch(handles, reps[0][0].moves)
    nrep = len(reps)
    temps = [280.0 * (400.0 / 280.0) ** (k / max(nrep - 1, 1)) for k in range(nrep)]
    b.set_ladder(temps)
    b.reset(2024, mcx.MC_PHILOX)
    threads, unroll, ms_tuned = b.tune()
    b.run(args.batch_warmup, exchange_every=100, use_graph=True)
    out = b.run(args.batch_sweeps, exchange_every=100, use_graph=True)
    bmps = out["moves"] / out["seconds"]
    tm = nm = 0.0
    for (s, _), h in zip(reps, handles):
        r = h.mc_run(s.moves, s.natom // 3, 99, rng_mode=mcx.MC_PHILOX, list_every=50, record=False)
        tm += r["seconds"]
        nm += s.natom // 3
    pmps = nm / tm
    natoms = sum(s.natom for s, _ in reps)
    # algorithmic FLOP of one sweep (one move per molecule in every replica)
    flop_sweep = 0.0
    for (s, _), h in zip(reps, handles):
        npairs, nin, row_mean = list_stats(h, s)
        flop_sweep += (s.natom // 3) * move_flop(3.0, row_mean, nin / max(npairs, 1))
    ms_sweep = out["seconds"] / args.batch_sweeps * 1e3
    tf = flop_sweep / (ms_sweep * 1e-3) / 1e12
    return {"workload": "%d water boxes, %d-%d molecules (%d atoms), T ladder 280-400 K, exchange every 100 sweeps, "
                        "residue TRN 0.3 A + ROT 15 deg, sweep = one move per molecule"
                        % (nrep, min(s.natom for s, _ in reps) // 3, max(s.natom for s, _ in reps) // 3, natoms),
            "moves_per_s": bmps, "sweeps": args.batch_sweeps, "moves": out["moves"], "seconds": out["seconds"],
            "ms_per_sweep": out["seconds"] / args.batch_sweeps * 1e3,
            "pair_evals_per_s": 2.0 * out["pair_evals"] / out["seconds"],
            "acceptance": out["accepted"] / out["moves"],
            "exchange_attempts": out["exchange_attempts"], "exchange_accepted": out["exchange_accepted"],
            "launch_shape": {"threads": threads, "tiles_in_flight": unroll, "tuned_ms_per_sweep": ms_tuned},
            "plain_loop_moves_per_s": pmps,
            "plain_loop": "each replica one sweep through the single-chain device loop (mcx_mc_run, list every 50)",
            "speedup_vs_plain": bmps / pmps,
            "roofline": {"bound": "valu", "kernel": "batched sweep kernel (one persistent launch per sweep, "
                                                    "FP64 pair change; time per sweep from the graph replays)",
                         "achieved": tf, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": tf / PEAK_FP64_TFLOPS,
                         "flop_per_launch": flop_sweep, "ms_per_launch": ms_sweep, "traffic": None}}


def force_bias_vs_translation(args):
    """Force-bias residue moves (dx = beta A F + sqrt(2A) xi, A = 0.01 A^2,
    proposal-density ratio in the acceptance) against plain residue
    translations (0.3 A) on the 12,288-atom water box, plus the adjoint
    kernel alone and a 64-system batched adjoint call."""
    from mc_cpp_amd import mcx, systems
    out = {}
    for name, mt in (("force_bias", systems.MoveType(systems.MV_FBIAS, systems.MODE_RESIDUE, 0.01, 50.0, 0)),
                     # drift at the test's own beta; A adapted toward 20 % acceptance every 100 moves
                     # (on the speculative kernel a lower acceptance resolves more moves per round:
                     # targets 50 / 30 / 20 / 10 / 5 % gave 1,116 / 1,440 / 1,487 / 1,463 / 1,151 A^2/s)
                     ("force_bias_test_beta", systems.MoveType(systems.MV_FBIAS_ACC, systems.MODE_RESIDUE, 5e-6,
                                                               20.0, 100)),
                     ("translation", systems.MoveType(systems.MV_TRN, systems.MODE_RESIDUE, 0.3, 50.0, 0))):
        s, _ = systems.water_box(seed=2)
        ff = mcx.FField(s)
        ff.h.mc_run([mt], 2000, 5, rng_mode=mcx.MC_PHILOX, list_every=50, record=False)   # warm-up
        r = ff.h.mc_run([mt], args.fb_moves, 6, rng_mode=mcx.MC_PHILOX, list_every=50, record=True)
        acc = r["acc"].astype(bool)
        sq = float((r["disp"][acc] ** 2).sum())          # A^2, accepted molecule displacements
        out[name] = {"moves_per_s": args.fb_moves / r["seconds"],
                     "acceptance": float(r["accepted"][0]) / args.fb_moves, "moves": args.fb_moves,
                     "final_step": float(r["dmax"][0]),
                     "accepted_sq_disp_A2_per_s": sq / r["seconds"],
                     "mean_sq_disp_per_move_A2": sq / args.fb_moves}
    out["force_bias_over_translation"] = out["force_bias"]["moves_per_s"] / out["translation"]["moves_per_s"]
    for name in ("force_bias", "force_bias_test_beta"):
        out[name]["sq_disp_per_s_over_translation"] = (out[name]["accepted_sq_disp_A2_per_s"] /
                                                        out["translation"]["accepted_sq_disp_A2_per_s"])
    # adjoint kernel: one launch over the 12,288-atom box (HIP events)
    s, _ = systems.water_box(seed=2)
    h = mcx.FField(s).h
    ms = h.time_kernel(4, 50)
    npairs = h.update_list()
    out["adjoint_kernel_ms"] = ms
    out["adjoint_pair_visits_per_s"] = 2.0 * npairs / (ms * 1e-3)
    # 64 configurations in one batched adjoint call (includes the copy of
    # six gradients per atom back to the host)
    boxes = [systems.water_box(seed=100 + k)[0] for k in range(64)]
    hs = [mcx.FField(b).h for b in boxes]
    mcx.grad_batch(hs)
    t0 = time.perf_counter()
    for _ in range(3):
        mcx.grad_batch(hs)
    out["batched_adjoint"] = {"systems": 64, "atoms_each": boxes[0].natom,
                              "ms_per_call": (time.perf_counter() - t0) / 3 * 1e3,
                              "note": "wall time of mcx_nonbonded_grad_batch incl. the device-to-host copy"}
    return out


def large_mixed(args):
    """The ~100k-atom mixed system (32,768-site water lattice minus the
    waters within 2.6 A of a 2,000-bead chain, L = 99.4 A): pair kernels,
    list build and the device Metropolis loop with the reference's 50-move
    list schedule and with displacement-triggered rebuilds."""
    from mc_cpp_amd import mcx, systems
    s, _ = systems.mixed_system()
    n = s.natom
    h = mcx.FField(s).h
    npairs = h.update_list()
    out = {"workload": "mixed_%d: water lattice + 2,000-bead chain, L=99.4 A, switch 10/12 A, skin 2 A, T=300 K, "
                       "residue TRN 0.3 A / ROT 15 deg (waters) + atom TRN 0.1 A (chain)" % n,
           "atoms": n, "list_pairs": npairs,
           "switch_energy_ms": h.time_kernel(1, 20), "list_build_ms": h.time_kernel(2, 5)}
    sf = s.copy()
    sf.cutmode = systems.CUT_FULL
    hf = mcx.from_system(sf)
    hf.energy()
    ms = hf.time_kernel(0, 3)
    out["full_energy_ms"] = ms
    out["full_pair_evals_per_s"] = n * (n - 1) / 2 / (ms * 1e-3)
    for name, every in (("mc_list_every_50", 50), ("mc_list_auto", mcx.LIST_AUTO)):
        h.mc_run(s.moves, 500, 1, rng_mode=mcx.MC_PHILOX, list_every=every, record=False)
        r = h.mc_run(s.moves, args.large_moves, 2, rng_mode=mcx.MC_PHILOX, list_every=every, record=False)
        out[name] = {"moves_per_s": args.large_moves / r["seconds"], "moves": args.large_moves,
                     "acceptance": float(r["accepted"].sum()) / args.large_moves,
                     "list_rebuilds": int(r["list_updates"])}
    return out


def list_stats(h, s):
    """(list entries, entries inside the cutoff, mean full-row length) of
    the handle's current pair list, from the list itself and the
    coordinates (min-image, the run's box)."""
    import numpy as np
    cnt, flat = h.get_list()
    x, y, z = h.get_coords()
    i = np.repeat(np.arange(s.natom, dtype=np.int64), cnt)
    j = flat.astype(np.int64)
    r2 = np.zeros(len(j))
    for c, L in zip((x, y, z), s.box):
        d = c[j] - c[i]
        if s.pbtype != 0:
            d -= L * np.rint(d / L)
        r2 += d * d
    nin = int((r2 <= s.ctoff * s.ctoff).sum())
    return int(len(j)), nin, 2.0 * len(j) / s.natom


def move_flop(atoms_per_move, row_mean, f_in):
    """Algorithmic FLOP of one move's dE: every moved atom's full-row
    candidates, old and new position, 17 per candidate + 23 in cutoff."""
    return atoms_per_move * row_mean * 2.0 * (FLOP_LIST_ENTRY + FLOP_LIST_INCUT * f_in)


def kernel_probe(args):
    """Fixed kernel sequence for the HBM-traffic counter runs
    (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE): the headline device loop for
    --probe-moves moves (Metropolis kernel launches of 50 moves with their
    list updates), 20 launches of the O(N^2) pair kernel and of the
    switched list-energy kernel, 5 list builds, then 5 launches of a plain
    device copy of COPY_BYTES (torch's elementwise copy of float32), whose
    counter readings calibrate the counters' units."""
    import torch
    from mc_cpp_amd import mcx, systems
    s, _ = systems.water_box(seed=2)
    h = mcx.FField(s).h
    h.mc_run(s.moves, 500, 1000, rng_mode=mcx.MC_REPLAY, list_every=50, record=False)
    h.mc_run(s.moves, args.probe_moves, 2000, rng_mode=mcx.MC_REPLAY, list_every=50, record=False)
    h.time_kernel(1, 20)
    h.time_kernel(2, 5)
    sf = s.copy()
    sf.cutmode = systems.CUT_FULL
    hf = mcx.from_system(sf)
    hf.energy()
    hf.time_kernel(0, 20)
    large = ""
    if args.probe_large:
        # the 92,384-atom mixed box: its switched energy (16 lanes per row)
        # and the sorted-candidate list build
        sl, _ = systems.mixed_system(seed=3)
        hl = mcx.from_system(sl)
        hl.update_list()
        hl.time_kernel(1, 20)
        hl.time_kernel(2, 5)
        large = ", mixed_92384: k_switch_energy x21, list build x6"
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "mc_run %d+500 moves, k_switch_energy x21, list build x6, k_full_energy x21 "
                               "(1 warm-up each)%s, copy x5" % (args.probe_moves, large),
                      "copy_bytes": COPY_BYTES, "atoms": s.natom}), flush=True)


def traffic_from_profile(name):
    """HBM bytes per launch of a kernel from its committed counter summary
    (a line 'traffic_bytes_per_launch: <int>' in profiles/<name>), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def host_info():
    nproc = len(os.sched_getaffinity(0))
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": nproc, "os_cpu_count": os.cpu_count(), "model": model}


def cpu_baselines(s, fp, args):
    """CPU timings on this host, bounded samples of the headline workload:
      * the reference's own MC_metropolis::run (oracle/_ref/libmcref.so, two
        full switched getE per move, list rebuilt every 50 moves), one
        thread -- the reference's energy loops are serial;
      * the FP64 restatement (oracle) in the same compat mode on all cores
        (rows of the pair loops dealt out by OpenMP);
      * the restatement's CPU dE mode (only the pairs of the moved atoms,
        old and new, same decisions), one thread and all cores;
      * one end-to-end reference getE() (src/mc_cpp.cpp:169-201 times this
        call).
    Returns the cpu_baseline object."""
    import ctypes
    from oracle import oracle
    from mc_cpp_amd import systems
    info = host_info()
    allc = max(1, min(info["nproc"], args.cpu_threads))
    modes = {}
    out = {"host": info}
    gomp = None
    try:
        gomp = ctypes.CDLL("libgomp.so.1")
    except OSError:
        pass

    def oracle_run(mode, nmoves, threads):
        sc = s.copy()
        o = oracle.Oracle(sc)
        o.update_list(threads)
        oracle.set_energy_threads(threads)
        try:
            t0 = time.perf_counter()
            o.mc_run(nmoves, 7, mode=mode, list_every=50, nthreads=threads)
            dt = time.perf_counter() - t0
        finally:
            oracle.set_energy_threads(1)
        return {"moves_per_s": nmoves / dt, "moves": nmoves, "seconds": dt, "threads": threads}

    ref = None
    if oracle.ref_available():
        d = tempfile.mkdtemp(prefix="mcx_bench_")
        cor, ff = systems.save_mdbas(os.path.join(d, "water"), s.copy(), fp)
        if gomp is not None:
            gomp.omp_set_num_threads(1)
        r = oracle.Reference(cor, ff, s.copy())
        t0 = time.perf_counter()
        r.getE()
        ge = time.perf_counter() - t0
        t0 = time.perf_counter()
        r.mc_run(s.moves, args.cpu_moves, 7)
        dt = time.perf_counter() - t0
        ref = {"moves_per_s": args.cpu_moves / dt, "moves": args.cpu_moves, "seconds": dt, "threads": 1}
        modes["reference_compat_1thread"] = ref
        out["reference_getE_ms"] = ge * 1e3
        if gomp is not None:
            gomp.omp_set_num_threads(info["nproc"])
    modes["port_compat_allcores"] = oracle_run(0, args.cpu_moves, allc)
    modes["port_dE_1thread"] = oracle_run(1, args.cpu_de_moves, 1)
    modes["port_dE_allcores"] = oracle_run(1, args.cpu_de_moves, allc)
    out["modes"] = modes
    if ref is not None:
        out.update({"value": ref["moves_per_s"], "unit": "moves/s", "cores": 1, "kind": "reference",
                    "sample": "%d moves of the reference's MC_metropolis::run on the 12,288-atom water box (two "
                              "full switched getE per move, list rebuilt every 50 moves), %.1f s on one core; "
                              "'modes' adds the FP64 restatement on %d cores and its CPU dE mode"
                              % (ref["moves"], ref["seconds"], allc)})
    else:
        m = modes["port_compat_allcores"]
        out.update({"value": m["moves_per_s"], "unit": "moves/s", "cores": allc, "kind": "port",
                    "sample": "%d moves of the FP64 restatement of MC_metropolis::run (two full switched "
                              "energies per move), %.1f s on %d cores" % (m["moves"], m["seconds"], allc)})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--moves", type=int, default=20000, help="Metropolis moves per step")
    ap.add_argument("--cpu-moves", type=int, default=200, help="moves of the CPU compat-mode runs")
    ap.add_argument("--cpu-de-moves", type=int, default=200, help="moves of the CPU dE-mode runs")
    ap.add_argument("--cpu-threads", type=int, default=16, help="at most this many threads for 'all cores'")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--event-every", type=int, default=8,
                    help="bracket every k-th kernel launch of the loop with HIP events (0: none)")
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--batch-sweeps", type=int, default=100)
    ap.add_argument("--batch-warmup", type=int, default=1)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--fb-moves", type=int, default=20000, help="moves per force-bias / translation run")
    ap.add_argument("--no-fbias", action="store_true")
    ap.add_argument("--probe", action="store_true", help="kernel sequence for the HBM counter runs only")
    ap.add_argument("--probe-moves", type=int, default=2000)
    ap.add_argument("--probe-large", action="store_true", help="the probe also runs the ~100k-atom system's kernels")
    ap.add_argument("--large-moves", type=int, default=10000, help="moves per run on the ~100k-atom system")
    ap.add_argument("--no-large", action="store_true")
    args = ap.parse_args()
    if args.probe:
        kernel_probe(args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the Metropolis chain is sequential and this build has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.event_every > 0:
        os.environ["MCX_KERNEL_EVENTS"] = str(args.event_every)

    import numpy as np
    from mc_cpp_amd import mcx, systems

    s, fp = systems.water_box(seed=2)
    s_cpu = s.copy()
    ff = mcx.FField(s)
    h = ff.h
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None

    # ---- headline: device Metropolis loop ----
    for w in range(args.warmup):
        h.mc_run(s.moves, args.moves, 1000 + w, rng_mode=mcx.MC_REPLAY, list_every=50, record=False)
    sync()
    t0 = time.perf_counter()
    acc = 0
    nlist = 0
    ev = {"chain_kernel_ms": 0.0, "chain_launches": 0, "list_kernel_ms": 0.0, "list_launches": 0}
    for k in range(args.steps):
        r = h.mc_run(s.moves, args.moves, 2000 + k, rng_mode=mcx.MC_REPLAY, list_every=50, record=False)
        acc += int(r["accepted"].sum())
        nlist += int(r["list_updates"])
        for key in ev:
            ev[key] += r[key]
    sync()
    dt = time.perf_counter() - t0
    nmoves = args.steps * args.moves
    moves_per_s = nmoves / dt

    # ---- algorithmic work of a move, from the list the run ended with ----
    npairs, nin, row_mean = list_stats(h, s)
    f_in = nin / npairs
    atoms_per_move = s.natom / len(np.unique(s.resid))        # residue moves of 3-atom waters
    flop_move = move_flop(atoms_per_move, row_mean, f_in)
    chain = None
    if ev["chain_launches"]:
        ms_launch = ev["chain_kernel_ms"] / ev["chain_launches"]
        moves_launch = args.moves / -(-args.moves // 50)          # a launch runs the moves between list updates
        achieved = flop_move * moves_launch / (ms_launch * 1e-3) / 1e12
        chain = {"bound": "valu", "kernel": "k_mc_wide (device Metropolis loop: %d moves per launch, "
                                            "FP64 pair change)" % round(moves_launch),
                 "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                 "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic_from_profile("k_mc_wide_water12288_pmc.md"),
                 "flop_per_move": flop_move, "flop_per_launch": flop_move * moves_launch,
                 "ms_per_launch": ms_launch, "launches_timed": ev["chain_launches"],
                 "us_per_move_in_kernel": ms_launch * 1e3 / moves_launch,
                 "frac_of_fp32_peak": achieved / PEAK_FP32_TFLOPS,
                 "work": "per move: %.2f moved atoms x %.1f full-row candidates x 2 (old, new) x "
                         "(17 + 23 x f_in), f_in = %.4f in-cutoff fraction of the list" % (atoms_per_move, row_mean, f_in)}

    # ---- end-to-end getE as the drop-in FField_MCX does it: the coordinates
    # the move code changed are uploaded, the energy summed over the list of
    # the last update (mcx_set_coords_moved + mcx_energy) ----
    x, y, z = h.get_coords()
    h.update_list()
    h.set_coords_moved(x, y, z)
    h.energy()
    reps_e = 200
    t0 = time.perf_counter()
    for _ in range(reps_e):
        h.set_coords_moved(x, y, z)
        h.energy()
    getE_ms = (time.perf_counter() - t0) / reps_e * 1e3

    # ---- pair kernel: full O(N^2) energy, HIP events on the library stream ----
    n = s.natom
    s_full = s.copy()
    s_full.cutmode = systems.CUT_FULL
    hf = mcx.from_system(s_full)
    hf.energy()
    reps = 50
    ms_full = hf.time_kernel(0, reps)
    pairs = n * (n - 1) / 2
    flop = FLOP_FULL_PAIR * pairs
    tflops = flop / (ms_full * 1e-3) / 1e12

    # ---- switched list energy kernel and list build ----
    ms_sw = h.time_kernel(1, reps)
    flop_sw = FLOP_LIST_ENTRY * npairs + FLOP_LIST_INCUT * nin
    tf_sw = flop_sw / (ms_sw * 1e-3) / 1e12
    ms_list = h.time_kernel(2, 20)
    npairs_b = h.update_list()
    list_bytes = 24 * n + 4 * (n + 1) + 4 * npairs_b
    list_gbs = list_bytes / (ms_list * 1e-3) / 1e9
    patch = None
    if ev["list_launches"]:
        ms_p = ev["list_kernel_ms"] / ev["list_launches"]
        patch = {"bound": "hbm", "kernel": "k_patch_rows + k_patch_apply (in-run list update every 50 moves; "
                                           "the same list as a full rebuild)",
                 "achieved": list_bytes / (ms_p * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                 "frac": list_bytes / (ms_p * 1e-3) / 1e9 / PEAK_HBM_GBS, "ms_per_update": ms_p,
                 "updates_timed": ev["list_launches"], "bytes_per_update": list_bytes}

    batch = None if args.no_batch else batched_replicas(args)
    fbias = None if args.no_fbias else force_bias_vs_translation(args)
    large = None if args.no_large else large_mixed(args)
    cpu = None if args.no_cpu else cpu_baselines(s_cpu, fp, args)
    if cpu is not None:
        cpu["gpu_over_cpu"] = moves_per_s / cpu["value"]

    secondary = {
        "k_full_energy": {"bound": "valu", "kernel": "k_full_energy (O(N^2) min-image LJ+Coulomb, 12,288 atoms)",
                          "achieved": tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                          "frac": tflops / PEAK_FP32_TFLOPS,
                          "traffic": traffic_from_profile("k_full_energy_water12288_pmc.md"),
                          "peak_without_packed_fp32": PEAK_FP32_TFLOPS / 2,
                          "frac_without_packed_fp32": tflops / (PEAK_FP32_TFLOPS / 2),
                          "flop_per_launch": flop, "ms_per_launch": ms_full,
                          "pair_evals_per_s": pairs / (ms_full * 1e-3)},
        "k_switch_energy": {"bound": "valu", "kernel": "k_switch_energy (switched list energy, the getE of "
                                                       "the drop-in path)",
                            "achieved": tf_sw, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                            "frac": tf_sw / PEAK_FP32_TFLOPS,
                            "traffic": traffic_from_profile("k_switch_energy_water12288_pmc.md"),
                            "flop_per_launch": flop_sw, "ms_per_launch": ms_sw, "list_entries": npairs,
                            "in_cutoff": nin},
        "list_build": {"bound": "hbm", "kernel": "full list build (bin, rows, compaction)",
                       "achieved": list_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                       "frac": list_gbs / PEAK_HBM_GBS, "bytes_per_launch": list_bytes, "ms_per_launch": ms_list,
                       "pairs": npairs_b},
        "list_update_in_run": patch,
    }
    if batch is not None:
        secondary["k_batch_sweep"] = batch.pop("roofline", None)

    out = {
        "metric": "MC moves/s, pair-evals/s (%FP32 peak) on 12k-atom water; batched vs plain",
        "value": moves_per_s,
        "unit": "moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Philox water lattice, random orientations)",
        "config": {"workload": "water_box_12288: 4096 rigid TIP3P-like waters, L=49.7 A, switch 10/12 A, "
                               "skin 2 A, T=300 K, residue TRN 0.3 A + ROT 15 deg, plain single-chain "
                               "device Metropolis loop (MT19937 replay of the reference's draws), list "
                               "updated every 50 moves",
                   "moves_per_step": args.moves, "acceptance": acc / nmoves,
                   "list_rebuilds": nlist, "kernel_events_every": args.event_every},
        "roofline": chain,
        "secondary_rooflines": secondary,
        "pair_evals_per_s": pairs / (ms_full * 1e-3),
        "full_energy_ms": ms_full,
        "switch_energy_ms": ms_sw,
        "getE_end_to_end": {"ms": getE_ms, "what": "mcx_set_coords_moved from host arrays + mcx_energy over the "
                                                   "list of the last update (the drop-in FField_MCX::getE), wall "
                                                   "clock, PCIe included",
                            "reference_cpu_ms": cpu.get("reference_getE_ms") if cpu else None},
        "batched": batch,
        "force_bias": fbias,
        "large_mixed": large,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
