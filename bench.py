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

Force-bias (same run): residue force-bias moves (A = 0.01 A^2) against
plain residue translations on the 12,288-atom box, moves/s and acceptance;
the adjoint kernel's time on that box and a 64-configuration batched
adjoint call -> "force_bias".

Also measured in the same run (HIP events on the library's own stream):
  * the O(N^2) pair kernel (full min-image LJ+Coulomb energy of the box):
    pair-evals/s and FP32 fraction of peak -> "roofline";
  * the list build: algorithmic HBM bytes/s -> "list_build";
  * the reference's own CPU code (oracle/_ref/libmcref.so, compiled from the
    mc_cpp sources) on this host: moves/s of its two-getE-per-move loop
    on a bounded sample -> "cpu_baseline".
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
PEAK_HBM_GBS = 8000.0
FLOP_FULL_PAIR = 32.0          # per pair, FMA = 2 (the constant frozen in csrc/mcx_internal.h)
COPY_BYTES = 256 << 20         # calibration copy of the counter runs


def cpu_reference_baseline(s, fp, nmoves, seed):
    """Times the reference's MC_metropolis::run (two full getE per move) on
    a bounded sample on one core; returns (moves_per_s, seconds)."""
    from oracle import oracle
    from mc_cpp_amd import systems
    if not oracle.ref_available():
        return None
    d = tempfile.mkdtemp(prefix="mcx_bench_")
    cor, ff = systems.save_mdbas(os.path.join(d, "water"), s, fp)
    r = oracle.Reference(cor, ff, s)
    t0 = time.perf_counter()
    r.mc_run(s.moves, nmoves, seed)
    dt = time.perf_counter() - t0
    return nmoves / dt, dt


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
            "speedup_vs_plain": bmps / pmps}


def force_bias_vs_translation(args):
    """Force-bias residue moves (dx = beta A F + sqrt(2A) xi, A = 0.01 A^2,
    proposal-density ratio in the acceptance) against plain residue
    translations (0.3 A) on the 12,288-atom water box, plus the adjoint
    kernel alone and a 64-system batched adjoint call."""
    from mc_cpp_amd import mcx, systems
    out = {}
    for name, mt in (("force_bias", systems.MoveType(systems.MV_FBIAS, systems.MODE_RESIDUE, 0.01, 50.0, 0)),
                     ("translation", systems.MoveType(systems.MV_TRN, systems.MODE_RESIDUE, 0.3, 50.0, 0))):
        s, _ = systems.water_box(seed=2)
        ff = mcx.FField(s)
        ff.h.mc_run([mt], 2000, 5, rng_mode=mcx.MC_PHILOX, list_every=50, record=False)   # warm-up
        r = ff.h.mc_run([mt], args.fb_moves, 6, rng_mode=mcx.MC_PHILOX, list_every=50, record=False)
        out[name] = {"moves_per_s": args.fb_moves / r["seconds"],
                     "acceptance": float(r["accepted"][0]) / args.fb_moves, "moves": args.fb_moves}
    out["force_bias_over_translation"] = out["force_bias"]["moves_per_s"] / out["translation"]["moves_per_s"]
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


def kernel_probe():
    """Fixed kernel sequence for the HBM-traffic counter runs
    (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE): 20 launches of the O(N^2)
    pair kernel on the 12,288-atom box, then 5 launches of a plain device
    copy of COPY_BYTES (torch's elementwise copy of float32), whose counter
    readings calibrate the counters' units."""
    import torch
    from mc_cpp_amd import mcx, systems
    s, _ = systems.water_box(seed=2)
    s.cutmode = systems.CUT_FULL
    h = mcx.from_system(s)
    h.energy()
    h.time_kernel(0, 20)
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "k_full_energy x21 (1 warm-up), copy x5", "copy_bytes": COPY_BYTES,
                      "atoms": s.natom}), flush=True)


def traffic_from_profile():
    """HBM bytes per launch of k_full_energy from the committed counter
    summary (a line 'traffic_bytes_per_launch: <int>'), or None."""
    path = os.path.join(ROOT, "profiles", "k_full_energy_water12288_pmc.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--moves", type=int, default=20000, help="Metropolis moves per step")
    ap.add_argument("--cpu-moves", type=int, default=24, help="moves of the reference CPU loop to time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--batch-sweeps", type=int, default=100)
    ap.add_argument("--batch-warmup", type=int, default=1)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--fb-moves", type=int, default=20000, help="moves per force-bias / translation run")
    ap.add_argument("--no-fbias", action="store_true")
    ap.add_argument("--probe", action="store_true", help="kernel sequence for the HBM counter runs only")
    ap.add_argument("--large-moves", type=int, default=10000, help="moves per run on the ~100k-atom system")
    ap.add_argument("--no-large", action="store_true")
    args = ap.parse_args()
    if args.probe:
        kernel_probe()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the Metropolis chain is sequential and this build has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

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
    for k in range(args.steps):
        r = h.mc_run(s.moves, args.moves, 2000 + k, rng_mode=mcx.MC_REPLAY, list_every=50, record=False)
        acc += int(r["accepted"].sum())
        nlist += int(r["list_updates"])
    sync()
    dt = time.perf_counter() - t0
    moves_per_s = args.steps * args.moves / dt

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
    ms_list = h.time_kernel(2, 20)
    npairs = h.update_list()
    list_bytes = 24 * n + 4 * (n + 1) + 4 * npairs
    list_gbs = list_bytes / (ms_list * 1e-3) / 1e9

    batch = None if args.no_batch else batched_replicas(args)
    fbias = None if args.no_fbias else force_bias_vs_translation(args)
    large = None if args.no_large else large_mixed(args)

    cpu = None
    if not args.no_cpu:
        res = cpu_reference_baseline(s_cpu, fp, args.cpu_moves, 7)
        if res is not None:
            mps, secs = res
            cpu = {"value": mps, "unit": "moves/s", "cores": 1, "kind": "reference",
                   "sample": "%d moves of MC_metropolis::run on this box (two full switched getE per move, "
                             "single-threaded energy loops as the reference builds them), %.1f s"
                             % (args.cpu_moves, secs)}

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
        "dtype": "f32 pair math, f64 Coulomb distance chain and sums",
        "data": "synthetic (seeded Philox water lattice, random orientations)",
        "config": {"workload": "water_box_12288: 4096 rigid TIP3P-like waters, L=49.7 A, switch 10/12 A, "
                               "skin 2 A, T=300 K, residue TRN 0.3 A + ROT 15 deg, plain single-chain "
                               "device Metropolis loop, list rebuilt every 50 moves",
                   "moves_per_step": args.moves, "acceptance": acc / (args.steps * args.moves),
                   "list_rebuilds": nlist},
        "pair_evals_per_s": pairs / (ms_full * 1e-3),
        "full_energy_ms": ms_full,
        "switch_energy_ms": ms_sw,
        "roofline": {"bound": "valu", "kernel": "k_full_energy (O(N^2) min-image LJ+Coulomb, 12,288 atoms)",
                     "achieved": tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / PEAK_FP32_TFLOPS, "traffic": traffic_from_profile(),
                     "peak_without_packed_fp32": PEAK_FP32_TFLOPS / 2,
                     "frac_without_packed_fp32": tflops / (PEAK_FP32_TFLOPS / 2),
                     "flop_per_launch": flop, "ms_per_launch": ms_full},
        "list_build": {"bound": "hbm", "achieved": list_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                       "frac": list_gbs / PEAK_HBM_GBS, "bytes_per_launch": list_bytes, "ms_per_launch": ms_list,
                       "pairs": npairs},
        "batched": batch,
        "force_bias": fbias,
        "large_mixed": large,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
