#!/usr/bin/env python3
"""Benchmark of the scBFM sweep on one MI355X.

Workload (default): the dense melt, 32768 linear chains of N=256
(8,388,608 monomers) on a 512^3 periodic lattice at occupied-site fraction
0.5, 2 species, built by ordered packing and relaxed for 1000 MCS (part of
the workload build, independent of --warmup).  A step is one Monte-Carlo
step (MCS): every species drawn once on average, i.e. n_colors substeps of
pass 1 + pass 2.

    python bench.py [--steps K] [--warmup W] [--lattice bit|u8]

Prints ONE JSON line.  `value` = attempted monomer moves per second over the
K timed MCS (hipEvent time of the graph replays, inputs resident in HBM,
every graph built before the start event).  Also reported:
  roofline      the sweep: algorithmic bytes of the K timed MCS (pass 1 +
                pass 2, the metric's count) / their GPU time, against 8 TB/s;
                per kernel beside it: bytes per launch over the mean launch
                time from HIP events around every launch of a profiled sweep,
                and over the steady-state mean of the committed rocprofv3
                kernel trace of this workload (profiles/)
  cpu_baseline  the CPU oracle (same colored scheme, OpenMP over all host
                threads) on a bounded sample of the same state
  parity        the GPU re-run of the CPU sample, compared bit for bit; the
                other lattice kind's state after the same MCS against this
                one's, and a sample of it against the oracle
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import pscbfm_amd as P  # noqa: E402
from pscbfm_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "attempted monomer moves/sec (1 GPU) + % HBM roofline; batched vs plain at bit-exact"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


SLOT_ORDER = {"auto": None, "spatial": True, "reference": False}


def make_updater(box, xyz, bonds, seed, kind, block=0, slots=0, spatial_order=None):
    sd = W.SystemData(tuple(box), xyz, bonds, seed, "ordered")
    return W.stage(sd, kind, block=block, slots=slots, initialize=True, spatial_order=spatial_order)


def algorithmic_bytes(stats, lattice_bytes_per_cell: float):
    """Bytes per the metric definition: pass 1 = n(12 + 9 c_L) + 16 D per
    substep; pass 2 = A (12 + 2 c_L)."""
    p1 = stats.attempted * (12.0 + 9.0 * lattice_bytes_per_cell) + 16.0 * stats.bond_terms
    p2 = stats.accepted * (12.0 + 2.0 * lattice_bytes_per_cell)
    return p1, p2


def cpu_baseline(u, box, bonds, colors, seed, target_seconds: float, max_mcs: int = 200):
    """Times the oracle on the current state; re-runs the same substeps on the
    GPU and compares (the oracle is the checker here, never the product)."""
    import oracle

    n = colors.size
    xyz0 = u.getPositions(n)
    t0, _ = u.counters()
    n_colors = int(colors.max()) + 1
    a = np.concatenate([bonds[:, 0], bonds[:, 1]]).astype(np.int64)
    b = np.concatenate([bonds[:, 1], bonds[:, 0]]).astype(np.int64)
    order = np.argsort(a, kind="stable")
    a, b = a[order], b[order]
    start = np.zeros(n + 1, dtype=np.int64)
    np.add.at(start, a + 1, 1)
    start = np.cumsum(start).astype(np.uint32)
    threads = oracle.max_threads()
    # calibrate on one MCS, then extend to about target_seconds
    probe = oracle.System(box, xyz0, start, b, colors, seed=seed)
    tic = time.perf_counter()
    probe.run(t0, n_colors, threads=threads)
    dt = time.perf_counter() - tic
    mcs = max(1, min(max_mcs, int(target_seconds / max(dt, 1e-3))))
    sysm = oracle.System(box, xyz0, start, b, colors, seed=seed)
    tic = time.perf_counter()
    _, olog = sysm.run(t0, mcs * n_colors, threads=threads)
    cpu_s = time.perf_counter() - tic
    attempted = 0
    sizes = np.bincount(colors, minlength=n_colors)
    for k in range(mcs * n_colors):
        s, _ = P.substep_draw(seed, t0 + k, n_colors)
        attempted += int(sizes[s])
    # GPU re-run of the same substeps from the same state
    _, glog = u.run(mcs, log=True)
    gxyz = u.getPositions(n)
    parity = bool(np.array_equal(glog, olog) and np.array_equal(gxyz, sysm.xyz))
    return {
        "value": attempted / cpu_s,
        "unit": "attempted monomer moves/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{mcs} MCS ({mcs * n_colors} substeps, {attempted} attempts) of the same melt state, "
                  f"colored two-pass oracle (oracle/oracle.c) with OpenMP, {cpu_s:.1f} s",
    }, parity, mcs


def cpu_single_core(u, box, bonds, colors, seed, target_seconds: float) -> dict:
    """The two single-core CPU figures beside the all-cores oracle: the colored
    oracle on one thread, and the sequential LeMonADE-style simulator (random
    monomer, random direction, apply at once; N attempts per MCS) on one
    thread.  Each runs whole MCS of the current melt state for about
    `target_seconds`."""
    import oracle

    n = colors.size
    xyz0 = u.getPositions(n)
    t0, _ = u.counters()
    n_colors = int(colors.max()) + 1
    a = np.concatenate([bonds[:, 0], bonds[:, 1]]).astype(np.int64)
    b = np.concatenate([bonds[:, 1], bonds[:, 0]]).astype(np.int64)
    order = np.argsort(a, kind="stable")
    a, b = a[order], b[order]
    start = np.zeros(n + 1, dtype=np.int64)
    np.add.at(start, a + 1, 1)
    start = np.cumsum(start).astype(np.uint32)
    sizes = np.bincount(colors, minlength=n_colors)
    out = {}
    # colored oracle, one thread: one substep to calibrate, then whole MCS
    probe = oracle.System(box, xyz0, start, b, colors, seed=seed)
    tic = time.perf_counter()
    probe.run(t0, 1, threads=1)
    per_substep = time.perf_counter() - tic
    mcs = max(1, int(target_seconds / max(per_substep * n_colors, 1e-3)))
    sysm = oracle.System(box, xyz0, start, b, colors, seed=seed)
    tic = time.perf_counter()
    sysm.run(t0, mcs * n_colors, threads=1)
    dt = time.perf_counter() - tic
    attempted = sum(int(sizes[P.substep_draw(seed, t0 + k, n_colors)[0]]) for k in range(mcs * n_colors))
    out["oracle_1_core"] = {"value": attempted / dt, "unit": "attempted monomer moves/s", "cores": 1,
                            "kind": "port", "sample": f"{mcs} MCS ({attempted} attempts), {dt:.1f} s"}
    # sequential simulator, one thread
    seq = oracle.System(box, xyz0, start, b, colors, seed=seed)
    tic = time.perf_counter()
    seq.sequential(seed, 1)
    per_mcs = time.perf_counter() - tic
    mcs = max(1, int(target_seconds / max(per_mcs, 1e-3)))
    seq = oracle.System(box, xyz0, start, b, colors, seed=seed)
    tic = time.perf_counter()
    acc = seq.sequential(seed, mcs)
    dt = time.perf_counter() - tic
    out["sequential_1_core"] = {"value": mcs * n / dt, "unit": "attempted monomer moves/s", "cores": 1,
                                "kind": "port", "acceptance": acc / (mcs * n),
                                "sample": f"{mcs} MCS ({mcs * n} attempts), {dt:.1f} s"}
    return out


def _oracle_of(sd, colors):
    import oracle

    n = sd.n
    pairs = W.directed_pairs(n, sd.bonds)
    start = np.concatenate([[0], np.cumsum(np.bincount(pairs[:, 0].astype(np.int64), minlength=n))]).astype(np.uint32)
    return oracle.System(sd.box, sd.xyz, start, pairs[:, 1].copy(), colors, seed=sd.seed)


def network_section(seed: int, mcs: int, kind: int, c_l: float, shape=None, clear_pass: int = 0,
                    spatial_order=None, dense: bool = True, after_tune=None) -> dict:
    """Star network of 12800 4-arm stars of arm length 29 (1,497,600
    monomers).  dense: on 256^3 (occupied fraction 0.714), built from an
    ordered packing of the stars, 100 sweeps of the stars on the GPU (checked
    against the oracle) and two rounds of arm-end crosslinks, four species.
    Else grown: stars grown as self-avoiding walks in a 512x256x256 box
    (occupied fraction 0.357, as dense as growth gets), one round of
    crosslinks, three species.  Times `mcs` sweeps and checks the first 3
    sweeps of a second copy against the oracle."""
    tic = time.perf_counter()
    relax_parity = None
    if dense:
        stars = W.dense_stars(seed)
        relax_mcs = W.DENSE_NETWORK["relax_mcs"]
        moved = W.relax_on_gpu(kind)(stars, relax_mcs)
        sd = W.crosslinked(stars, moved, W.DENSE_NETWORK["n_stars"], W.DENSE_NETWORK["arm_length"])
        built = time.perf_counter() - tic
        # the run that built it, on the oracle
        v = W.stage(stars, kind, initialize=True)
        o = _oracle_of(stars, v.colors(stars.n))
        v.cleanup()
        o.run(0, relax_mcs * o.n_colors)
        relax_parity = bool(np.array_equal(o.xyz, moved))
        log(f"dense network: stars' {relax_mcs}-sweep build run equal to the oracle's: {relax_parity}")
    else:
        xyz, bonds = P.gen_star_network((512, 256, 256), 12800, 4, 29, seed)
        sd = W.SystemData((512, 256, 256), xyz, bonds, seed, "saw")
        built = time.perf_counter() - tic
    box = sd.box
    u = W.stage(sd, kind, initialize=True, spatial_order=spatial_order)
    n = sd.n
    log(f"network: {n} monomers, {u.nColors()} species {u.colorSizes().tolist()}, "
        f"built in {built:.1f} s")
    if clear_pass:
        u.setClearPass(clear_pass)
    if shape:
        block, slots = shape
        trials = {}
        u.setLaunchShape(block, slots)
        log(f"network shape fixed: block {block} slots {slots}")
    else:
        block, slots, trials = u.autotune(3)
        log(f"network autotune: block {block} slots {slots}; "
            + ", ".join(f"{b}x{s_}:{ms:.3f}" for (b, s_), ms in sorted(trials.items())))
        if after_tune:
            # the tuner ran as usual; the network takes the stated shape
            # (profiled runs: the tuner's choice drifts under a profiler)
            block, slots = after_tune
            u.setLaunchShape(block, slots)
            log(f"network shape after tuning fixed at block {block} slots {slots}")
    u.run(20)
    st, _ = u.run(mcs)
    p1, p2 = algorithmic_bytes(st, c_l)
    value = st.attempted / (st.gpu_ms * 1e-3)
    gbs = (p1 + p2) / (st.gpu_ms * 1e-3) / 1e9
    log(f"network {mcs} MCS: {st.gpu_ms:.1f} ms GPU, {value / 1e9:.2f} G attempts/s, "
        f"{gbs:.0f} GB/s algorithmic ({gbs / HBM_PEAK_GBS:.2f} of peak)")
    u.cleanup()
    # bit-exact check of the first sweeps from the initial state
    v = W.stage(sd, kind, initialize=True)
    colors = v.colors(n)
    _, glog = v.run(3, log=True)
    gx = v.getPositions(n)
    v.cleanup()
    import oracle

    nc = int(colors.max()) + 1
    threads = oracle.max_threads()
    o = _oracle_of(sd, colors)
    tic = time.perf_counter()
    _, olog = o.run(0, 3 * nc, threads=threads)
    parity = bool(np.array_equal(gx, o.xyz) and np.array_equal(glog, olog))
    log(f"network parity over 3 MCS: {parity}")
    # the oracle continues from there: the CPU baseline of this workload
    cpu_mcs = 30
    o.run(3 * nc, (cpu_mcs - 3) * nc, threads=threads)
    cpu_s = time.perf_counter() - tic
    sizes = np.bincount(colors, minlength=nc)
    cpu_attempted = sum(int(sizes[P.substep_draw(seed, k, nc)[0]]) for k in range(cpu_mcs * nc))
    log(f"network CPU oracle: {cpu_attempted / cpu_s / 1e6:.1f} M attempts/s on {threads} threads")
    fraction = 8 * n / (box[0] * box[1] * box[2])
    if dense:
        what = (f"dense star network: 12800 4-arm stars (arm 29) = {n} monomers on {box[0]}^3 periodic, "
                f"occupied fraction {fraction:.3f}; ordered packing of the stars, "
                f"{W.DENSE_NETWORK['relax_mcs']} sweeps, then two rounds of arm-end crosslinks "
                f"({len(sd.bonds) - 12800 * 4 * 29} crosslinks)")
    else:
        what = (f"grown star network: 12800 4-arm stars (arm 29) = {n} monomers, arm ends crosslinked, "
                f"{box[0]}x{box[1]}x{box[2]} periodic, occupied fraction {fraction:.3f}")
    out = {
        "workload": f"{what}, {int(colors.max()) + 1} species, seed {seed}",
        "value": value,
        "unit": "attempted monomer moves/s",
        "mcs": mcs,
        "gpu_ms": st.gpu_ms,
        "sweep_algorithmic_GBps": gbs,
        "sweep_roofline_frac": gbs / HBM_PEAK_GBS,
        "acceptance": st.accepted / st.attempted,
        "shape": {"block": block, "slots": slots},
        "parity_vs_oracle": f"{'bit-exact' if parity else 'MISMATCH'} over 3 MCS",
    }
    out["cpu_baseline"] = {"value": cpu_attempted / cpu_s, "unit": "attempted monomer moves/s", "cores": threads,
                           "kind": "port",
                           "sample": f"the first {cpu_mcs} MCS ({cpu_attempted} attempts) of the network, colored "
                                     f"two-pass oracle (oracle/oracle.c) with OpenMP, {cpu_s:.1f} s"}
    if relax_parity is not None:
        out["build_run_vs_oracle"] = (f"{'bit-exact' if relax_parity else 'MISMATCH'} over "
                                      f"{W.DENSE_NETWORK['relax_mcs']} MCS of the stars")
    return out


def batched_vs_plain(n_rep: int, mcs: int, seed: int, kind: int, tune_sample: int, tune_mcs: int,
                     plain: bool = True, shape=(0, 0), eager_sample: int = 64, after_tune=None) -> dict:
    """Replica mix: one batch (autotuned shape, graph-replayed launches over
    every replica) against the plain loop of one updater per system, each
    called with runSimulationOnGPU(mcs) in turn.  Every replica's final
    coordinates are compared bit for bit between the two."""
    tic = time.perf_counter()
    specs = W.replica_specs(n_rep, seed)
    reps = W.build_replicas(specs)
    n_total = sum(r.n for r in reps)
    n_ordered = sum(r.method == "ordered" for r in reps)
    staged = [W.stage(r, kind) for r in reps]
    log(f"replica mix: {n_rep} systems, {n_total} monomers ({n_ordered} by ordered fallback), "
        f"built and staged in {time.perf_counter() - tic:.1f} s")

    if shape[0] and shape[1]:
        # a fixed shape (profiling runs pin it: the autotuner's timings shift
        # under a profiler)
        block, slots = shape
        trials = {}
        log(f"batch shape fixed: block {block} slots {slots}")
    else:
        rng = np.random.default_rng(seed + 1)
        pick = sorted(rng.choice(n_rep, size=min(tune_sample, n_rep), replace=False).tolist())
        tic = time.perf_counter()
        block, slots, trials = P.autotune_batch([staged[i] for i in pick], tune_mcs, kind)
        log(f"autotune on {len(pick)} replicas: block {block} slots {slots} ({time.perf_counter() - tic:.1f} s); "
            + ", ".join(f"{b}x{s_}:{ms:.3f}" for (b, s_), ms in sorted(trials.items())))
        if after_tune:
            # the tuner ran as usual; the batch takes the stated shape (profiled
            # runs: the tuner's choice drifts under a profiler)
            block, slots = after_tune
            log(f"batch shape after tuning fixed at block {block} slots {slots}")

    batch = P.Batch(0, n_rep, kind, block, slots)
    tic = time.perf_counter()
    ids = [batch.add(u) for u in staged]
    live, used_tiles, cap, dev_bytes = batch.usage()
    log(f"batch: {live} replicas, {used_tiles} workgroups, {dev_bytes / 2**20:.0f} MiB "
        f"(added in {time.perf_counter() - tic:.1f} s)")
    tic = time.perf_counter()
    batch_ms = batch.run(mcs)
    b_xyz = batch.positions_many(ids, [r.n for r in reps])
    batch_wall = time.perf_counter() - tic

    if not plain:  # batch only (profiling runs)
        attempted_est = sum(r.n for r in reps) * mcs
        log(f"batched {mcs} MCS: {batch_ms:.1f} ms GPU, {batch_wall * 1e3:.1f} ms wall incl. read-back")
        return {"value": attempted_est / batch_wall, "unit": "monomer sweeps/s (plain loop skipped)",
                "batch_ms": batch_ms, "shape": {"block": block, "slots": slots, "workgroups": used_tiles}}
    for u in staged:
        u.initialize()
    # The plain loop's graphs are built first, outside its timing: a caller
    # that runs each system repeatedly builds them once.
    for u in staged:
        u.prepareGraphs(mcs)
    tic = time.perf_counter()
    plain_ms = 0.0
    attempted = 0
    for u in staged:
        u.runSimulationOnGPU(mcs)
        st = u.lastRunStats()
        plain_ms += st.gpu_ms
        attempted += st.attempted
    plain_wall = time.perf_counter() - tic
    same = sum(bool(np.array_equal(u.getPositions(r.n), x)) for u, r, x in zip(staged, reps, b_xyz))
    accepted_same = sum(int(batch.info(i).accepted == u.counters()[1]) for i, u in zip(ids, staged))
    for u in staged:
        u.cleanup()
    # The plain loop with eager launches (one host launch per kernel, as the
    # reference issues them), on a sample of the replicas.
    n_eager = min(eager_sample, n_rep)
    eager_att = 0
    eager_same = 0
    eager_wall = 0.0
    for k in range(n_eager):
        u = W.stage(reps[k], kind, initialize=True)
        tic = time.perf_counter()
        st, _ = u.run(mcs, use_graph=False)
        x = u.getPositions(reps[k].n)
        eager_wall += time.perf_counter() - tic
        eager_att += st.attempted
        eager_same += int(np.array_equal(x, b_xyz[k]))
        u.cleanup()
    batched = attempted / batch_wall
    plain = attempted / plain_wall
    eager = eager_att / eager_wall if eager_wall > 0 else None
    log(f"batched {mcs} MCS: {batch_ms:.1f} ms GPU, {batch_wall * 1e3:.1f} ms wall incl. read-back "
        f"-> {batched / 1e9:.2f} G attempts/s")
    log(f"plain loop (graphs prebuilt): {plain_ms:.1f} ms GPU, {plain_wall * 1e3:.1f} ms wall -> "
        f"{plain / 1e9:.3f} G attempts/s; speedup {batched / plain:.1f}x; bit-identical {same}/{n_rep}, "
        f"accepted equal {accepted_same}/{n_rep}")
    if eager:
        log(f"plain loop, eager launches, first {n_eager} replicas: {eager_wall * 1e3:.1f} ms wall -> "
            f"{eager / 1e9:.3f} G attempts/s; speedup {batched / eager:.1f}x; bit-identical {eager_same}/{n_eager}")
    return {
        "workload": f"replica mix: {n_rep} systems, box edges log-uniform in {{32,64,128}} per axis, "
                    f"chains N ~ U[16,512], phi ~ U[0.1,0.5], SAW growth (ordered fallback for {n_ordered}), "
                    f"{n_total} monomers, {mcs} MCS each, per-replica seeds from {seed}",
        "value": batched,
        "unit": "attempted monomer moves/s (wall, incl. coordinate read-back)",
        "gpu_value": attempted / (batch_ms * 1e-3),
        "plain_value": plain,
        "plain_gpu_value": attempted / (plain_ms * 1e-3),
        "speedup_vs_plain": batched / plain,
        "plain_note": "each system's graphs built before the timed loop (prepareGraphs); runSimulationOnGPU "
                      "per system in turn, synchronous with copy-back",
        "plain_eager": {"value": eager, "unit": "attempted monomer moves/s (wall, incl. read-back)",
                        "replicas": n_eager, "speedup_vs": batched / eager if eager else None,
                        "parity": f"{eager_same}/{n_eager} bit-identical to the batch",
                        "note": "eager launches, one host launch per kernel"},
        "batch_ms": batch_ms,
        "plain_ms_wall": plain_wall * 1e3,
        "shape": {"block": block, "slots": slots, "workgroups": used_tiles},
        "autotune_ms_per_mcs": {f"{b}x{s_}": ms for (b, s_), ms in sorted(trials.items())},
        "parity": f"{same}/{n_rep} replicas bit-identical to their plain single-system runs "
                  f"(accepted counts equal for {accepted_same}/{n_rep})",
        "device_bytes": dev_bytes,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed MCS")
    ap.add_argument("--warmup", type=int, default=20, help="untimed MCS after the pre-relaxation")
    ap.add_argument("--prerelax", type=int, default=1000,
                    help="MCS of pre-relaxation after the ordered packing (part of building the workload)")
    ap.add_argument("--lattice", choices=["bit", "u8"], default="bit")
    ap.add_argument("--box", type=int, default=512)
    ap.add_argument("--chain", type=int, default=256)
    ap.add_argument("--phi", type=float, default=0.5)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--block", type=int, default=0, help="threads per workgroup (0: default)")
    ap.add_argument("--slots", type=int, default=0, help="monomers per pass-1 thread (0: default)")
    ap.add_argument("--perform-block", type=int, default=0, help="pass-2 threads per workgroup (0: default)")
    ap.add_argument("--graph-chunk", type=int, default=0, help="substeps per captured graph (0: default)")
    ap.add_argument("--list-group", type=int, default=0,
                    help="pass-2 threads per pass-1 list (0: default; -1: one list per workgroup)")
    ap.add_argument("--list-depth", type=int, default=0, help="pass-2 entries per thread in flight (0: default)")
    ap.add_argument("--slot-order", choices=["auto", "spatial", "reference"], default="auto",
                    help="slot order inside each species (results are the same)")
    ap.add_argument("--clear-pass", type=int, default=0, choices=[0, 1, 2],
                    help="pass that clears the other parity's proposals (0: default)")
    ap.add_argument("--profile-mcs", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="target seconds of the all-cores oracle sample (its calibration run overestimates, "
                         "so the sample lands at about half)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-single-seconds", type=float, default=5.0,
                    help="seconds for each single-core CPU figure (0 = skip them)")
    ap.add_argument("--replicas", type=int, default=1024, help="systems of the batched comparison")
    ap.add_argument("--replica-mcs", type=int, default=500)
    ap.add_argument("--tune-sample", type=int, default=128)
    ap.add_argument("--batch-block", type=int, default=0, help="fixed batch block (with --batch-slots; 0: autotune)")
    ap.add_argument("--batch-slots", type=int, default=0, help="fixed batch slots (with --batch-block; 0: autotune)")
    ap.add_argument("--tune-mcs", type=int, default=5)
    ap.add_argument("--no-plain", action="store_true", help="batched section without the plain per-system loop")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--network-mcs", type=int, default=2000)
    ap.add_argument("--network-shape", default="", help="BLOCKxSLOTS of the network (default: autotuned)")
    ap.add_argument("--network-shape-after-tune", default="",
                    help="BLOCKxSLOTS the networks use after the autotuner has run (default: the tuner's pick)")
    ap.add_argument("--batch-shape-after-tune", default="",
                    help="BLOCKxSLOTS the batch uses after the autotuner has run (default: the tuner's pick)")
    ap.add_argument("--no-network", action="store_true")
    ap.add_argument("--network", choices=("both", "dense", "grown"), default="both",
                    help="star networks to run: dense (256^3, occupied fraction 0.714) and/or grown "
                         "(512x256x256, 0.357)")
    ap.add_argument("--compare-mcs", type=int, default=300, help="MCS of the other lattice kind on the same melt")
    ap.add_argument("--compare-cpu-seconds", type=float, default=3.0,
                    help="target seconds of the oracle sample that checks the other lattice kind")
    ap.add_argument("--eager-sample", type=int, default=64,
                    help="replicas of the plain loop re-run with eager launches")
    ap.add_argument("--serve-rates", default="20,80,160,320",
                    help="Poisson arrival rates (jobs/s) of the served stream")
    ap.add_argument("--serve-seconds", type=float, default=8.0, help="arrival window per rate")
    ap.add_argument("--pressure-budget-gb", type=float, default=0.7,
                    help="HBM budget of the stream where admission by footprint binds (0: skip it)")
    ap.add_argument("--pressure-melts", type=int, default=6, help="512^3 melts in the budget-bound stream")
    ap.add_argument("--pressure-melt-mcs", type=int, default=20000, help="sweeps each of those melts asks for")
    ap.add_argument("--pressure-rate", type=float, default=20.0, help="arrival rate of the budget-bound stream")
    ap.add_argument("--no-serve", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: the sweep of one system is not sharded across GPUs "
              "(a z-slab decomposition with halo exchange would be the way; it is not implemented)",
              file=sys.stderr)
        return 2

    L = args.box
    box = (L, L, L)
    n_mono = int(round(args.phi * L ** 3 / 8))
    n_chains = n_mono // args.chain
    kind = P.LATTICE_BIT if args.lattice == "bit" else P.LATTICE_U8
    c_l = 0.125 if kind == P.LATTICE_BIT else 1.0

    log(f"building {n_chains} chains x {args.chain} on {L}^3 (ordered packing, phi={args.phi})")
    tic = time.perf_counter()
    xyz = P.gen_linear_chains_ordered(box, n_chains, args.chain)
    bonds = P.chain_bonds(n_chains, args.chain)
    u = make_updater(box, xyz, bonds, args.seed, kind, args.block, args.slots,
                     spatial_order=SLOT_ORDER[args.slot_order])
    if args.perform_block:
        u.setPerformBlock(args.perform_block)
    if args.graph_chunk:
        u.setGraphChunk(args.graph_chunk)
    if args.list_group or args.list_depth:
        u.setPerformWalk(P.LEGACY_WALK if args.list_group < 0 else args.list_group, args.list_depth)
    if args.clear_pass:
        u.setClearPass(args.clear_pass)
    colors = u.colors(xyz.shape[0])
    n = int(xyz.shape[0])
    log(f"initialize done in {time.perf_counter() - tic:.1f} s; species sizes {u.colorSizes().tolist()}; "
        f"device bytes {u.deviceBytes() / 2**20:.0f} MiB")

    # The graphs of the warm-up and the timed run first, then the
    # pre-relaxation (part of building the workload, whatever --warmup says)
    # and the warm-up back to back with the timed run: those two skip the
    # coordinate read-back, so the GPU runs without a gap of more than a host
    # synchronization up to the timed window (after an idle gap of a few ms
    # its first tens of MCS run up to 6% slower).
    u.prepareGraphs(args.steps)
    if args.warmup > 0:
        u.prepareGraphs(args.warmup)
    u.setReadBack(False)
    if args.prerelax > 0:
        st, _ = u.run(args.prerelax)
        log(f"pre-relax {args.prerelax} MCS: {st.gpu_ms:.1f} ms GPU, acceptance {st.accepted / max(st.attempted, 1):.4f}")
    if args.warmup > 0:
        st, _ = u.run(args.warmup)
        log(f"warm-up {args.warmup} MCS: {st.gpu_ms:.1f} ms GPU, acceptance {st.accepted / max(st.attempted, 1):.4f}")
    u.setReadBack(True)

    # timed region: K MCS of graph replays between two HIP events; the run
    # then copies the coordinates back (runSimulationOnGPU semantics)
    tic = time.perf_counter()
    st, _ = u.run(args.steps)
    wall = time.perf_counter() - tic
    timed_xyz = u.getPositions(n)
    gpu_s = st.gpu_ms / 1e3
    value = st.attempted / gpu_s
    p1, p2 = algorithmic_bytes(st, c_l)
    sweep_gbs = (p1 + p2) / gpu_s / 1e9
    log(f"timed {args.steps} MCS: {st.gpu_ms:.2f} ms GPU ({wall * 1e3:.1f} ms wall incl. copy-back), "
        f"{value / 1e9:.2f} G attempts/s, acceptance {st.accepted / st.attempted:.4f}, "
        f"sweep {sweep_gbs:.0f} GB/s algorithmic ({sweep_gbs / HBM_PEAK_GBS:.3f} of peak)")

    # per-kernel timing: HIP events around every launch of a separate sweep
    pst, _ = u.run(args.profile_mcs, profile=True)
    k1_mean_ms = pst.check_ms / pst.check_launches
    k2_mean_ms = pst.perform_ms / pst.perform_launches
    pp1, pp2 = algorithmic_bytes(pst, c_l)
    k1_bytes = pp1 / pst.check_launches
    k2_bytes = pp2 / pst.perform_launches
    log(f"pass 1 mean {k1_mean_ms * 1e3:.1f} us ({k1_bytes / (k1_mean_ms * 1e-3) / 1e9:.0f} GB/s algorithmic), "
        f"pass 2 mean {k2_mean_ms * 1e3:.1f} us ({k2_bytes / (k2_mean_ms * 1e-3) / 1e9:.0f} GB/s algorithmic)")

    # Plain streaming copy of 512 MiB: achievable-bandwidth reference, and the
    # kernel whose byte counters scale the profiler's traffic figures.
    copy_bytes = 1 << 29
    copy_ms = P.copy_probe(0, copy_bytes, 10)
    copy_gbs = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9
    log(f"copy reference: {copy_gbs:.0f} GB/s (read+write)")

    # Steady-state kernel durations and HBM bytes per launch from the
    # committed rocprofv3 runs of this workload (profiles/), when present.
    prof = {}
    pfile = ROOT / "profiles" / f"dense_melt_{L}_{args.lattice}.json"
    if pfile.exists() and args.chain == 256 and abs(args.phi - 0.5) < 1e-9:
        prof = json.loads(pfile.read_text())
    # the profile's names of this lattice's kernels
    names = {"k_check": "k_check" if args.lattice == "bit" else "k_check.u8",
             "k_perform": "k_perform" if args.lattice == "bit" else "k_perform.u8"}
    pdur = prof.get("durations", {})
    ptraf = (prof.get("traffic") or {}).get("kernels", {})
    pcount = prof.get("counters", {})

    def kernel_entry(key, bytes_per_launch, mean_ms):
        e = {"algorithmic_bytes_per_launch": bytes_per_launch, "mean_launch_us_events": mean_ms * 1e3,
             "achieved_GBps_events": bytes_per_launch / (mean_ms * 1e-3) / 1e9}
        e["frac_events"] = e["achieved_GBps_events"] / HBM_PEAK_GBS
        d = pdur.get(names[key])
        if d:
            e["mean_launch_us_rocprof"] = d["mean_us"]
            e["achieved_GBps_rocprof"] = bytes_per_launch / (d["mean_us"] * 1e-6) / 1e9
            e["frac_rocprof"] = e["achieved_GBps_rocprof"] / HBM_PEAK_GBS
        t = ptraf.get(names[key])
        e["traffic"] = t.get("traffic_bytes") if t else None
        if t and d:
            # measured HBM rate of the kernel: rocprof bytes over rocprof time
            e["hbm_GBps_rocprof"] = t["traffic_bytes"] / (d["mean_us"] * 1e-6) / 1e9
        c = pcount.get(names[key], {})
        if c.get("TCC_HIT_sum", 0) + c.get("TCC_MISS_sum", 0) > 0:
            e["l2_hit_rate_rocprof"] = c["TCC_HIT_sum"] / (c["TCC_HIT_sum"] + c["TCC_MISS_sum"])
        return e

    kernels = {"k_check (pass 1)": kernel_entry("k_check", k1_bytes, k1_mean_ms),
               "k_perform (pass 2)": kernel_entry("k_perform", k2_bytes, k2_mean_ms)}
    substeps_per_mcs = st.substeps / args.steps
    traffic = None
    t1 = kernels["k_check (pass 1)"]["traffic"]
    t2 = kernels["k_perform (pass 2)"]["traffic"]
    if t1 is not None and t2 is not None:
        traffic = t1 + t2  # one substep: one launch of each

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "attempted monomer moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": st.gpu_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"dense melt: {n_chains} linear chains of N={args.chain} ({n} monomers) "
                        f"on {L}^3 periodic, occupied-site fraction {8 * n / L**3:.3f}, "
                        f"ordered packing + {args.prerelax} MCS pre-relax, {u.nColors()} species, "
                        f"{args.lattice} Z-curve lattice, seed {args.seed}",
            "lattice": args.lattice,
            "monomers": n,
            "box": L,
            "prerelax_mcs": args.prerelax,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": sweep_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": sweep_gbs / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "sweep: one substep = one launch of k_check (pass 1) and one of k_perform (pass 2)",
            "algorithmic_bytes_per_substep": (p1 + p2) / st.substeps,
            "mean_substep_us": st.gpu_ms * 1e3 / st.substeps,
            "substeps_per_mcs": substeps_per_mcs,
            "definition": "achieved = algorithmic bytes of the timed MCS (pass 1: n(12+9c_L)+16D per substep, "
                          "pass 2: A(12+2c_L)) / their GPU time; traffic = rocprofv3 HBM bytes of one "
                          "launch of each kernel",
            "kernels": kernels,
        },
        "acceptance": st.accepted / st.attempted,
        "copy_reference_GBps": copy_gbs,
        "wall_ms_incl_copyback": wall * 1e3,
    }
    if args.compare_mcs > 0:
        # the same melt on the other occupancy lattice (uint8 vs 1-bit): after
        # the same MCS its state must equal this one's at the end of the timed
        # run (the lattice kind does not change a trajectory), a sample is
        # checked against the oracle, then timed
        other = "u8" if args.lattice == "bit" else "bit"
        okind = P.LATTICE_U8 if other == "u8" else P.LATTICE_BIT
        oc_l = 1.0 if other == "u8" else 0.125
        v = make_updater(box, xyz, bonds, args.seed, okind, args.block, args.slots,
                         spatial_order=SLOT_ORDER[args.slot_order])
        same_mcs = args.prerelax + args.warmup + args.steps
        v.run(same_mcs)
        cross = bool(np.array_equal(v.getPositions(n), timed_xyz))
        log(f"{other} lattice after {same_mcs} MCS: "
            f"{'identical to' if cross else 'DIFFERENT from'} the {args.lattice} lattice's state")
        del timed_xyz
        oparity = None
        if not args.no_cpu and args.compare_cpu_seconds > 0:
            _, oparity, omcs = cpu_baseline(v, box, bonds, colors, args.seed, args.compare_cpu_seconds, max_mcs=50)
            log(f"{other} lattice vs the oracle over {omcs} MCS: {oparity}")
        v.prepareGraphs(args.compare_mcs)
        v.setReadBack(False)
        v.run(20)
        v.setReadBack(True)
        ost, _ = v.run(args.compare_mcs)
        v.cleanup()
        op1, op2 = algorithmic_bytes(ost, oc_l)
        ogbs = (op1 + op2) / (ost.gpu_ms * 1e-3) / 1e9
        out["other_lattice"] = {
            "lattice": other, "mcs": args.compare_mcs, "value": ost.attempted / (ost.gpu_ms * 1e-3),
            "unit": "attempted monomer moves/s", "sweep_algorithmic_GBps": ogbs,
            "sweep_roofline_frac": ogbs / HBM_PEAK_GBS,
            "acceptance": ost.accepted / ost.attempted,
        }
        out["other_lattice"]["state_vs_main_lattice"] = (
            f"{'bit-identical' if cross else 'MISMATCH'} after {same_mcs} MCS ({n} monomers)")
        if oparity is not None:
            out["other_lattice"]["parity_vs_oracle"] = f"{'bit-exact' if oparity else 'MISMATCH'} over {omcs} MCS"
        log(f"{other} lattice: {out['other_lattice']['value'] / 1e9:.2f} G attempts/s, "
            f"sweep {ogbs:.0f} GB/s algorithmic ({ogbs / HBM_PEAK_GBS:.3f} of peak)")
    if not args.no_cpu:
        log("CPU baseline (oracle) on a bounded sample ...")
        cpu, parity, mcs = cpu_baseline(u, box, bonds, colors, args.seed, args.cpu_seconds)
        out["cpu_baseline"] = cpu
        out["parity_vs_oracle"] = f"{'bit-exact' if parity else 'MISMATCH'} over {mcs} MCS"
        log(f"CPU {cpu['value'] / 1e6:.1f} M attempts/s on {cpu['cores']} threads; parity {parity}")
        if args.cpu_single_seconds > 0:
            one = cpu_single_core(u, box, bonds, colors, args.seed, args.cpu_single_seconds)
            out["cpu_single_core"] = one
            log(f"CPU one core: oracle {one['oracle_1_core']['value'] / 1e6:.1f} M attempts/s, sequential "
                f"{one['sequential_1_core']['value'] / 1e6:.1f} M attempts/s "
                f"(acceptance {one['sequential_1_core']['acceptance']:.3f})")
    u.cleanup()
    if not args.no_network and args.network_mcs > 0:
        nshape = tuple(int(v) for v in args.network_shape.split("x")) if args.network_shape else None
        nafter = (tuple(int(v) for v in args.network_shape_after_tune.split("x"))
                  if args.network_shape_after_tune else None)
        if args.network in ("both", "dense"):
            out["network"] = network_section(args.seed, args.network_mcs, kind, c_l, nshape, args.clear_pass,
                                             SLOT_ORDER[args.slot_order], dense=True, after_tune=nafter)
        if args.network in ("both", "grown"):
            out["network_grown"] = network_section(args.seed, args.network_mcs, kind, c_l, nshape,
                                                   args.clear_pass, SLOT_ORDER[args.slot_order], dense=False,
                                                   after_tune=nafter)
    if not args.no_batch and args.replicas > 0:
        out["batched"] = batched_vs_plain(args.replicas, args.replica_mcs, args.seed, kind, args.tune_sample,
                                          args.tune_mcs, plain=not args.no_plain,
                                          shape=(args.batch_block, args.batch_slots),
                                          eager_sample=args.eager_sample,
                                          after_tune=tuple(int(v) for v in args.batch_shape_after_tune.split("x"))
                                          if args.batch_shape_after_tune else None)
    if not args.no_serve and args.serve_rates:
        from pscbfm_amd import serving
        out["serving"] = {}
        for k, rate in enumerate(float(r) for r in args.serve_rates.split(",")):
            jobs = serving.draw_jobs(rate, args.serve_seconds, args.seed + 100 + k)
            log(f"served stream at {rate:g} jobs/s: {len(jobs)} jobs "
                f"({sum(j.kind == 'melt' for j in jobs)} dense 512^3 melts)")
            res = serving.run_stream(jobs, log=log, seed=args.seed, spatial_order=SLOT_ORDER[args.slot_order])
            log(f"  {res['jobs_per_s']:.1f} jobs/s, latency p50 {res['latency_p50_s']:.2f} s "
                f"p99 {res['latency_p99_s']:.2f} s, queue wait p99 {res['queue_wait_p99_s']:.2f} s, GPU busy "
                f"{res['gpu_busy_fraction']:.2f}, peak resident {res['peak_resident_GB']:.2f} GB; {res['parity']} "
                f"(sweeps {res['verified_mcs']})")
            if res["melts"]:
                log(f"  melt jobs ({res['melt_mcs']} MCS) done after {[round(x, 2) for x in res['melt_latency_s']]} s; "
                    f"{res['melt_parity']}")
            plain = out.get("batched", {}).get("plain_value")
            if plain:
                res["speedup_vs_plain_loop"] = res["monomer_sweeps_per_s"] / plain
            out["serving"][f"{rate:g}/s"] = res
        if args.pressure_budget_gb > 0:
            # a stated budget small enough that admission by device footprint
            # binds: several long 512^3 melts (0.25 GB each) among the jobs, so
            # that a melt arriving while another is resident has to wait
            rate = args.pressure_rate
            jobs = serving.draw_jobs(rate, args.serve_seconds, args.seed + 99, melts=args.pressure_melts,
                                     melt_mcs=args.pressure_melt_mcs)
            budget = int(args.pressure_budget_gb * 1e9)
            log(f"served stream at {rate:g} jobs/s under a {args.pressure_budget_gb:g} GB budget: {len(jobs)} jobs "
                f"({sum(j.kind == 'melt' for j in jobs)} dense 512^3 melts)")
            res = serving.run_stream(jobs, hbm_budget=budget, log=log, seed=args.seed,
                                     spatial_order=SLOT_ORDER[args.slot_order])
            log(f"  {res['jobs_per_s']:.1f} jobs/s, latency p50 {res['latency_p50_s']:.2f} s "
                f"p99 {res['latency_p99_s']:.2f} s, queue wait p99 {res['queue_wait_p99_s']:.2f} s, "
                f"peak resident {res['peak_resident_GB']:.3f} of {args.pressure_budget_gb:g} GB; {res['parity']}; "
                f"{res['melt_parity']}")
            out["serving"][f"{rate:g}/s, {args.pressure_budget_gb:g} GB budget"] = res
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
