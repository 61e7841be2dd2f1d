#!/usr/bin/env python3
"""Benchmark of the scBFM sweep on one MI355X.

Workload (default): the dense melt, 32768 linear chains of N=256
(8,388,608 monomers) on a 512^3 periodic lattice at occupied-site fraction
0.5, 2 species.  A step is one Monte-Carlo step (MCS): every species drawn
once on average, i.e. n_colors substeps of pass 1 + pass 2.

    python bench.py [--steps K] [--warmup W] [--lattice bit|u8]

Prints ONE JSON line.  `value` = attempted monomer moves per second over the
K timed MCS (hipEvent time of the launched substeps, inputs resident in HBM).
Also reported:
  roofline      pass-1 kernel: algorithmic bytes per launch / mean launch time
                (HIP events around every pass-1 launch of a separate profiled
                sweep), against 8 TB/s
  cpu_baseline  the CPU oracle (same colored scheme, OpenMP over all host
                threads) on a bounded sample of the same state
  parity        the GPU re-run of the CPU sample, compared bit for bit
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

HBM_PEAK_GBS = 8000.0
METRIC = "attempted monomer moves/sec (1 GPU) + % HBM roofline; batched vs plain at bit-exact"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def undirected_to_directed(n: int, bonds: np.ndarray) -> np.ndarray:
    a = np.concatenate([bonds[:, 0], bonds[:, 1]]).astype(np.int64)
    b = np.concatenate([bonds[:, 1], bonds[:, 0]]).astype(np.int64)
    order = np.argsort(a, kind="stable")
    return np.stack([a[order], b[order]], axis=1).astype(np.uint32)


def standard_forbidden() -> np.ndarray:
    forb = np.ones(512, dtype=np.uint8)
    for v in P.standard_bond_vectors():
        dx, dy, dz = (int(c) for c in v)
        forb[(dx & 7) | ((dy & 7) << 3) | ((dz & 7) << 6)] = 0
    return forb


def make_updater(box, xyz, bonds, seed, kind, block=0, slots=0):
    u = P.UpdaterGPUScBFM_AB_Type()
    u.setGpu(0)
    u.setPeriodicity(True, True, True)
    u.setNrOfAllMonomers(xyz.shape[0])
    u.setMonomerCoordinatesBulk(xyz)
    u.setConnectivityBulk(undirected_to_directed(xyz.shape[0], bonds))
    u.setLatticeSize(*box)
    forb = standard_forbidden()
    for dx in range(-4, 4):
        for dy in range(-4, 4):
            for dz in range(-4, 4):
                u.copyBondSet(dx, dy, dz, bool(forb[(dx & 7) | ((dy & 7) << 3) | ((dz & 7) << 6)]))
    u.setSeed(seed)
    u.setLatticeKind(kind)
    u.setCheckOnInitialize(False)
    if block or slots:
        u.setLaunchShape(block, slots)
    u.initialize()
    return u, forb


def algorithmic_bytes(stats, lattice_bytes_per_cell: float):
    """Bytes per the metric definition: pass 1 = n(12 + 9 c_L) + 16 D per
    substep; pass 2 = A (12 + 2 c_L)."""
    p1 = stats.attempted * (12.0 + 9.0 * lattice_bytes_per_cell) + 16.0 * stats.bond_terms
    p2 = stats.accepted * (12.0 + 2.0 * lattice_bytes_per_cell)
    return p1, p2


def cpu_baseline(u, box, bonds, colors, seed, target_seconds: float):
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
    mcs = max(1, min(200, int(target_seconds / max(dt, 1e-3))))
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed MCS")
    ap.add_argument("--warmup", type=int, default=1000, help="untimed pre-relaxation MCS")
    ap.add_argument("--lattice", choices=["bit", "u8"], default="bit")
    ap.add_argument("--box", type=int, default=512)
    ap.add_argument("--chain", type=int, default=256)
    ap.add_argument("--phi", type=float, default=0.5)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--block", type=int, default=0, help="threads per workgroup (0: default)")
    ap.add_argument("--slots", type=int, default=0, help="monomers per pass-1 thread (0: default)")
    ap.add_argument("--profile-mcs", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: the sweep of one system is not sharded across GPUs "
              "in this build (see DESIGN.md for how it would be)", file=sys.stderr)
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
    u, _ = make_updater(box, xyz, bonds, args.seed, kind, args.block, args.slots)
    colors = u.colors(xyz.shape[0])
    log(f"initialize done in {time.perf_counter() - tic:.1f} s; species sizes {u.colorSizes().tolist()}; "
        f"device bytes {u.deviceBytes() / 2**20:.0f} MiB")

    if args.warmup > 0:
        tic = time.perf_counter()
        st, _ = u.run(args.warmup)
        log(f"pre-relax {args.warmup} MCS: {st.gpu_ms:.1f} ms GPU, acceptance {st.accepted / max(st.attempted, 1):.4f}")

    # timed region: K MCS, graph replays, synchronous at both ends
    tic = time.perf_counter()
    st, _ = u.run(args.steps)
    wall = time.perf_counter() - tic
    gpu_s = st.gpu_ms / 1e3
    value = st.attempted / gpu_s
    p1, p2 = algorithmic_bytes(st, c_l)
    log(f"timed {args.steps} MCS: {st.gpu_ms:.1f} ms GPU ({wall * 1e3:.1f} ms wall incl. copy-back), "
        f"{value / 1e9:.2f} G attempts/s, acceptance {st.accepted / st.attempted:.4f}")

    # per-kernel timing: HIP events around every launch of a separate sweep
    pst, _ = u.run(args.profile_mcs, profile=True)
    k1_mean_ms = pst.check_ms / pst.check_launches
    k2_mean_ms = pst.perform_ms / pst.perform_launches
    pp1, pp2 = algorithmic_bytes(pst, c_l)
    k1_bytes_per_launch = pp1 / pst.check_launches
    k1_gbs = k1_bytes_per_launch / (k1_mean_ms * 1e-3) / 1e9
    log(f"pass 1 mean {k1_mean_ms * 1e3:.1f} us ({k1_gbs:.0f} GB/s algorithmic), "
        f"pass 2 mean {k2_mean_ms * 1e3:.1f} us")

    # Plain streaming copy of 512 MiB: achievable-bandwidth reference, and the
    # kernel whose byte counters scale the profiler's traffic figures.
    copy_bytes = 1 << 29
    copy_ms = P.copy_probe(0, copy_bytes, 10)
    copy_gbs = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9
    log(f"copy reference: {copy_gbs:.0f} GB/s (read+write)")

    # HBM bytes per pass-1 launch, from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs of this workload (profiles/), when present.
    traffic = None
    tfile = ROOT / "profiles" / f"dense_melt_{L}_{args.lattice}.json"
    if tfile.exists() and args.chain == 256 and abs(args.phi - 0.5) < 1e-9:
        tr = json.loads(tfile.read_text()).get("traffic") or {}
        traffic = tr.get("kernels", {}).get("k_check", {}).get("traffic_bytes")

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
            "workload": f"dense melt: {n_chains} linear chains of N={args.chain} ({xyz.shape[0]} monomers) "
                        f"on {L}^3 periodic, occupied-site fraction {8 * xyz.shape[0] / L**3:.3f}, "
                        f"ordered packing + {args.warmup} MCS pre-relax, {u.nColors()} species, "
                        f"{args.lattice} Z-curve lattice, seed {args.seed}",
            "lattice": args.lattice,
            "monomers": int(xyz.shape[0]),
            "box": L,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": k1_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": k1_gbs / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_check (pass 1)",
            "bytes_per_launch": k1_bytes_per_launch,
            "mean_launch_us": k1_mean_ms * 1e3,
        },
        "acceptance": st.accepted / st.attempted,
        "sweep_algorithmic_GBps": (p1 + p2) / gpu_s / 1e9,
        "sweep_roofline_frac": (p1 + p2) / gpu_s / 1e9 / HBM_PEAK_GBS,
        "pass2_mean_us": k2_mean_ms * 1e3,
        "copy_reference_GBps": copy_gbs,
        "wall_ms_incl_copyback": wall * 1e3,
    }
    if not args.no_cpu:
        log("CPU baseline (oracle) on a bounded sample ...")
        cpu, parity, mcs = cpu_baseline(u, box, bonds, colors, args.seed, args.cpu_seconds)
        out["cpu_baseline"] = cpu
        out["parity_vs_oracle"] = f"{'bit-exact' if parity else 'MISMATCH'} over {mcs} MCS"
        log(f"CPU {cpu['value'] / 1e6:.1f} M attempts/s on {cpu['cores']} threads; parity {parity}")
    u.cleanup()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
