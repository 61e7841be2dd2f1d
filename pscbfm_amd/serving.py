"""Served stream of equilibration jobs over the request loop (Server).

Jobs are drawn like the replica mix (workloads.replica_specs).  `melts` of
them, at arrivals picked at random, are instead a dense 512^3 melt (32768
chains of 256 at occupied fraction 0.5, ordered packing) asking for
`melt_mcs` sweeps.  Every other job asks for a number of sweeps drawn
log-uniformly from [mcs_lo, mcs_hi].  Arrivals are a Poisson process of the
given rate; a submitter thread hands each job to the server at its arrival
time (the staging of the job's updater is done beforehand, so the arrival
process is not delayed by host work).  Reported per rate: completed jobs/s,
p50/p99 completion latency (submit -> result ready), p50/p99 queueing delay,
monomer-sweeps/s, the GPU busy fraction and the peak resident device bytes
against the budget; a random sample of the jobs, the longest jobs and every
melt are re-run alone in an updater and compared bit for bit.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import LATTICE_BIT, Server
from . import workloads as W

MELT_BOX = (512, 512, 512)
MELT_CHAINS, MELT_LEN = 32768, 256


@dataclass
class Job:
    system: W.SystemData
    mcs: int
    arrival: float      # seconds after the stream starts
    kind: str           # "replica" or "melt"


def draw_jobs(rate: float, duration: float, seed: int, mcs_lo: int = 1000, mcs_hi: int = 100000,
              melts: int = 1, melt_mcs: int = 1000) -> List[Job]:
    rng = np.random.default_rng(seed)
    times = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        times.append(t)
    specs = W.replica_specs(len(times), seed + 7)
    systems = W.build_replicas(specs)
    melt_at = set(rng.choice(len(times), size=min(melts, len(times)), replace=False).tolist()) if times else set()
    jobs = []
    for k, (at, sd) in enumerate(zip(times, systems)):
        mcs = int(round(math.exp(rng.uniform(math.log(mcs_lo), math.log(mcs_hi)))))
        kind = "replica"
        if k in melt_at:
            from . import chain_bonds, gen_linear_chains_ordered
            xyz = gen_linear_chains_ordered(MELT_BOX, MELT_CHAINS, MELT_LEN)
            sd = W.SystemData(MELT_BOX, xyz, chain_bonds(MELT_CHAINS, MELT_LEN), int(rng.integers(1, 2**62)),
                              "ordered")
            kind = "melt"
            mcs = melt_mcs
        jobs.append(Job(sd, mcs, at, kind))
    return jobs


def run_stream(jobs: List[Job], hbm_budget: int = 288 * 10**9, tile_capacity: int = 32768,
               replica_capacity: int = 8192, block: int = 256, slots: int = 2, quantum: int = 64,
               lattice_kind: int = LATTICE_BIT, verify: int = 8, verify_longest: int = 2, seed: int = 0,
               log=None, spatial_order=None) -> dict:
    staged = [W.stage(j.system, lattice_kind, spatial_order=spatial_order) for j in jobs]
    srv = Server(hbm_budget, tile_capacity, replica_capacity, lattice_kind, block, slots, quantum)
    ids: List[Optional[int]] = [None] * len(jobs)
    t_start = time.perf_counter()

    def submitter():
        for k, (job, u) in enumerate(zip(jobs, staged)):
            delay = job.arrival - (time.perf_counter() - t_start)
            if delay > 0:
                time.sleep(delay)
            ids[k] = srv.submit(u, job.mcs)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    if log:
        log(f"  all {len(jobs)} jobs submitted after {time.perf_counter() - t_start:.1f} s; draining")
    srv.drain()
    infos = [srv.job(i) for i in ids]
    st = srv.stats()
    lat = np.array([inf.t_done - inf.t_submit for inf in infos])
    wait = np.array([inf.t_admit - inf.t_submit for inf in infos])
    first = min(inf.t_submit for inf in infos)
    last = max(inf.t_done for inf in infos)
    span = last - first
    sweeps = sum(j.system.n * j.mcs for j in jobs)
    # bit-exact check against the same system run alone: a random sample of
    # the replica jobs plus the longest ones (up to the 10^5-sweep end of the
    # range), and every melt below
    rng = np.random.default_rng(seed)
    reps = [k for k, j in enumerate(jobs) if j.kind == "replica"]
    longest = sorted(reps, key=lambda k: -jobs[k].mcs)[:verify_longest]
    rest = [k for k in reps if k not in longest]
    pick = sorted(rng.choice(rest, size=min(verify, len(rest)), replace=False).tolist()) if rest else []
    pick = sorted(set(pick) | set(longest))
    def alone_equals_served(k: int) -> bool:
        u = W.stage(jobs[k].system, lattice_kind, initialize=True)
        u.runSimulationOnGPU(jobs[k].mcs)
        ok = np.array_equal(u.getPositions(jobs[k].system.n), srv.result(ids[k], jobs[k].system.n))
        u.cleanup()
        return bool(ok)

    same = sum(alone_equals_served(k) for k in pick)
    melt_ids = [k for k, j in enumerate(jobs) if j.kind == "melt"]
    melt_same = sum(alone_equals_served(k) for k in melt_ids)
    srv.close()
    return {
        "jobs": len(jobs),
        "melts": sum(j.kind == "melt" for j in jobs),
        "jobs_per_s": len(jobs) / span,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "queue_wait_p50_s": float(np.percentile(wait, 50)),
        "queue_wait_p99_s": float(np.percentile(wait, 99)),
        "monomer_sweeps_per_s": sweeps / span,
        "gpu_busy_fraction": st.gpu_busy_seconds / span,
        "peak_resident_GB": st.peak_resident_bytes / 1e9,
        "hbm_budget_GB": hbm_budget / 1e9,
        "launches": st.launches,
        "substeps": st.substeps,
        "parity": f"{same}/{len(pick)} sampled jobs bit-identical to their runs alone",
        "verified_mcs": sorted(jobs[k].mcs for k in pick),
        "melt_latency_s": [float(lat[k]) for k in melt_ids],
        "melt_mcs": [jobs[k].mcs for k in melt_ids],
        "melt_parity": f"{melt_same}/{len(melt_ids)} melts bit-identical to their runs alone",
    }
