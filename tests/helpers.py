"""Test helpers: build systems, drive the GPU updater through the reference's
call order, and run the same substeps on the CPU oracle."""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np

import oracle
import pscbfm_amd as P

GOLDEN = Path(__file__).resolve().parent / "golden"


@dataclass
class Sys:
    box: tuple
    xyz: np.ndarray          # (n,3) int32 unfolded
    bonds: np.ndarray        # (m,2) undirected

    @property
    def n(self) -> int:
        return int(self.xyz.shape[0])

    def csr(self):
        n = self.n
        if self.bonds.size == 0:
            return np.zeros(n + 1, np.uint32), np.zeros(0, np.uint32)
        a = np.concatenate([self.bonds[:, 0], self.bonds[:, 1]]).astype(np.int64)
        b = np.concatenate([self.bonds[:, 1], self.bonds[:, 0]]).astype(np.int64)
        order = np.argsort(a, kind="stable")
        a, b = a[order], b[order]
        start = np.zeros(n + 1, dtype=np.int64)
        np.add.at(start, a + 1, 1)
        return np.cumsum(start).astype(np.uint32), b.astype(np.uint32)

    def directed_pairs(self) -> np.ndarray:
        start, ids = self.csr()
        src = np.repeat(np.arange(self.n, dtype=np.uint32), np.diff(start.astype(np.int64)))
        return np.stack([src, ids], axis=1).astype(np.uint32)


def chains(box, n_chains, length, seed) -> Sys:
    xyz = P.gen_linear_chains(box, n_chains, length, seed)
    return Sys(tuple(box), xyz, P.chain_bonds(n_chains, length))


def ordered_chains(box, n_chains, length) -> Sys:
    xyz = P.gen_linear_chains_ordered(box, n_chains, length)
    return Sys(tuple(box), xyz, P.chain_bonds(n_chains, length))


def stars(box, n_stars, arms, arm_length, seed) -> Sys:
    xyz, bonds = P.gen_star_network(box, n_stars, arms, arm_length, seed)
    return Sys(tuple(box), xyz, bonds)


def oracle_relax(sd, mcs: int) -> np.ndarray:
    """A relax(system, mcs) for the workload builders on the CPU oracle, with
    the coloring the updater computes; returns the unfolded corners."""
    s = Sys(tuple(sd.box), sd.xyz, sd.bonds)
    start, ids = s.csr()
    colors = P.color_graph(s.n, start, ids)
    o = oracle.System(s.box, s.xyz, start, ids, colors, forbidden_table(), sd.seed)
    o.run(0, int(mcs) * (int(colors.max()) + 1))
    return o.xyz


def forbidden_table() -> np.ndarray:
    return oracle.standard_bondset()


def stage_updater(s: Sys, seed: int, kind: int = P.LATTICE_BIT, colors=None, gpu: int = 0, bulk: bool = True,
                  block: int = 0, slots: int = 0, spatial=None) -> P.UpdaterGPUScBFM_AB_Type:
    """Stages a system in the call order of GPUScBFM_AB_Type.h:83-130, up to
    (not including) initialize.  spatial: slot order inside each species
    (True: Z-curve order, False: monomer order, None: the library's
    automatic choice)."""
    u = P.UpdaterGPUScBFM_AB_Type()
    u.setGpu(gpu)
    u.setPeriodicity(True, True, True)
    u.setNrOfAllMonomers(s.n)
    if bulk:
        u.setMonomerCoordinatesBulk(s.xyz)
    else:
        for i, (x, y, z) in enumerate(s.xyz):
            u.setMonomerCoordinates(i, int(x), int(y), int(z))
    for i in range(min(s.n, 4)):
        u.setAttribute(i, 1 + (i & 1))
    u.setConnectivityBulk(s.directed_pairs())
    u.setLatticeSize(*s.box)
    u.populateLattice()
    forb = forbidden_table()
    for dx in range(-4, 4):
        for dy in range(-4, 4):
            for dz in range(-4, 4):
                u.copyBondSet(dx, dy, dz, bool(forb[oracle.bond_index(dx, dy, dz)]))
    u.setSeed(seed)
    u.setLatticeKind(kind)
    if colors is not None:
        u.setColors(colors)
    if block or slots:
        u.setLaunchShape(block, slots)
    u.setSlotOrder(spatial)
    return u


def make_updater(s: Sys, seed: int, kind: int = P.LATTICE_BIT, colors=None, gpu: int = 0, bulk: bool = True,
                 block: int = 0, slots: int = 0, spatial=None) -> P.UpdaterGPUScBFM_AB_Type:
    """Configures and initializes an updater (GPUScBFM_AB_Type.h:83-130)."""
    u = stage_updater(s, seed, kind, colors, gpu, bulk, block, slots, spatial)
    u.initialize()
    return u


def oracle_for(s: Sys, colors, seed: int) -> oracle.System:
    start, ids = s.csr()
    return oracle.System(s.box, s.xyz, start, ids, colors, forbidden_table(), seed)


def gpu_vs_oracle(s: Sys, seed: int, n_mcs: int, kind: int = P.LATTICE_BIT, chunks=(None,), use_graph=True,
                  block: int = 0, slots: int = 0, spatial=None):
    """Runs n_mcs sweeps on the GPU (split into `chunks` calls) and on the
    oracle; returns (gpu_xyz, gpu_log, oracle_xyz, oracle_log)."""
    u = make_updater(s, seed, kind, block=block, slots=slots, spatial=spatial)
    colors = u.colors(s.n)
    logs = []
    plan = [n_mcs] if chunks == (None,) else list(chunks)
    assert sum(plan) == n_mcs
    for m in plan:
        _, log = u.run(m, log=True, use_graph=use_graph)
        logs.append(log)
    gx = u.getPositions(s.n)
    glog = np.concatenate(logs) if logs else np.zeros(0, np.uint32)
    o = oracle_for(s, colors, seed)
    _, olog = o.run(0, n_mcs * (int(colors.max()) + 1))
    u.cleanup()
    return gx, glog, o.xyz, olog
