"""Synthetic scBFM workloads and the staging of a system into an updater.

Replica mix (batched and served runs): each box edge drawn log-uniformly
from {32, 64, 128}, chain length N uniform in [16, 512], occupied-site
fraction phi uniform in [0.1, 0.5], round(phi * V / (8 N)) chains (at least
one), a per-replica seed.  Chains are grown as seeded self-avoiding walks
with bond vectors drawn uniformly from the 108 set; a box too dense for the
growth falls back to the ordered serpentine packing (recorded per replica).

Dense star network: 12,800 four-arm stars of arm 29 (1,497,600 monomers)
on 256^3, occupied fraction 0.714, beyond what seeded growth reaches.  The
stars start on an ordered packing of the sites with even coordinates, run
100 sweeps as a star system, and then get two rounds of crosslinks from
arm ends to other stars' arm interiors at valid bond vectors.  On the
packing every bond joins sites of opposite parity, so the stars alone have
two species; the crosslinks made after the run close odd cycles, and the
reference's coloring splits the network into four species.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import (LATTICE_BIT, UpdaterGPUScBFM_AB_Type, chain_bonds, crosslink_star_ends, gen_linear_chains,
               gen_linear_chains_ordered, gen_stars_ordered, standard_bond_vectors)

EDGES = (32, 64, 128)

# The dense star network's build parameters.
DENSE_NETWORK = {"box": (256, 256, 256), "n_stars": 12800, "arm_length": 29, "relax_mcs": 100, "rounds": 2}


@dataclass
class ReplicaSpec:
    box: tuple
    length: int        # monomers per chain
    phi: float         # requested occupied-site fraction
    n_chains: int
    seed: int          # growth and sweep seed

    @property
    def n_monomers(self) -> int:
        return self.n_chains * self.length


@dataclass
class SystemData:
    box: tuple
    xyz: np.ndarray                       # (n, 3) int32, unfolded corners
    bonds: np.ndarray                     # (m, 2) uint32, undirected
    seed: int
    method: str = "saw"                   # "saw" or "ordered"
    spec: Optional[ReplicaSpec] = field(default=None, repr=False)

    @property
    def n(self) -> int:
        return int(self.xyz.shape[0])


def replica_specs(count: int, seed: int, edges=EDGES, n_range=(16, 512), phi_range=(0.1, 0.5)) -> List[ReplicaSpec]:
    rng = np.random.default_rng(seed)
    out = []
    for r in range(count):
        box = tuple(int(rng.choice(edges)) for _ in range(3))
        length = int(rng.integers(n_range[0], n_range[1] + 1))
        phi = float(rng.uniform(*phi_range))
        vol = box[0] * box[1] * box[2]
        n_chains = max(1, int(round(phi * vol / (8 * length))))
        out.append(ReplicaSpec(box, length, phi, n_chains, int(rng.integers(1, 2**63))))
    return out


def build_chains(spec: ReplicaSpec, max_restarts: int = 200) -> SystemData:
    try:
        xyz = gen_linear_chains(spec.box, spec.n_chains, spec.length, spec.seed, max_restarts)
        method = "saw"
    except RuntimeError:
        xyz = gen_linear_chains_ordered(spec.box, spec.n_chains, spec.length)
        method = "ordered"
    return SystemData(spec.box, xyz, chain_bonds(spec.n_chains, spec.length), spec.seed, method, spec)


def build_replicas(specs: List[ReplicaSpec], workers: int = 0) -> List[SystemData]:
    """Builds every replica (the generators release the GIL, so threads help)."""
    workers = workers or min(16, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(build_chains, specs))


def dense_stars(seed: int, box=DENSE_NETWORK["box"], n_stars: int = DENSE_NETWORK["n_stars"],
                arm_length: int = DENSE_NETWORK["arm_length"]) -> SystemData:
    """The dense network's stars on their ordered packing, without crosslinks."""
    xyz, bonds = gen_stars_ordered(box, n_stars, arm_length, seed)
    return SystemData(tuple(box), xyz, bonds, seed, "ordered")


def crosslinked(stars: SystemData, moved: np.ndarray, n_stars: int, arm_length: int,
                rounds: int = DENSE_NETWORK["rounds"]) -> SystemData:
    """The stars at the corners `moved` (a run of `stars`), with `rounds`
    rounds of arm-end crosslinks made at those corners."""
    moved = np.ascontiguousarray(moved, dtype=np.int32)
    cross = crosslink_star_ends(stars.box, moved, n_stars, 4, arm_length, stars.seed, rounds)
    return SystemData(stars.box, moved, np.concatenate([stars.bonds, cross]), stars.seed, "ordered+run")


def dense_network(seed: int, relax, box=DENSE_NETWORK["box"], n_stars: int = DENSE_NETWORK["n_stars"],
                  arm_length: int = DENSE_NETWORK["arm_length"], relax_mcs: int = DENSE_NETWORK["relax_mcs"],
                  rounds: int = DENSE_NETWORK["rounds"]) -> SystemData:
    """The dense star network.  relax(system, mcs) runs the star system and
    returns its unfolded corners (relax_on_gpu, or any bit-exact runner)."""
    stars = dense_stars(seed, box, n_stars, arm_length)
    return crosslinked(stars, relax(stars, relax_mcs), n_stars, arm_length, rounds)


def relax_on_gpu(kind: int = LATTICE_BIT, gpu: int = 0):
    """A relax(system, mcs) that runs the system in an updater on `gpu`."""
    def run(sd: SystemData, mcs: int) -> np.ndarray:
        u = stage(sd, kind, gpu=gpu, initialize=True)
        try:
            u.runSimulationOnGPU(int(mcs))
            return u.getPositions(sd.n)
        finally:
            u.cleanup()
    return run


def standard_forbidden() -> np.ndarray:
    """512-entry table (1 = forbidden) indexed by (dx&7)|(dy&7)<<3|(dz&7)<<6."""
    forb = np.ones(512, dtype=np.uint8)
    for v in standard_bond_vectors():
        dx, dy, dz = (int(c) for c in v)
        forb[(dx & 7) | ((dy & 7) << 3) | ((dz & 7) << 6)] = 0
    return forb


def directed_pairs(n: int, bonds: np.ndarray) -> np.ndarray:
    """Both directions of every undirected bond, grouped by source monomer."""
    if bonds.size == 0:
        return np.zeros((0, 2), dtype=np.uint32)
    a = np.concatenate([bonds[:, 0], bonds[:, 1]]).astype(np.int64)
    b = np.concatenate([bonds[:, 1], bonds[:, 0]]).astype(np.int64)
    order = np.argsort(a, kind="stable")
    return np.stack([a[order], b[order]], axis=1).astype(np.uint32)


_FORBIDDEN = None


def stage(sd: SystemData, kind: int = LATTICE_BIT, gpu: int = 0, block: int = 0, slots: int = 0,
          check: bool = False, initialize: bool = False, spatial_order=None) -> UpdaterGPUScBFM_AB_Type:
    """Stages a system in the reference's call order (GPUScBFM_AB_Type.h:83-130);
    initialize() only if asked (a batch takes the staged system as it is).
    spatial_order: slot order of a species, True (Z-curve order), False (the
    reference's monomer order) or None (the library's automatic choice);
    results are the same."""
    global _FORBIDDEN
    if _FORBIDDEN is None:
        _FORBIDDEN = standard_forbidden()
    u = UpdaterGPUScBFM_AB_Type()
    u.setGpu(gpu)
    u.setPeriodicity(True, True, True)
    u.setNrOfAllMonomers(sd.n)
    u.setMonomerCoordinatesBulk(sd.xyz)
    u.setConnectivityBulk(directed_pairs(sd.n, sd.bonds))
    u.setLatticeSize(*sd.box)
    for dx in range(-4, 4):
        for dy in range(-4, 4):
            for dz in range(-4, 4):
                u.copyBondSet(dx, dy, dz, bool(_FORBIDDEN[(dx & 7) | ((dy & 7) << 3) | ((dz & 7) << 6)]))
    u.setSeed(sd.seed)
    u.setLatticeKind(kind)
    u.setCheckOnInitialize(check)
    u.setSlotOrder(spatial_order)
    if block or slots:
        u.setLaunchShape(block, slots)
    if initialize:
        u.initialize()
    return u
