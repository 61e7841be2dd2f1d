"""Reader and writer of LeMonADE's BFM text format (the reference's I/O format).

The reference reads and writes these files through LeMonADE's
UpdaterReadBfmFile / AnalyzerWriteBfmFile (mainCUDASimulatorGPUScBFM_AB_Type.cpp:154-160);
the keys handled here are the ones its test files use (tests/*.bfm):

    !number_of_monomers=N
    !bonds                      extra bonds, "a b" (1-based), one per line
    !box_x= / !box_y= / !box_z=
    !periodic_x= / _y= / _z=
    !set_of_bondvectors         "dx dy dz:id", id = the byte that encodes it
    !attributes                 "a-b:value" ranges (1-based)
    !mcs=K                      strands: "x y z" then one byte per bond, each
                                byte the id of the vector to the next monomer

Strand lines are split positionally: after the z token one separating space,
then every remaining byte (spaces and control bytes included) is a bond.
"""
from __future__ import annotations

import gzip
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Tuple

import numpy as np


@dataclass
class BfmSystem:
    box: Tuple[int, int, int]
    periodic: Tuple[bool, bool, bool]
    xyz: np.ndarray                      # (n, 3) int32, unfolded, last !mcs block
    bonds: np.ndarray                    # (m, 2) uint32, undirected, 0-based
    bondset: Dict[int, Tuple[int, int, int]]
    attributes: np.ndarray               # (n,) int32 (0 where unset)
    mcs: int = 0
    frames: List[Tuple[int, np.ndarray]] = field(default_factory=list)

    @property
    def n(self) -> int:
        return int(self.xyz.shape[0])

    def adjacency(self):
        """Directed CSR (start, ids) with both directions of every bond, in
        the order: for each monomer its partners by increasing bond index."""
        n = self.n
        a = np.concatenate([self.bonds[:, 0], self.bonds[:, 1]]).astype(np.int64)
        b = np.concatenate([self.bonds[:, 1], self.bonds[:, 0]]).astype(np.int64)
        order = np.argsort(a, kind="stable")
        a, b = a[order], b[order]
        start = np.zeros(n + 1, dtype=np.uint32)
        np.add.at(start, a + 1, 1)
        start = np.cumsum(start).astype(np.uint32)
        return start, b.astype(np.uint32)


_SEEDS = [(2, 0, 0), (2, 1, 0), (2, -1, 0), (1, 2, 0), (1, -2, 0), (2, 1, 1), (2, -1, 1), (2, 1, -1), (2, -1, -1),
          (3, 0, 0), (2, 2, 1), (2, -2, 1), (2, 2, -1), (2, -2, -1), (3, 1, 0), (3, -1, 0), (1, 3, 0), (1, -3, 0)]


def lemonade_bond_ids() -> Dict[int, Tuple[int, int, int]]:
    """The ids LeMonADE writes for the 108 scBFM vectors (17..124): each of 18
    seed vectors (a,b,c) followed by (b,c,a), (c,a,b), (-a,b,c), (b,c,-a), (c,-a,b)."""
    out = {}
    k = 17
    for a, b, c in _SEEDS:
        for v in ((a, b, c), (b, c, a), (c, a, b), (-a, b, c), (b, c, -a), (c, -a, b)):
            out[k] = v
            k += 1
    return out


def _open(path: Path) -> bytes:
    data = Path(path).read_bytes()
    if data[:2] == b"\x1f\x8b":
        data = gzip.decompress(data)
    return data


def read_bfm(path) -> BfmSystem:
    lines = _open(path).split(b"\n")
    n = None
    box = [0, 0, 0]
    periodic = [True, True, True]
    extra: List[Tuple[int, int]] = []
    bondset: Dict[int, Tuple[int, int, int]] = {}
    attr_ranges: List[Tuple[int, int, int]] = []
    frames: List[Tuple[int, np.ndarray, List[Tuple[int, int]]]] = []
    i = 0
    while i < len(lines):
        raw = lines[i]
        line = raw.rstrip(b"\r")
        if line.startswith(b"#") or not line.strip():
            i += 1
            continue
        if line.startswith(b"!number_of_monomers="):
            n = int(line.split(b"=")[1])
        elif line.startswith(b"!box_"):
            axis = b"xyz".index(line[5:6])
            box[axis] = int(line.split(b"=")[1])
        elif line.startswith(b"!periodic_"):
            axis = b"xyz".index(line[10:11])
            periodic[axis] = int(line.split(b"=")[1]) != 0
        elif line == b"!bonds":
            i += 1
            while i < len(lines) and lines[i].strip() and not lines[i].startswith(b"!"):
                a, b = lines[i].split()
                extra.append((int(a) - 1, int(b) - 1))
                i += 1
            continue
        elif line == b"!set_of_bondvectors":
            i += 1
            while i < len(lines) and lines[i].strip() and not lines[i].startswith(b"!"):
                vec, ident = lines[i].rsplit(b":", 1)
                dx, dy, dz = (int(t) for t in vec.split())
                bondset[int(ident)] = (dx, dy, dz)
                i += 1
            continue
        elif line == b"!attributes":
            i += 1
            while i < len(lines) and lines[i].strip() and not lines[i].startswith(b"!"):
                rng, val = lines[i].split(b":")
                a, b = rng.split(b"-")
                attr_ranges.append((int(a) - 1, int(b) - 1, int(val)))
                i += 1
            continue
        elif line.startswith(b"!mcs="):
            age = int(line.split(b"=")[1].split()[0])
            coords: List[Tuple[int, int, int]] = []
            strand_bonds: List[Tuple[int, int]] = []
            i += 1
            while i < len(lines) and lines[i] and not lines[i].startswith(b"!"):
                s = lines[i]
                p1 = s.index(b" ")
                p2 = s.index(b" ", p1 + 1)
                p3 = s.find(b" ", p2 + 1)
                x, y = int(s[:p1]), int(s[p1 + 1:p2])
                if p3 < 0:
                    z, chars = int(s[p2 + 1:]), b""
                else:
                    z, chars = int(s[p2 + 1:p3]), s[p3 + 1:]
                coords.append((x, y, z))
                for ch in chars:
                    dx, dy, dz = bondset[ch]
                    px, py, pz = coords[-1]
                    coords.append((px + dx, py + dy, pz + dz))
                    strand_bonds.append((len(coords) - 2, len(coords) - 1))
                i += 1
            frames.append((age, np.asarray(coords, dtype=np.int32).reshape(-1, 3), strand_bonds))
            continue
        i += 1
    if not frames:
        raise ValueError(f"{path}: no !mcs block")
    age, xyz, strand_bonds = frames[-1]
    if n is not None and xyz.shape[0] != n:
        raise ValueError(f"{path}: {xyz.shape[0]} monomers in !mcs, header says {n}")
    all_bonds = strand_bonds + extra
    bonds = np.asarray(all_bonds, dtype=np.uint32).reshape(-1, 2)
    attributes = np.zeros(xyz.shape[0], dtype=np.int32)
    for a, b, v in attr_ranges:
        attributes[a:b + 1] = v
    return BfmSystem(box=tuple(box), periodic=tuple(periodic), xyz=xyz, bonds=bonds, bondset=bondset,
                     attributes=attributes, mcs=age, frames=[(f[0], f[1]) for f in frames])


def write_bfm(path, box, xyz: np.ndarray, bonds: np.ndarray, bondset: Optional[Dict[int, Tuple[int, int, int]]] = None,
              attributes: Optional[np.ndarray] = None, mcs: int = 0, periodic=(True, True, True)) -> None:
    """Writes one frame.  Bonds (i, i+1) whose vector is in the bond set are
    written inside strands; every other bond goes to !bonds."""
    if bondset is None:
        bondset = lemonade_bond_ids()
    code_of = {v: k for k, v in bondset.items()}
    xyz = np.asarray(xyz, dtype=np.int64).reshape(-1, 3)
    n = xyz.shape[0]
    bonds = np.asarray(bonds, dtype=np.int64).reshape(-1, 2)
    chain_link = np.zeros(n, dtype=bool)      # monomer i is bonded to i-1 inside a strand
    extra = []
    for a, b in bonds:
        lo, hi = (a, b) if a < b else (b, a)
        vec = tuple(int(c) for c in (xyz[hi] - xyz[lo]))
        if hi == lo + 1 and vec in code_of and not chain_link[hi]:
            chain_link[hi] = True
        else:
            extra.append((lo, hi))
    out = bytearray()
    out += b"#!version=2.0\n#Bond Fluctuation Model - simulation data file\n#monomer numbering starts at 1\n\n"
    out += f"!number_of_monomers={n}\n\n!bonds\n".encode()
    for a, b in extra:
        out += f"{a + 1} {b + 1}\n".encode()
    out += b"\n"
    for axis, L in zip("xyz", box):
        out += f"!box_{axis}={int(L)}\n\n".encode()
    for axis, p in zip("xyz", periodic):
        out += f"!periodic_{axis}={int(bool(p))}\n\n".encode()
    out += b"!set_of_bondvectors\n"
    for k in sorted(bondset):
        dx, dy, dz = bondset[k]
        out += f"{dx} {dy} {dz}:{k}\n".encode()
    out += b"\n"
    if attributes is not None:
        out += b"!attributes\n"
        attr = np.asarray(attributes)
        start = 0
        for i in range(1, n + 1):
            if i == n or attr[i] != attr[start]:
                out += f"{start + 1}-{i}:{int(attr[start])}\n".encode()
                start = i
        out += b"\n"
    out += f"\n!mcs={int(mcs)}\n".encode()
    i = 0
    while i < n:
        out += f"{int(xyz[i, 0])} {int(xyz[i, 1])} {int(xyz[i, 2])}".encode()
        j = i + 1
        chars = bytearray()
        while j < n and chain_link[j]:
            chars.append(code_of[tuple(int(c) for c in (xyz[j] - xyz[j - 1]))])
            j += 1
        if chars:
            out += b" " + bytes(chars)
        out += b"\n"
        i = j
    out += b"\n"
    Path(path).write_bytes(bytes(out))
