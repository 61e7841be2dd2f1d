"""ctypes wrapper of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, never by the pscbfm_amd library.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path
from typing import Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_COLORING_PATH = HERE / "_ref" / "libref_coloring.so"

_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class OracleSystem(C.Structure):
    _fields_ = [
        ("box", C.c_uint32 * 3), ("n", C.c_uint32), ("xyz", _i32p), ("nbr_start", _u32p), ("nbr", _u32p),
        ("color", _u8p), ("n_colors", C.c_uint32), ("forbidden", _u8p), ("seed", C.c_uint64),
        ("draws", _u32p), ("draw_t0", C.c_uint64),
    ]


class R250State(C.Structure):
    _fields_ = [("buf", C.c_uint32 * 250), ("index", C.c_uint32)]


_LIB: Optional[C.CDLL] = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        h.oracle_hash.restype = C.c_uint32
        h.oracle_hash.argtypes = [C.c_uint32]
        h.oracle_splitmix64.restype = C.c_uint64
        h.oracle_splitmix64.argtypes = [C.c_uint64, C.c_uint64]
        h.oracle_substep_draw.restype = None
        h.oracle_substep_draw.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, _u32p, _u32p]
        h.oracle_bond_index.restype = C.c_uint32
        h.oracle_bond_index.argtypes = [C.c_int32] * 3
        h.oracle_standard_bondset.restype = C.c_uint32
        h.oracle_standard_bondset.argtypes = [_u8p]
        h.oracle_dilute_bits.restype = C.c_uint32
        h.oracle_dilute_bits.argtypes = [C.c_uint32]
        h.oracle_zcurve_index.restype = C.c_uint32
        h.oracle_zcurve_index.argtypes = [C.c_uint32] * 6
        h.oracle_direction.restype = None
        h.oracle_direction.argtypes = [C.c_uint32, _i32p, _i32p, _i32p]
        h.oracle_front_blocked.restype = C.c_int
        h.oracle_front_blocked.argtypes = [_u8p, _u32p, _i32p, C.c_uint32]
        h.oracle_run.restype = C.c_int64
        h.oracle_run.argtypes = [C.POINTER(OracleSystem), C.c_uint64, C.c_uint64, _u32p, C.c_int, C.c_uint64]
        h.oracle_run_naive.restype = C.c_int64
        h.oracle_run_naive.argtypes = [C.POINTER(OracleSystem), C.c_uint64, C.c_uint64, _u32p]
        h.oracle_check_system.restype = C.c_int
        h.oracle_check_system.argtypes = [C.POINTER(OracleSystem)]
        h.oracle_sequential_run.restype = C.c_int64
        h.oracle_sequential_run.argtypes = [C.POINTER(OracleSystem), C.c_uint64, C.c_uint64]
        h.oracle_max_threads.restype = C.c_int
        h.oracle_max_threads.argtypes = []
        h.oracle_r250_load.restype = None
        h.oracle_r250_load.argtypes = [C.POINTER(R250State), _u32p]
        h.oracle_r250_rand32.restype = C.c_uint32
        h.oracle_r250_rand32.argtypes = [C.POINTER(R250State)]
        _LIB = h
    return _LIB


def hash32(a: int) -> int:
    return int(lib().oracle_hash(a & 0xFFFFFFFF))


def splitmix64(seed: int, t: int) -> int:
    return int(lib().oracle_splitmix64(seed & 0xFFFFFFFFFFFFFFFF, t))


def substep_draw(seed: int, t: int, n_colors: int):
    s, k = C.c_uint32(), C.c_uint32()
    lib().oracle_substep_draw(seed & 0xFFFFFFFFFFFFFFFF, t, n_colors, C.byref(s), C.byref(k))
    return s.value, k.value


def standard_bondset() -> np.ndarray:
    out = np.empty(512, dtype=np.uint8)
    n = lib().oracle_standard_bondset(out.ctypes.data_as(_u8p))
    assert n == 108
    return out


def bond_index(dx: int, dy: int, dz: int) -> int:
    return int(lib().oracle_bond_index(dx, dy, dz))


def zcurve_index(x, y, z, lx, ly, lz) -> int:
    return int(lib().oracle_zcurve_index(x, y, z, lx, ly, lz))


def front_blocked(occ: np.ndarray, box, p, d: int) -> bool:
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    b = np.asarray(box, dtype=np.uint32)
    pp = np.asarray(p, dtype=np.int32)
    return bool(lib().oracle_front_blocked(occ.ctypes.data_as(_u8p), b.ctypes.data_as(_u32p),
                                           pp.ctypes.data_as(_i32p), d))


def max_threads() -> int:
    return int(lib().oracle_max_threads())


class R250:
    """The oracle's R250 (oracle_r250_*): load 250 state words, then draw."""

    def __init__(self, state) -> None:
        self._s = R250State()
        arr = np.ascontiguousarray(state, dtype=np.uint32)
        if arr.size != 250:
            raise ValueError("R250 needs 250 state words")
        lib().oracle_r250_load(C.byref(self._s), arr.ctypes.data_as(_u32p))

    @classmethod
    def from_seed_file(cls, path) -> "R250":
        """State = words 7..256 of a LeMonADE seed file (little-endian uint32)."""
        words = np.frombuffer(open(path, "rb").read(), dtype="<u4")
        return cls(words[7:257])

    def rand32(self) -> int:
        return int(lib().oracle_r250_rand32(C.byref(self._s)))

    def substep_draws(self, n_substeps: int, n_colors: int) -> np.ndarray:
        """Per substep (species, kernel seed) = (rand32 % n_colors, rand32), as CRITICAL.cu:509-510."""
        out = np.empty(2 * n_substeps, dtype=np.uint32)
        for k in range(n_substeps):
            out[2 * k] = self.rand32() % n_colors
            out[2 * k + 1] = self.rand32()
        return out


class System:
    """A system in oracle form: unfolded corners, CSR bonds, proper coloring."""

    def __init__(self, box, xyz, start, ids, colors, forbidden=None, seed: int = 0):
        self.box = tuple(int(v) for v in box)
        self.xyz = np.ascontiguousarray(xyz, dtype=np.int32).reshape(-1, 3).copy()
        self.start = np.ascontiguousarray(start, dtype=np.uint32)
        self.ids = np.ascontiguousarray(ids if len(ids) else np.zeros(1), dtype=np.uint32)
        self.colors = np.ascontiguousarray(colors, dtype=np.uint8)
        self.n_colors = int(self.colors.max()) + 1 if self.colors.size else 0
        self.forbidden = standard_bondset() if forbidden is None else np.ascontiguousarray(forbidden, np.uint8)
        self.seed = int(seed)
        self.draws = None      # optional explicit draws: 2 words per substep from draw_t0 on
        self.draw_t0 = 0

    def _struct(self) -> OracleSystem:
        s = OracleSystem()
        for a in range(3):
            s.box[a] = self.box[a]
        s.n = self.xyz.shape[0]
        s.xyz = self.xyz.ctypes.data_as(_i32p)
        s.nbr_start = self.start.ctypes.data_as(_u32p)
        s.nbr = self.ids.ctypes.data_as(_u32p)
        s.color = self.colors.ctypes.data_as(_u8p)
        s.n_colors = self.n_colors
        s.forbidden = self.forbidden.ctypes.data_as(_u8p)
        s.seed = self.seed & 0xFFFFFFFFFFFFFFFF
        if self.draws is not None:
            self.draws = np.ascontiguousarray(self.draws, dtype=np.uint32)
            s.draws = self.draws.ctypes.data_as(_u32p)
            s.draw_t0 = int(self.draw_t0)
        return s

    def run(self, t0: int, n_substeps: int, threads: int = 0, shuffle_seed: int = 0):
        log = np.zeros(max(n_substeps, 1), dtype=np.uint32)
        s = self._struct()
        total = lib().oracle_run(C.byref(s), t0, n_substeps, log.ctypes.data_as(_u32p), threads, shuffle_seed)
        if total < 0:
            raise ValueError("oracle rejected the system")
        return int(total), log[:n_substeps]

    def run_naive(self, t0: int, n_substeps: int):
        log = np.zeros(max(n_substeps, 1), dtype=np.uint32)
        s = self._struct()
        total = lib().oracle_run_naive(C.byref(s), t0, n_substeps, log.ctypes.data_as(_u32p))
        if total < 0:
            raise ValueError("oracle rejected the system")
        return int(total), log[:n_substeps]

    def check(self) -> int:
        s = self._struct()
        return int(lib().oracle_check_system(C.byref(s)))

    def sequential(self, seed: int, n_mcs: int) -> int:
        s = self._struct()
        return int(lib().oracle_sequential_run(C.byref(s), seed & 0xFFFFFFFFFFFFFFFF, n_mcs))


def ref_coloring_available() -> bool:
    return REF_COLORING_PATH.exists()


def ref_coloring(n: int, start: np.ndarray, ids: np.ndarray, uniform: bool = True) -> np.ndarray:
    """Coloring computed by the reference's own graphColoring.tpp (oracle/_ref)."""
    h = C.CDLL(str(REF_COLORING_PATH))
    h.ref_graph_coloring.restype = C.c_int
    h.ref_graph_coloring.argtypes = [C.c_uint32, _u32p, _u32p, C.c_int, _u8p]
    s = np.ascontiguousarray(start, dtype=np.uint32)
    i = np.ascontiguousarray(ids if len(ids) else np.zeros(1), dtype=np.uint32)
    out = np.empty(max(n, 1), dtype=np.uint8)
    code = h.ref_graph_coloring(n, s.ctypes.data_as(_u32p), i.ctypes.data_as(_u32p), int(uniform),
                                out.ctypes.data_as(_u8p))
    if code:
        raise RuntimeError("reference coloring failed")
    return out[:n]
