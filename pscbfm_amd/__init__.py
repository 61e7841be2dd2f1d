"""MI355X (gfx950) scBFM sweep: Python mirror of the reference's updater API.

`UpdaterGPUScBFM_AB_Type` keeps the method names, argument meaning and error
behaviour of src/pscbfm/UpdaterGPUScBFM_AB_Type.h:316-361, and calls the
hand-written HIP kernels through the C ABI of include/pscbfm.h (ctypes over
pscbfm_amd/libpscbfm.so).  There is no fallback: without the built library
every call raises.

Error mapping (reference exception -> Python):
    std::invalid_argument -> InvalidArgument (a ValueError)
    std::runtime_error    -> RuntimeError
    std::out_of_range     -> IndexError
    CUDA error + exit()   -> DeviceError (a RuntimeError)
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

__all__ = [
    "UpdaterGPUScBFM_AB_Type", "InvalidArgument", "DeviceError", "RunStats", "lib", "library_path",
    "LATTICE_U8", "LATTICE_BIT", "standard_bond_vectors", "move_hash", "substep_draw",
    "zcurve_encode", "color_graph", "gen_linear_chains", "gen_linear_chains_ordered",
    "gen_star_network", "gen_stars_ordered", "crosslink_star_ends", "chain_bonds", "Batch", "ReplicaInfo",
    "autotune_batch", "TUNE_CANDIDATES", "Server", "JobInfo", "ServerCounters",
]

LATTICE_U8 = 0
LATTICE_BIT = 1
LEGACY_WALK = 0xFFFFFFFF  # setPerformWalk: one pass-1 list per pass-2 workgroup

_HERE = Path(__file__).resolve().parent
library_path = _HERE / "libpscbfm.so"


class InvalidArgument(ValueError):
    """std::invalid_argument of the reference."""


class DeviceError(RuntimeError):
    """HIP runtime failure (the reference exits the process instead)."""


class RunStats(C.Structure):
    _fields_ = [
        ("gpu_ms", C.c_double), ("substeps", C.c_uint64), ("attempted", C.c_uint64),
        ("accepted", C.c_uint64), ("bond_terms", C.c_uint64), ("check_ms", C.c_double),
        ("perform_ms", C.c_double), ("check_launches", C.c_uint64), ("perform_launches", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ReplicaInfo(C.Structure):
    _fields_ = [
        ("live", C.c_int32), ("n_colors", C.c_uint32), ("n_monomers", C.c_uint32), ("tiles", C.c_uint32),
        ("substeps", C.c_uint64), ("stop_at", C.c_uint64), ("accepted", C.c_uint64),
        ("device_bytes", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class JobInfo(C.Structure):
    _fields_ = [
        ("state", C.c_int32), ("n_monomers", C.c_uint32), ("n_mcs", C.c_uint64), ("accepted", C.c_uint64),
        ("device_bytes", C.c_uint64), ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerCounters(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("done", C.c_uint64), ("failed", C.c_uint64), ("queued", C.c_uint32),
        ("running", C.c_uint32), ("resident_bytes", C.c_uint64), ("peak_resident_bytes", C.c_uint64),
        ("used_tiles", C.c_uint32), ("tile_capacity", C.c_uint32), ("launches", C.c_uint64),
        ("substeps", C.c_uint64), ("gpu_busy_seconds", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_LIB: Optional[C.CDLL] = None

_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)
_u8p = C.POINTER(C.c_uint8)
_H = C.c_void_p

_SIGNATURES = {
    "pscbfm_create": (C.c_int, [C.POINTER(_H)]),
    "pscbfm_destroy": (None, [_H]),
    "pscbfm_last_error": (C.c_char_p, [_H]),
    "pscbfm_create_error": (C.c_char_p, []),
    "pscbfm_destruct": (C.c_int, [_H]),
    "pscbfm_set_gpu": (C.c_int, [_H, C.c_int]),
    "pscbfm_copy_bond_set": (C.c_int, [_H, C.c_int, C.c_int, C.c_int, C.c_int]),
    "pscbfm_set_nr_of_all_monomers": (C.c_int, [_H, C.c_uint32]),
    "pscbfm_set_attribute": (C.c_int, [_H, C.c_uint32, C.c_int32]),
    "pscbfm_set_monomer_coordinates": (C.c_int, [_H, C.c_uint32, C.c_int32, C.c_int32, C.c_int32]),
    "pscbfm_set_monomer_coordinates_bulk": (C.c_int, [_H, _i32p, C.c_uint32]),
    "pscbfm_set_connectivity": (C.c_int, [_H, C.c_uint32, C.c_uint32]),
    "pscbfm_set_connectivity_bulk": (C.c_int, [_H, _u32p, C.c_uint64]),
    "pscbfm_set_lattice_size": (C.c_int, [_H, C.c_uint32, C.c_uint32, C.c_uint32]),
    "pscbfm_set_periodicity": (C.c_int, [_H, C.c_int, C.c_int, C.c_int]),
    "pscbfm_populate_lattice": (C.c_int, [_H]),
    "pscbfm_initialize": (C.c_int, [_H]),
    "pscbfm_execute": (C.c_int, [_H]),
    "pscbfm_run": (C.c_int, [_H, C.c_int32, _u64p]),
    "pscbfm_get_monomer_position": (C.c_int, [_H, C.c_uint32, _i32p, _i32p, _i32p]),
    "pscbfm_get_positions_bulk": (C.c_int, [_H, _i32p, C.c_uint32]),
    "pscbfm_check_system": (C.c_int, [_H]),
    "pscbfm_cleanup": (C.c_int, [_H]),
    "pscbfm_set_seed": (C.c_int, [_H, C.c_uint64]),
    "pscbfm_set_start_mcs": (C.c_int, [_H, C.c_uint64]),
    "pscbfm_set_r250_state": (C.c_int, [_H, _u32p]),
    "pscbfm_r250_from_seed_file": (C.c_int, [_H, C.c_char_p, C.c_uint32]),
    "pscbfm_r250_next": (C.c_int, [_H, _u32p, C.c_uint64]),
    "pscbfm_r250_get_state": (C.c_int, [_H, _u32p]),
    "pscbfm_set_draw_chunk": (C.c_int, [_H, C.c_uint64]),
    "pscbfm_set_lattice_kind": (C.c_int, [_H, C.c_int]),
    "pscbfm_set_uniform_coloring": (C.c_int, [_H, C.c_int]),
    "pscbfm_set_check_on_initialize": (C.c_int, [_H, C.c_int]),
    "pscbfm_set_colors": (C.c_int, [_H, _u8p, C.c_uint32]),
    "pscbfm_set_launch_shape": (C.c_int, [_H, C.c_uint32, C.c_uint32]),
    "pscbfm_set_perform_block": (C.c_int, [_H, C.c_uint32]),
    "pscbfm_set_graph_chunk": (C.c_int, [_H, C.c_uint32]),
    "pscbfm_set_read_back": (C.c_int, [_H, C.c_int]),
    "pscbfm_read_positions_back": (C.c_int, [_H]),
    "pscbfm_set_perform_walk": (C.c_int, [_H, C.c_uint32, C.c_uint32]),
    "pscbfm_set_clear_pass": (C.c_int, [_H, C.c_uint32]),
    "pscbfm_set_slot_order": (C.c_int, [_H, C.c_int]),
    "pscbfm_prepare_graphs": (C.c_int, [_H, C.c_int32]),
    "pscbfm_run_ex": (C.c_int, [_H, C.c_int32, _u32p, C.c_int, C.c_int, C.POINTER(RunStats)]),
    "pscbfm_last_run_stats": (C.c_int, [_H, C.POINTER(RunStats)]),
    "pscbfm_run_ex_noalloc": (C.c_int, [_H, C.c_int32, _u32p, C.POINTER(RunStats)]),
    "pscbfm_get_n_colors": (C.c_int, [_H, _u32p]),
    "pscbfm_get_colors": (C.c_int, [_H, _u8p, C.c_uint32]),
    "pscbfm_get_color_sizes": (C.c_int, [_H, _u32p, C.c_uint32]),
    "pscbfm_get_counters": (C.c_int, [_H, _u64p, _u64p]),
    "pscbfm_get_device_bytes": (C.c_int, [_H, _u64p]),
    "pscbfm_download_lattice": (C.c_int, [_H, _u8p, C.c_uint64]),
    "pscbfm_count_proposal_bits": (C.c_int, [_H, _u64p, _u64p]),
    "pscbfm_count_partner_escapes": (C.c_int, [_H, _u64p, _u64p]),
    "pscbfm_zcurve_encode": (C.c_int, [C.c_uint32] * 6 + [_u32p]),
    "pscbfm_move_hash": (C.c_uint32, [C.c_uint32]),
    "pscbfm_substep_draw": (None, [C.c_uint64, C.c_uint64, C.c_uint32, _u32p, _u32p]),
    "pscbfm_color_graph": (C.c_int, [C.c_uint32, _u32p, _u32p, C.c_int, _u8p]),
    "pscbfm_standard_bond_vectors": (C.c_int, [_i32p]),
    "pscbfm_gen_linear_chains": (C.c_int, [C.c_uint32] * 5 + [C.c_uint64, C.c_uint32, _i32p]),
    "pscbfm_gen_linear_chains_ordered": (C.c_int, [C.c_uint32] * 5 + [_i32p]),
    "pscbfm_gen_star_network": (C.c_int, [C.c_uint32] * 6 + [C.c_uint64, _i32p, _u32p, C.c_uint64, _u64p]),
    "pscbfm_gen_stars_ordered": (C.c_int, [C.c_uint32] * 5 + [C.c_uint64, _i32p, _u32p]),
    "pscbfm_crosslink_star_ends": (C.c_int, [C.c_uint32] * 7 + [C.c_uint64, _i32p, _u32p, C.c_uint64, _u64p]),
    "pscbfm_copy_probe": (C.c_int, [C.c_int, C.c_uint64, C.c_int, C.POINTER(C.c_double)]),
    "pscbfm_batch_create": (C.c_int, [C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.POINTER(_H)]),
    "pscbfm_batch_destroy": (None, [_H]),
    "pscbfm_batch_last_error": (C.c_char_p, [_H]),
    "pscbfm_batch_footprint": (C.c_int, [_H, _H, _u64p, _u32p]),
    "pscbfm_batch_add": (C.c_int, [_H, _H, _u32p]),
    "pscbfm_batch_remove": (C.c_int, [_H, C.c_uint32]),
    "pscbfm_batch_set_perform_block": (C.c_int, [_H, C.c_uint32]),
    "pscbfm_batch_set_target": (C.c_int, [_H, C.c_uint32, C.c_uint64]),
    "pscbfm_batch_launch": (C.c_int, [_H, C.c_uint64]),
    "pscbfm_batch_synchronize": (C.c_int, [_H]),
    "pscbfm_batch_run": (C.c_int, [_H, C.c_uint64, C.POINTER(C.c_double)]),
    "pscbfm_batch_info": (C.c_int, [_H, C.c_uint32, C.POINTER(ReplicaInfo)]),
    "pscbfm_batch_get_positions": (C.c_int, [_H, C.c_uint32, _i32p, C.c_uint32]),
    "pscbfm_batch_get_positions_many": (C.c_int, [_H, _u32p, C.c_uint32, C.POINTER(_i32p), _u32p]),
    "pscbfm_batch_usage": (C.c_int, [_H, _u32p, _u32p, _u32p, _u64p]),
    "pscbfm_autotune": (C.c_int, [_H, C.c_uint32, _u32p, _u32p, C.POINTER(C.c_double)]),
    "pscbfm_batch_autotune": (C.c_int, [C.c_int, C.c_int, C.POINTER(_H), C.c_uint32, C.c_uint32, _u32p, _u32p,
                                        C.POINTER(C.c_double)]),
    "pscbfm_server_create": (C.c_int, [C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_uint32, C.POINTER(_H)]),
    "pscbfm_server_destroy": (None, [_H]),
    "pscbfm_server_last_error": (C.c_char_p, [_H]),
    "pscbfm_server_submit": (C.c_int, [_H, _H, C.c_uint64, _u64p]),
    "pscbfm_server_wait": (C.c_int, [_H, C.c_uint64, C.c_double, C.POINTER(C.c_int)]),
    "pscbfm_server_job": (C.c_int, [_H, C.c_uint64, C.POINTER(JobInfo)]),
    "pscbfm_server_result": (C.c_int, [_H, C.c_uint64, _i32p, C.c_uint32]),
    "pscbfm_server_release": (C.c_int, [_H, C.c_uint64]),
    "pscbfm_server_drain": (C.c_int, [_H]),
    "pscbfm_server_stats": (C.c_int, [_H, C.POINTER(ServerCounters)]),
    "pscbfm_server_now": (C.c_int, [_H, C.POINTER(C.c_double)]),
    "pscbfm_bfm_stage": (C.c_int, [_H, C.c_char_p, _u64p]),
    "pscbfm_bfm_write": (C.c_int, [_H, C.c_char_p, C.c_uint64, C.c_int]),
    "pscbfm_version": (C.c_char_p, []),
}


def lib() -> C.CDLL:
    """Loads the in-tree HIP library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not library_path.exists():
            raise RuntimeError(f"{library_path} is missing; build it with `python -m pscbfm_amd.build` "
                               "(the sweep has no CPU fallback)")
        handle = C.CDLL(str(library_path))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def _ptr(arr: np.ndarray, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


def _raise(code: int, message: str) -> None:
    if code == 0:
        return
    if code == 1:
        raise InvalidArgument(message)
    if code == 2:
        raise RuntimeError(message)
    if code == 3:
        raise IndexError(message)
    if code == 4:
        raise DeviceError(message)
    raise RuntimeError(f"pscbfm error {code}: {message}")


class UpdaterGPUScBFM_AB_Type:
    """Drop-in mirror of the reference class (UpdaterGPUScBFM_AB_Type.h:124-362)."""

    def __init__(self) -> None:
        self._lib = lib()
        h = C.c_void_p()
        code = self._lib.pscbfm_create(C.byref(h))
        _raise(code, (self._lib.pscbfm_create_error() or b"").decode())
        self._h = h

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            self._lib.pscbfm_destroy(h)
            self._h = None

    def _call(self, name: str, *args) -> None:
        code = getattr(self._lib, name)(self._h, *args)
        if code:
            _raise(code, (self._lib.pscbfm_last_error(self._h) or b"").decode())

    # -- reference surface ----------------------------------------------------
    def destruct(self) -> None:
        self._call("pscbfm_destruct")

    def setGpu(self, iGpuToUse: int) -> None:
        self._call("pscbfm_set_gpu", int(iGpuToUse))

    def copyBondSet(self, dx: int, dy: int, dz: int, bondForbidden: bool) -> None:
        self._call("pscbfm_copy_bond_set", int(dx), int(dy), int(dz), int(bool(bondForbidden)))

    def setNrOfAllMonomers(self, nAllMonomers: int) -> None:
        self._call("pscbfm_set_nr_of_all_monomers", int(nAllMonomers))

    def setAttribute(self, i: int, attribute: int) -> None:
        self._call("pscbfm_set_attribute", int(i), int(attribute))

    def setMonomerCoordinates(self, i: int, x: int, y: int, z: int) -> None:
        self._call("pscbfm_set_monomer_coordinates", int(i), int(x), int(y), int(z))

    def setConnectivity(self, monoidx1: int, monoidx2: int) -> None:
        self._call("pscbfm_set_connectivity", int(monoidx1), int(monoidx2))

    def setLatticeSize(self, boxX: int, boxY: int, boxZ: int) -> None:
        self._call("pscbfm_set_lattice_size", int(boxX), int(boxY), int(boxZ))

    def setPeriodicity(self, isPeriodicX: bool, isPeriodicY: bool, isPeriodicZ: bool) -> None:
        self._call("pscbfm_set_periodicity", int(bool(isPeriodicX)), int(bool(isPeriodicY)),
                   int(bool(isPeriodicZ)))

    def populateLattice(self) -> None:
        self._call("pscbfm_populate_lattice")

    def initialize(self) -> None:
        self._call("pscbfm_initialize")

    def execute(self) -> bool:
        self._call("pscbfm_execute")
        return True

    def runSimulationOnGPU(self, nrMCS_per_Call: int) -> None:
        self._call("pscbfm_run", int(nrMCS_per_Call), None)

    def getMonomerPositionInX(self, i: int) -> int:
        return self._position(i)[0]

    def getMonomerPositionInY(self, i: int) -> int:
        return self._position(i)[1]

    def getMonomerPositionInZ(self, i: int) -> int:
        return self._position(i)[2]

    def checkSystem(self) -> None:
        self._call("pscbfm_check_system")

    def cleanup(self) -> None:
        self._call("pscbfm_cleanup")

    def _position(self, i: int):
        x, y, z = C.c_int32(), C.c_int32(), C.c_int32()
        self._call("pscbfm_get_monomer_position", int(i), C.byref(x), C.byref(y), C.byref(z))
        return x.value, y.value, z.value

    # -- extensions -----------------------------------------------------------
    def setSeed(self, seed: int) -> None:
        self._call("pscbfm_set_seed", int(seed) & 0xFFFFFFFFFFFFFFFF)

    def setStartMcs(self, mcs: int) -> None:
        """Age of the input state in sweeps (before initialize): the counter
        draw continues from sweep `mcs`, so a resumed run equals the
        uninterrupted one."""
        self._call("pscbfm_set_start_mcs", int(mcs))

    def setR250State(self, words: Sequence[int]) -> None:
        """R250 mode: the substep draws come from this R250 state (250 words in
        draw order), species = r250 % nColors then seed = r250 (CRITICAL.cu:509-510)."""
        arr = np.ascontiguousarray(words, dtype=np.uint32)
        if arr.size != 250:
            raise ValueError("an R250 state is 250 words")
        self._call("pscbfm_set_r250_state", _ptr(arr, C.c_uint32))

    def loadR250SeedFile(self, path, discard: int = 0) -> None:
        """R250 mode from a LeMonADE seed file (state = words 7..256), then
        `discard` draws (the reference simulator discards 2)."""
        self._call("pscbfm_r250_from_seed_file", str(path).encode(), int(discard))

    def r250Next(self, n: int) -> np.ndarray:
        out = np.empty(max(int(n), 1), dtype=np.uint32)
        self._call("pscbfm_r250_next", _ptr(out, C.c_uint32), int(n))
        return out[:int(n)]

    def setDrawChunk(self, substeps: int) -> None:
        """R250 mode: substeps of draws kept on the device at once (0: 2^22)."""
        self._call("pscbfm_set_draw_chunk", int(substeps))

    def r250State(self) -> np.ndarray:
        out = np.empty(250, dtype=np.uint32)
        self._call("pscbfm_r250_get_state", _ptr(out, C.c_uint32))
        return out

    def setLatticeKind(self, kind: int) -> None:
        self._call("pscbfm_set_lattice_kind", int(kind))

    def setUniformColoring(self, on: bool) -> None:
        self._call("pscbfm_set_uniform_coloring", int(bool(on)))

    def setCheckOnInitialize(self, on: bool) -> None:
        self._call("pscbfm_set_check_on_initialize", int(bool(on)))

    def setColors(self, colors: Sequence[int]) -> None:
        arr = np.ascontiguousarray(colors, dtype=np.uint8)
        self._call("pscbfm_set_colors", _ptr(arr, C.c_uint8), int(arr.size))

    def setLaunchShape(self, block: int = 0, slots: int = 0) -> None:
        self._call("pscbfm_set_launch_shape", int(block), int(slots))

    def setPerformBlock(self, block: int = 0) -> None:
        """Pass-2 threads per workgroup (0: automatic, 128 for systems at least a
        quarter occupied, else the pass-1 block)."""
        self._call("pscbfm_set_perform_block", int(block))

    def setPerformWalk(self, threads_per_list: int = 0, entries_per_thread: int = 0) -> None:
        """Pass-2 walk: threads per pass-1 list (0 or LEGACY_WALK, the default:
        one list per workgroup) and entries in flight per thread (1 or 4; 0: 4)."""
        self._call("pscbfm_set_perform_walk", int(threads_per_list), int(entries_per_thread))

    def setSlotOrder(self, order=None) -> None:
        """Slot order inside each species (before initialize): "spatial" (or
        True), by starting Z-curve code; "reference" (or False), the
        reference's monomer order; None or "auto" (the default), spatial on
        the bit lattice and on the byte lattice only for scattered partners.
        Results are the same in every order."""
        codes = {None: 2, "auto": 2, True: 1, "spatial": 1, False: 0, "reference": 0}
        if order not in codes:
            raise InvalidArgument(f"unknown slot order {order!r}")
        self._call("pscbfm_set_slot_order", codes[order])

    def setClearPass(self, pass_: int = 0) -> None:
        """The pass (1 or 2; 0: automatic, pass 1 on the bit lattice with at
        most two bonds per monomer, else pass 2) that zeroes the other
        parity's proposal buffer each substep."""
        self._call("pscbfm_set_clear_pass", int(pass_))

    def setReadBack(self, on: bool) -> None:
        """Whether runs copy the positions back (default on, as runSimulationOnGPU)."""
        self._call("pscbfm_set_read_back", int(bool(on)))

    def readPositionsBack(self) -> None:
        self._call("pscbfm_read_positions_back")

    def setGraphChunk(self, substeps: int) -> None:
        self._call("pscbfm_set_graph_chunk", int(substeps))

    def prepareGraphs(self, n_mcs: int) -> None:
        """Builds the graphs a run of n_mcs sweeps replays (after initialize)."""
        self._call("pscbfm_prepare_graphs", int(n_mcs))

    def setMonomerCoordinatesBulk(self, xyz: np.ndarray) -> None:
        arr = np.ascontiguousarray(xyz, dtype=np.int32).reshape(-1, 3)
        self._call("pscbfm_set_monomer_coordinates_bulk", _ptr(arr, C.c_int32), int(arr.shape[0]))

    def setConnectivityBulk(self, pairs: np.ndarray) -> None:
        arr = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        self._call("pscbfm_set_connectivity_bulk", _ptr(arr, C.c_uint32), int(arr.shape[0]))

    def getPositions(self, n: int) -> np.ndarray:
        out = np.empty((n, 3), dtype=np.int32)
        self._call("pscbfm_get_positions_bulk", _ptr(out, C.c_int32), int(n))
        return out

    def run(self, n_mcs: int, log: bool = False, profile: bool = False, use_graph: bool = True):
        """Runs n_mcs sweeps; returns (RunStats, per-substep accepted log or None)."""
        stats = RunStats()
        n_colors = self.nColors() if log else 0
        buf = np.zeros(int(n_mcs) * n_colors, dtype=np.uint32) if log else None
        self._call("pscbfm_run_ex", int(n_mcs), _ptr(buf, C.c_uint32) if log else None,
                   int(bool(profile)), int(bool(use_graph)), C.byref(stats))
        return stats, buf

    def runNoAlloc(self, n_mcs: int, log: bool = False):
        """run(n_mcs, log, use_graph=False) without any device allocation
        (pscbfm_run_ex_noalloc); returns (RunStats, log or None)."""
        stats = RunStats()
        n_colors = self.nColors() if log else 0
        buf = np.zeros(int(n_mcs) * n_colors, dtype=np.uint32) if log else None
        self._call("pscbfm_run_ex_noalloc", int(n_mcs), _ptr(buf, C.c_uint32) if log else None, C.byref(stats))
        return stats, buf

    def autotune(self, mcs: int = 20):
        """Picks and applies the fastest launch shape (scratch copies only);
        returns (block, slots, {(block, slots): ms per sweep})."""
        b, sl = C.c_uint32(), C.c_uint32()
        ms = (C.c_double * len(TUNE_CANDIDATES))()
        self._call("pscbfm_autotune", int(mcs), C.byref(b), C.byref(sl), ms)
        return b.value, sl.value, {c: ms[i] for i, c in enumerate(TUNE_CANDIDATES) if ms[i] >= 0}

    def loadBfm(self, path) -> int:
        """Stages the last configuration of a BFM file (reference call order);
        returns the frame's MCS age."""
        mcs = C.c_uint64()
        self._call("pscbfm_bfm_stage", str(path).encode(), C.byref(mcs))
        return mcs.value

    def writeBfm(self, path, mcs: int, append: bool = False) -> None:
        """Writes the current coordinates as a BFM frame (header first unless appending to a non-empty file)."""
        self._call("pscbfm_bfm_write", str(path).encode(), int(mcs), int(bool(append)))

    def lastRunStats(self) -> RunStats:
        s = RunStats()
        self._call("pscbfm_last_run_stats", C.byref(s))
        return s

    def nColors(self) -> int:
        v = C.c_uint32()
        self._call("pscbfm_get_n_colors", C.byref(v))
        return v.value

    def colors(self, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.uint8)
        self._call("pscbfm_get_colors", _ptr(out, C.c_uint8), int(n))
        return out

    def colorSizes(self) -> np.ndarray:
        k = self.nColors()
        out = np.empty(k, dtype=np.uint32)
        self._call("pscbfm_get_color_sizes", _ptr(out, C.c_uint32), int(k))
        return out

    def counters(self):
        s, a = C.c_uint64(), C.c_uint64()
        self._call("pscbfm_get_counters", C.byref(s), C.byref(a))
        return s.value, a.value

    def deviceBytes(self) -> int:
        v = C.c_uint64()
        self._call("pscbfm_get_device_bytes", C.byref(v))
        return v.value

    def downloadLattice(self, n_cells: int) -> np.ndarray:
        out = np.empty(n_cells, dtype=np.uint8)
        self._call("pscbfm_download_lattice", _ptr(out, C.c_uint8), int(n_cells))
        return out

    def proposalBits(self):
        a, b = C.c_uint64(), C.c_uint64()
        self._call("pscbfm_count_proposal_bits", C.byref(a), C.byref(b))
        return a.value, b.value

    def partnerEscapes(self):
        """(directed bonded partners, those without a 16-bit partner code) of
        the species layout; works on a staged handle without a GPU."""
        p, e = C.c_uint64(), C.c_uint64()
        self._call("pscbfm_count_partner_escapes", C.byref(p), C.byref(e))
        return p.value, e.value


TUNE_CANDIDATES = [(b, s) for b in (64, 128, 192, 256) for s in (1, 2, 4)]


def autotune_batch(staged, mcs: int = 5, lattice_kind: int = LATTICE_BIT, gpu: int = 0):
    """Fastest (block, slots) for a batch of the staged systems (a sample of
    the real mix); returns (block, slots, {(block, slots): ms per sweep})."""
    arr = (C.c_void_p * len(staged))(*[u._h.value for u in staged])
    b, sl = C.c_uint32(), C.c_uint32()
    ms = (C.c_double * len(TUNE_CANDIDATES))()
    code = lib().pscbfm_batch_autotune(int(gpu), int(lattice_kind), arr, len(staged), int(mcs), C.byref(b),
                                       C.byref(sl), ms)
    _raise(code, (lib().pscbfm_create_error() or b"").decode())
    return b.value, sl.value, {c: ms[i] for i, c in enumerate(TUNE_CANDIDATES) if ms[i] >= 0}


class Batch:
    """Many independent systems swept by one launch per pass (include/pscbfm.h
    pscbfm_batch_*).  A replica is the system staged in an updater (no
    initialize needed) with that updater's seed; it evolves bit-identically to
    the same updater run alone.  Replicas join and leave between launches."""

    def __init__(self, tile_capacity: int = 0, replica_capacity: int = 1024, lattice_kind: int = LATTICE_BIT,
                 block: int = 256, slots: int = 2, gpu: int = 0) -> None:
        """tile_capacity 0: the workgroup table grows with the replicas."""
        self._lib = lib()
        h = C.c_void_p()
        code = self._lib.pscbfm_batch_create(int(gpu), int(lattice_kind), int(tile_capacity),
                                             int(replica_capacity), int(block), int(slots), C.byref(h))
        _raise(code, (self._lib.pscbfm_create_error() or b"").decode())
        self._h = h

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            self._lib.pscbfm_batch_destroy(h)
            self._h = None

    def _call(self, name: str, *args) -> None:
        code = getattr(self._lib, name)(self._h, *args)
        if code:
            _raise(code, (self._lib.pscbfm_batch_last_error(self._h) or b"").decode())

    def footprint(self, staged: "UpdaterGPUScBFM_AB_Type"):
        """(device bytes, workgroups) the staged system would take."""
        b, t = C.c_uint64(), C.c_uint32()
        self._call("pscbfm_batch_footprint", staged._h, C.byref(b), C.byref(t))
        return b.value, t.value

    def add(self, staged: "UpdaterGPUScBFM_AB_Type") -> int:
        rid = C.c_uint32()
        self._call("pscbfm_batch_add", staged._h, C.byref(rid))
        return rid.value

    def remove(self, rid: int) -> None:
        self._call("pscbfm_batch_remove", int(rid))

    def set_perform_block(self, block: int) -> None:
        self._call("pscbfm_batch_set_perform_block", int(block))

    def set_target(self, rid: int, n_mcs: int) -> None:
        self._call("pscbfm_batch_set_target", int(rid), int(n_mcs))

    def launch(self, substeps: int) -> None:
        self._call("pscbfm_batch_launch", int(substeps))

    def synchronize(self) -> None:
        self._call("pscbfm_batch_synchronize")

    def run(self, n_mcs: int) -> float:
        ms = C.c_double()
        self._call("pscbfm_batch_run", int(n_mcs), C.byref(ms))
        return ms.value

    def info(self, rid: int) -> ReplicaInfo:
        out = ReplicaInfo()
        self._call("pscbfm_batch_info", int(rid), C.byref(out))
        return out

    def positions(self, rid: int, n: int) -> np.ndarray:
        out = np.empty((n, 3), dtype=np.int32)
        self._call("pscbfm_batch_get_positions", int(rid), _ptr(out, C.c_int32), int(n))
        return out

    def positions_many(self, rids, ns) -> list:
        """positions() of several replicas: decoded on the device, copied back
        in rounds of up to 64 MiB."""
        ids = np.ascontiguousarray(rids, dtype=np.uint32)
        counts = np.ascontiguousarray(ns, dtype=np.uint32)
        if ids.shape != counts.shape:
            raise InvalidArgument("one count per replica id")
        outs = [np.empty((int(n), 3), dtype=np.int32) for n in counts]
        ptrs = (_i32p * max(len(outs), 1))(*[_ptr(o, C.c_int32) for o in outs])
        self._call("pscbfm_batch_get_positions_many", _ptr(ids, C.c_uint32), int(len(ids)), ptrs,
                   _ptr(counts, C.c_uint32))
        return outs

    def usage(self):
        """(live replicas, used workgroups, workgroup capacity, device bytes)."""
        a, b, c, d = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        self._call("pscbfm_batch_usage", C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return a.value, b.value, c.value, d.value


JOB_QUEUED, JOB_PREPARED, JOB_RUNNING, JOB_DONE, JOB_FAILED = range(5)


class Server:
    """Continuous-batching request loop (include/pscbfm.h pscbfm_server_*):
    submit staged systems with a number of sweeps from any thread; they are
    admitted into one batch as HBM budget and workgroups allow and run
    concurrently; each result is the one the system has alone."""

    def __init__(self, hbm_budget: int, tile_capacity: int, replica_capacity: int = 4096,
                 lattice_kind: int = LATTICE_BIT, block: int = 256, slots: int = 2, quantum: int = 64,
                 gpu: int = 0) -> None:
        self._lib = lib()
        h = C.c_void_p()
        code = self._lib.pscbfm_server_create(int(gpu), int(lattice_kind), int(hbm_budget), int(tile_capacity),
                                              int(replica_capacity), int(block), int(slots), int(quantum),
                                              C.byref(h))
        _raise(code, (self._lib.pscbfm_create_error() or b"").decode())
        self._h = h

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            self._lib.pscbfm_server_destroy(h)
            self._h = None

    __del__ = close

    def _call(self, name: str, *args) -> None:
        code = getattr(self._lib, name)(self._h, *args)
        if code:
            _raise(code, (self._lib.pscbfm_server_last_error(self._h) or b"").decode())

    def submit(self, staged: "UpdaterGPUScBFM_AB_Type", n_mcs: int) -> int:
        job = C.c_uint64()
        self._call("pscbfm_server_submit", staged._h, int(n_mcs), C.byref(job))
        return job.value

    def wait(self, job: int, timeout: float = -1.0) -> bool:
        done = C.c_int()
        self._call("pscbfm_server_wait", int(job), float(timeout), C.byref(done))
        return bool(done.value)

    def job(self, job: int) -> JobInfo:
        out = JobInfo()
        self._call("pscbfm_server_job", int(job), C.byref(out))
        return out

    def result(self, job: int, n: int) -> np.ndarray:
        out = np.empty((n, 3), dtype=np.int32)
        self._call("pscbfm_server_result", int(job), _ptr(out, C.c_int32), int(n))
        return out

    def release(self, job: int) -> None:
        self._call("pscbfm_server_release", int(job))

    def drain(self) -> None:
        self._call("pscbfm_server_drain")

    def stats(self) -> ServerCounters:
        out = ServerCounters()
        self._call("pscbfm_server_stats", C.byref(out))
        return out

    def now(self) -> float:
        t = C.c_double()
        self._call("pscbfm_server_now", C.byref(t))
        return t.value


# -- host helpers -----------------------------------------------------------------
def standard_bond_vectors() -> np.ndarray:
    out = np.empty((108, 3), dtype=np.int32)
    _raise(lib().pscbfm_standard_bond_vectors(_ptr(out, C.c_int32)), "bond vectors")
    return out


def move_hash(a: int) -> int:
    return int(lib().pscbfm_move_hash(int(a) & 0xFFFFFFFF))


def substep_draw(seed: int, t: int, n_colors: int):
    s, k = C.c_uint32(), C.c_uint32()
    lib().pscbfm_substep_draw(int(seed) & 0xFFFFFFFFFFFFFFFF, int(t), int(n_colors), C.byref(s), C.byref(k))
    return s.value, k.value


def zcurve_encode(box, x: int, y: int, z: int) -> int:
    out = C.c_uint32()
    _raise(lib().pscbfm_zcurve_encode(int(box[0]), int(box[1]), int(box[2]), int(x), int(y), int(z),
                                      C.byref(out)), "zcurve")
    return out.value


def color_graph(n: int, start: np.ndarray, ids: np.ndarray, uniform: bool = True) -> np.ndarray:
    s = np.ascontiguousarray(start, dtype=np.uint32)
    i = np.ascontiguousarray(ids, dtype=np.uint32)
    if i.size == 0:
        i = np.zeros(1, dtype=np.uint32)
    out = np.empty(max(n, 1), dtype=np.uint8)
    _raise(lib().pscbfm_color_graph(int(n), _ptr(s, C.c_uint32), _ptr(i, C.c_uint32), int(bool(uniform)),
                                    _ptr(out, C.c_uint8)), "graph coloring")
    return out[:n]


def copy_probe(gpu: int = 0, nbytes: int = 1 << 29, repeats: int = 10) -> float:
    """Mean time (ms) of a plain device copy of nbytes (read nbytes + write nbytes)."""
    ms = C.c_double()
    _raise(lib().pscbfm_copy_probe(int(gpu), int(nbytes), int(repeats), C.byref(ms)), "copy probe failed")
    return ms.value


def chain_bonds(n_chains: int, length: int) -> np.ndarray:
    """Undirected bond pairs (i, i+1) inside each of n_chains consecutive chains."""
    if length < 2 or n_chains == 0:
        return np.zeros((0, 2), dtype=np.uint32)
    base = (np.arange(n_chains, dtype=np.uint64) * length)[:, None]
    k = np.arange(length - 1, dtype=np.uint64)[None, :]
    a = (base + k).reshape(-1)
    return np.stack([a, a + 1], axis=1).astype(np.uint32)


def gen_linear_chains(box, n_chains: int, length: int, seed: int, max_restarts: int = 10000) -> np.ndarray:
    """Chains grown as self-avoiding walks (raises RuntimeError if the box jams)."""
    out = np.empty((n_chains * length, 3), dtype=np.int32)
    code = lib().pscbfm_gen_linear_chains(int(box[0]), int(box[1]), int(box[2]), int(n_chains), int(length),
                                          int(seed) & 0xFFFFFFFFFFFFFFFF, int(max_restarts),
                                          _ptr(out, C.c_int32))
    _raise(code, "self-avoiding growth failed (box too dense for the requested chains)")
    return out


def gen_linear_chains_ordered(box, n_chains: int, length: int) -> np.ndarray:
    out = np.empty((n_chains * length, 3), dtype=np.int32)
    code = lib().pscbfm_gen_linear_chains_ordered(int(box[0]), int(box[1]), int(box[2]), int(n_chains),
                                                  int(length), _ptr(out, C.c_int32))
    _raise(code, "ordered placement does not fit the requested chains")
    return out


def gen_star_network(box, n_stars: int, arms: int, arm_length: int, seed: int):
    per = 1 + arms * arm_length
    xyz = np.empty((n_stars * per, 3), dtype=np.int32)
    cap = n_stars * arms * (arm_length + 1)
    bonds = np.empty((max(cap, 1), 2), dtype=np.uint32)
    nb = C.c_uint64()
    code = lib().pscbfm_gen_star_network(int(box[0]), int(box[1]), int(box[2]), int(n_stars), int(arms),
                                         int(arm_length), int(seed) & 0xFFFFFFFFFFFFFFFF, _ptr(xyz, C.c_int32),
                                         _ptr(bonds, C.c_uint32), int(cap), C.byref(nb))
    _raise(code, "star network growth failed")
    return xyz, bonds[: nb.value].copy()


def gen_stars_ordered(box, n_stars: int, arm_length: int, seed: int):
    """Four-arm stars on a dense ordered packing (corners at even
    coordinates, up to 91% of those sites filled) and their bonds; no
    crosslinks (see crosslink_star_ends)."""
    per = 1 + 4 * arm_length
    xyz = np.empty((n_stars * per, 3), dtype=np.int32)
    bonds = np.empty((max(n_stars * 4 * arm_length, 1), 2), dtype=np.uint32)
    code = lib().pscbfm_gen_stars_ordered(int(box[0]), int(box[1]), int(box[2]), int(n_stars), int(arm_length),
                                          int(seed) & 0xFFFFFFFFFFFFFFFF, _ptr(xyz, C.c_int32),
                                          _ptr(bonds, C.c_uint32))
    _raise(code, "ordered stars do not fit the box")
    return xyz, bonds[: n_stars * 4 * arm_length].copy()


def crosslink_star_ends(box, xyz: np.ndarray, n_stars: int, arms: int, arm_length: int, seed: int,
                        rounds: int = 1) -> np.ndarray:
    """Crosslink pairs (k, 2) joining arm ends to other stars' arm interiors
    at valid bond vectors, for stars at the corners xyz (unfolded); each
    round gives every arm end at most one more partner."""
    pts = np.ascontiguousarray(xyz, dtype=np.int32)
    if pts.shape != (n_stars * (1 + arms * arm_length), 3):
        raise InvalidArgument("corners do not match the star count and size")
    cap = max(n_stars * arms * int(rounds), 1)
    bonds = np.empty((cap, 2), dtype=np.uint32)
    nb = C.c_uint64()
    code = lib().pscbfm_crosslink_star_ends(int(box[0]), int(box[1]), int(box[2]), int(n_stars), int(arms),
                                            int(arm_length), int(rounds), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                            _ptr(pts, C.c_int32),
                                            _ptr(bonds, C.c_uint32), int(cap), C.byref(nb))
    _raise(code, "crosslinking failed")
    return bonds[: nb.value].copy()
