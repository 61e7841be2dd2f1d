"""CPU tests of the library boundary and host logic (no GPU work).

Covers: every symbol of include/pscbfm.h is exported; the reference's error
behaviour (UpdaterGPUScBFM_AB_Type.cu: 472-487, 521-528, 699-726, 728-753,
777-795, 797-848, 882-986) through the Python mirror and the C++ drop-in
header; the graph coloring against the reference's own graphColoring.tpp
(compiled from the reference tree into oracle/_ref); BFM file I/O on the
reference's fixtures; the synthetic system builders.
"""
import ctypes as C
import itertools
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
import pscbfm_amd as P
from pscbfm_amd.bfm import read_bfm, write_bfm
import pscbfm_amd.workloads as W
from helpers import GOLDEN, Sys, chains, oracle_for, oracle_relax, ordered_chains, stars

ROOT = Path(__file__).resolve().parents[1]


# ---- exports ----------------------------------------------------------------------
def _declared(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pscbfm_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = _declared(ROOT / "include" / "pscbfm.h")
    assert len(names) > 40
    lib = C.CDLL(str(P.library_path))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_a_gfx950_code_object():
    data = P.library_path.read_bytes()
    assert b"gfx950" in data


# ---- error behaviour of the reference surface ------------------------------------
def _filled(n=2, box=(16, 16, 16)):
    u = P.UpdaterGPUScBFM_AB_Type()
    u.setPeriodicity(True, True, True)
    u.setNrOfAllMonomers(n)
    return u


def test_set_gpu_out_of_range_is_invalid_argument():
    u = P.UpdaterGPUScBFM_AB_Type()
    with pytest.raises(P.InvalidArgument, match="not present"):
        u.setGpu(64)
    with pytest.raises(P.InvalidArgument):
        u.setGpu(-1)


def test_non_periodic_box_is_invalid_argument():
    u = P.UpdaterGPUScBFM_AB_Type()
    for flags in [(False, True, True), (True, False, True), (True, True, False)]:
        with pytest.raises(P.InvalidArgument, match="periodic"):
            u.setPeriodicity(*flags)


def test_set_nr_of_all_monomers_twice_is_runtime_error():
    u = _filled()
    with pytest.raises(RuntimeError, match="already set"):
        u.setNrOfAllMonomers(5)


def test_coordinate_index_is_range_checked():
    u = _filled(3)
    with pytest.raises(IndexError):
        u.setMonomerCoordinates(3, 0, 0, 0)
    with pytest.raises(IndexError):
        u.getMonomerPositionInX(7)


def test_connectivity_limit_is_invalid_argument():
    u = _filled(10)
    for k in range(1, 9):
        u.setConnectivity(0, k)
    with pytest.raises(P.InvalidArgument, match="maximum amount of bonds"):
        u.setConnectivity(0, 9)


def test_lattice_size_checks():
    u = P.UpdaterGPUScBFM_AB_Type()
    with pytest.raises(RuntimeError, match="power of 2"):
        u.setLatticeSize(16, 24, 16)
    with pytest.raises(RuntimeError, match="power of 2"):
        u.setLatticeSize(16, 16, 48)   # the reference does not check z; here it is checked
    with pytest.raises(P.InvalidArgument, match="larger than"):
        u.setLatticeSize(65536, 16, 16)
    with pytest.raises(RuntimeError, match="at least 8"):
        u.setLatticeSize(4, 16, 16)
    u.setLatticeSize(32, 64, 16)
    u.setLatticeSize(32, 64, 16)   # unchanged: no-op


def test_initialize_requires_108_bonds():
    u = _filled(1)
    u.setLatticeSize(16, 16, 16)
    # no copyBondSet calls: every entry allowed -> 512 allowed bonds
    with pytest.raises(RuntimeError, match="Expected 108 allowed bonds, but got 512"):
        u.initialize()


def test_check_system_detects_overlap_and_forbidden_bond():
    forb = oracle.standard_bondset()

    def build(xyz, bonds):
        u = P.UpdaterGPUScBFM_AB_Type()
        u.setNrOfAllMonomers(len(xyz))
        for i, p in enumerate(xyz):
            u.setMonomerCoordinates(i, *p)
        for a, b in bonds:
            u.setConnectivity(a, b)
            u.setConnectivity(b, a)
        u.setLatticeSize(16, 16, 16)
        for dx, dy, dz in itertools.product(range(-4, 4), repeat=3):
            u.copyBondSet(dx, dy, dz, bool(forb[oracle.bond_index(dx, dy, dz)]))
        return u

    build([(0, 0, 0), (2, 0, 0)], [(0, 1)]).checkSystem()
    build([(15, 0, 0), (1, 0, 0)], [(0, 1)]).checkSystem()   # bond across the boundary (minimum image)
    with pytest.raises(RuntimeError, match="excluded volume"):
        build([(0, 0, 0), (1, 1, 0)], []).checkSystem()
    with pytest.raises(RuntimeError, match="excluded volume"):
        build([(0, 0, 0), (15, 15, 15)], []).checkSystem()   # overlap through the periodic boundary
    with pytest.raises(RuntimeError, match="forbidden bond"):
        build([(0, 0, 0), (2, 2, 0)], [(0, 1)]).checkSystem()


def test_colors_override_must_be_proper_and_gapless():
    u = _filled(2)
    u.setConnectivity(0, 1)
    u.setColors([0, 0])   # stored; validated at initialize (needs a GPU) -- only the count is checked here
    with pytest.raises(P.InvalidArgument):
        u.setColors([0])


def test_run_before_initialize_is_runtime_error():
    u = _filled(1)
    with pytest.raises(RuntimeError, match="initialize"):
        u.runSimulationOnGPU(1)


def test_cleanup_allows_replaying_the_setter_sequence():
    u = _filled(2)
    u.cleanup()
    u.setNrOfAllMonomers(4)   # the reference forgets to reset the count here


# ---- C++ drop-in header -----------------------------------------------------------
CPP_DRIVER = r"""
#include <cstdio>
#include <stdexcept>
#include "UpdaterGPUScBFM_AB_Type.h"
int main(int argc, char** argv) {
    int fails = 0;
    UpdaterGPUScBFM_AB_Type u;
    if (argc > 1) {   // argv[1]: a seed file; the loaded state is its words 7..256
        u.loadR250SeedFile(argv[1], 0);
        uint32_t st[250];
        u.getR250State(st);
        std::FILE* f = std::fopen(argv[1], "rb");
        uint32_t w[257];
        if (!f || std::fread(w, 4, 257, f) != 257) fails++;
        if (f) std::fclose(f);
        for (int k = 0; k < 250; ++k) fails += st[k] != w[7 + k];
        try { u.setR250State(nullptr); fails++; } catch (const std::invalid_argument&) {}
    }
    try { u.setPeriodicity(true, false, true); fails++; } catch (const std::invalid_argument&) {}
    u.setNrOfAllMonomers(2);
    try { u.setNrOfAllMonomers(2); fails++; } catch (const std::runtime_error&) {}
    try { u.setMonomerCoordinates(5, 0, 0, 0); fails++; } catch (const std::out_of_range&) {}
    try { u.setLatticeSize(12, 16, 16); fails++; } catch (const std::runtime_error&) {}
    try { u.setGpu(99); fails++; } catch (const std::invalid_argument&) {}
    u.setMonomerCoordinates(0, 0, 0, 0);
    u.setMonomerCoordinates(1, 2, 0, 0);
    u.setConnectivity(0, 1);
    u.setConnectivity(1, 0);
    u.setLatticeSize(16, 16, 16);
    u.populateLattice();
    if (u.getMonomerPositionInX(1) != 2) fails++;
    u.cleanup();
    std::printf("fails=%d\n", fails);
    return fails;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_cpp_dropin_header_compiles_links_and_throws(tmp_path):
    src = tmp_path / "driver.cpp"
    src.write_text(CPP_DRIVER)
    exe = tmp_path / "driver"
    libdir = P.library_path.parent
    subprocess.run(["g++", "-std=c++11", "-O1", "-I", str(ROOT / "include"), str(src), "-o", str(exe),
                    "-L", str(libdir), "-lpscbfm", f"-Wl,-rpath,{libdir}"], check=True)
    res = subprocess.run([str(exe), str(GOLDEN / "resultNormPscBFM.seeds")], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "fails=0" in res.stdout


# ---- coloring ---------------------------------------------------------------------
def _csr_from_pairs(n, pairs):
    s = Sys((8, 8, 8), np.zeros((n, 3), np.int32), np.asarray(pairs, np.uint32).reshape(-1, 2))
    return s.csr()


def _proper(colors, start, ids):
    for i in range(len(start) - 1):
        for k in range(start[i], start[i + 1]):
            if colors[i] == colors[ids[k]]:
                return False
    return True


def _graphs():
    rng = np.random.default_rng(3)
    yield "chain", 50, [(i, i + 1) for i in range(49)]
    yield "odd ring", 7, [(i, (i + 1) % 7) for i in range(7)]
    yield "stars", 40, [(0, k) for k in range(1, 5)] + [(k, k + 4) for k in range(1, 30)]
    yield "isolated", 5, []
    for trial in range(6):
        n = 200
        pairs = set()
        deg = np.zeros(n, int)
        for _ in range(300):
            a, b = (int(v) for v in rng.integers(0, n, 2))
            if a != b and deg[a] < 8 and deg[b] < 8 and (a, b) not in pairs and (b, a) not in pairs:
                pairs.add((a, b))
                deg[a] += 1
                deg[b] += 1
        yield f"random{trial}", n, sorted(pairs)
    for name in ["Ring_N6_n6.bfm", "Hydrogel_NStar_117_NoPerXYZ64.bfm"]:
        b = read_bfm(GOLDEN / name)
        yield name, b.n, [tuple(int(v) for v in p) for p in b.bonds]


@pytest.mark.parametrize("uniform", [True, False])
def test_coloring_is_proper(uniform):
    for name, n, pairs in _graphs():
        start, ids = _csr_from_pairs(n, pairs)
        col = P.color_graph(n, start, ids, uniform)
        assert _proper(col, start, ids), name


@pytest.mark.skipif(not oracle.ref_coloring_available(), reason="reference tree not present (oracle/_ref not built)")
@pytest.mark.parametrize("uniform", [True, False])
def test_coloring_equals_reference_graph_coloring(uniform):
    for name, n, pairs in _graphs():
        start, ids = _csr_from_pairs(n, pairs)
        assert np.array_equal(P.color_graph(n, start, ids, uniform), oracle.ref_coloring(n, start, ids, uniform)), name
    b = read_bfm(GOLDEN / "LinearChain_l1024_n64.bfm.gz")
    start, ids = b.adjacency()
    assert np.array_equal(P.color_graph(b.n, start, ids), oracle.ref_coloring(b.n, start, ids))


def test_hydrogel_needs_more_than_two_species():
    b = read_bfm(GOLDEN / "Hydrogel_NStar_117_NoPerXYZ64.bfm")
    start, ids = b.adjacency()
    # more bonds per monomer than the reference's MAX_CONNECTIVITY of 4
    assert int(np.diff(start.astype(np.int64)).max()) > 4
    col = P.color_graph(b.n, start, ids)
    assert col.max() + 1 >= 2 and _proper(col, start, ids)


# ---- BFM files ----------------------------------------------------------------------
def test_read_reference_fixtures():
    lc = read_bfm(GOLDEN / "LinearChain_l1024_n64.bfm.gz")
    assert lc.n == 65536 and lc.box == (128, 128, 128) and lc.periodic == (True, True, True)
    assert lc.bonds.shape == (1024 * 63, 2) and len(lc.bondset) == 108
    assert np.array_equal(np.bincount(lc.attributes), [0, 32768, 32768])
    ring = read_bfm(GOLDEN / "Ring_N6_n6.bfm")
    assert ring.n == 42 and ring.periodic == (False, False, False) and ring.bonds.shape[0] == 36
    gel = read_bfm(GOLDEN / "Hydrogel_NStar_117_NoPerXYZ64.bfm")
    assert gel.n == 621 and gel.box == (64, 64, 64)


def test_linear_chain_fixture_is_a_valid_state():
    lc = read_bfm(GOLDEN / "LinearChain_l1024_n64.bfm.gz")
    start, ids = lc.adjacency()
    o = oracle.System(lc.box, lc.xyz, start, ids, P.color_graph(lc.n, start, ids))
    assert o.check() == 0


def test_bfm_round_trip(tmp_path):
    for name in ["LinearChain_l1024_n64.bfm.gz", "Ring_N6_n6.bfm", "Hydrogel_NStar_117_NoPerXYZ64.bfm"]:
        a = read_bfm(GOLDEN / name)
        out = tmp_path / "rt.bfm"
        write_bfm(out, a.box, a.xyz, a.bonds, a.bondset, a.attributes, mcs=a.mcs, periodic=a.periodic)
        b = read_bfm(out)
        assert np.array_equal(a.xyz, b.xyz) and a.box == b.box and a.periodic == b.periodic
        key = lambda arr: sorted(tuple(sorted(map(int, p))) for p in arr)
        assert key(a.bonds) == key(b.bonds)
        assert np.array_equal(a.attributes, b.attributes)


# ---- synthetic systems --------------------------------------------------------------
def test_generated_systems_are_valid():
    for s in [chains((64, 64, 64), 64, 64, seed=1), ordered_chains((64, 64, 64), 512, 32),
              ordered_chains((32, 64, 128), 200, 100), stars((64, 64, 64), 30, 4, 29, seed=2)]:
        start, ids = s.csr()
        o = oracle_for(s, P.color_graph(s.n, start, ids), 0)
        assert o.check() == 0


def test_ordered_packing_reaches_half_filling():
    box = (64, 64, 64)
    n = 64 ** 3 // 16
    s = ordered_chains(box, n // 64, 64)
    assert 8 * s.n / 64 ** 3 == pytest.approx(0.5)


def test_saw_growth_is_seeded_and_reports_jams():
    a = P.gen_linear_chains((32, 32, 32), 10, 20, seed=4)
    b = P.gen_linear_chains((32, 32, 32), 10, 20, seed=4)
    assert np.array_equal(a, b)
    with pytest.raises(RuntimeError):
        P.gen_linear_chains((16, 16, 16), 200, 20, seed=1, max_restarts=50)


def test_star_network_degrees():
    s = stars((64, 64, 64), 20, 4, 29, seed=3)
    start, _ = s.csr()
    deg = np.diff(start.astype(np.int64))
    assert deg.max() <= 4 and deg[0] == 4
    assert s.bonds.shape[0] > 20 * 4 * 29   # some crosslinks formed


def test_dense_stars_fill_the_ordered_packing():
    """Stars at occupied fraction 0.714 on 256^3: a valid state (no overlap,
    every bond in the set), all bonds inside stars, two species (every bond
    of the packing joins sites of opposite parity); seeded; a box that holds
    fewer stars is refused."""
    sd = W.dense_stars(42)
    assert sd.n == 12800 * 117 and 8 * sd.n / 256 ** 3 == pytest.approx(0.714, abs=1e-3)
    assert sd.bonds.shape == (12800 * 4 * 29, 2)
    assert np.all(sd.bonds[:, 0] // 117 == sd.bonds[:, 1] // 117)
    s = Sys(sd.box, sd.xyz, sd.bonds)
    start, ids = s.csr()
    colors = P.color_graph(s.n, start, ids)
    assert int(colors.max()) + 1 == 2
    assert oracle_for(s, colors, 0).check() == 0
    assert np.array_equal(sd.xyz, W.dense_stars(42).xyz) and not np.array_equal(sd.xyz, W.dense_stars(43).xyz)
    assert W.dense_stars(1, (256, 256, 256), 16384).n == 16384 * 117   # every place filled: 91% of the sites
    with pytest.raises(RuntimeError):
        W.dense_stars(1, (256, 256, 256), 16385)
    for box, n_stars, arm in [((32, 32, 32), 40, 5), ((64, 16, 32), 20, 3), ((16, 64, 16), 10, 2), ((8, 8, 8), 4, 1)]:
        t = W.dense_stars(3, box, n_stars, arm)
        u = Sys(t.box, t.xyz, t.bonds)
        st, ids = u.csr()
        assert oracle_for(u, P.color_graph(u.n, st, ids), 0).check() == 0


def test_dense_network_after_a_run_has_four_species():
    """The dense network as built for the bench: 100 sweeps of the stars on
    the oracle, then two rounds of arm-end crosslinks.  Valid state, at most
    two crosslinks per arm end, no bond twice, crosslinks only between
    stars, and the reference's coloring gives four species."""
    sd = W.dense_network(42, oracle_relax)
    n_star_bonds = 12800 * 4 * 29
    cross = sd.bonds[n_star_bonds:]
    assert len(cross) > 1.5 * 12800 * 4     # most arm ends linked in both rounds
    assert np.all(cross[:, 0] // 117 != cross[:, 1] // 117)
    key = np.sort(sd.bonds, axis=1)
    assert np.unique(key, axis=0).shape[0] == key.shape[0]
    assert np.bincount(cross[:, 0]).max() <= 2
    s = Sys(sd.box, sd.xyz, sd.bonds)
    start, ids = s.csr()
    assert np.diff(start.astype(np.int64)).max() <= 4
    colors = P.color_graph(s.n, start, ids)
    assert int(colors.max()) + 1 == 4
    assert oracle_for(s, colors, 0).check() == 0


def test_crosslink_rounds_and_arguments():
    sd = W.dense_stars(5, (64, 64, 64), 120, 8)
    moved = oracle_relax(sd, 30)
    one = P.crosslink_star_ends(sd.box, moved, 120, 4, 8, seed=5, rounds=1)
    two = P.crosslink_star_ends(sd.box, moved, 120, 4, 8, seed=5, rounds=2)
    assert len(P.crosslink_star_ends(sd.box, moved, 120, 4, 8, seed=5, rounds=0)) == 0
    assert np.array_equal(two[: len(one)], one)           # round one is the same either way
    assert np.bincount(one[:, 0]).max() == 1 and np.bincount(two[:, 0]).max() <= 2 and len(two) > len(one)
    with pytest.raises(P.InvalidArgument):
        P.crosslink_star_ends(sd.box, moved[:-1], 120, 4, 8, seed=5)


def test_native_bfm_reader_and_writer_match_python(tmp_path):
    src = GOLDEN / "LinearChain_l1024_n64.bfm.gz"
    ref = read_bfm(src)
    u = P.UpdaterGPUScBFM_AB_Type()
    assert u.loadBfm(src) == ref.mcs
    assert np.array_equal(u.getPositions(ref.n), ref.xyz)
    out = tmp_path / "native.bfm"
    u.writeBfm(out, 7)
    u.writeBfm(out, 9, append=True)
    back = read_bfm(out)
    assert back.mcs == 9 and [f[0] for f in back.frames] == [7, 9]
    assert np.array_equal(back.xyz, ref.xyz) and back.box == ref.box
    key = lambda arr: sorted(tuple(sorted(map(int, p))) for p in arr)
    assert key(back.bonds) == key(ref.bonds)
    assert np.array_equal(back.attributes, ref.attributes)
    assert {k: tuple(v) for k, v in back.bondset.items()} == {k: tuple(v) for k, v in ref.bondset.items()}


def test_native_bfm_reader_rejects_non_periodic_fixtures():
    # the reference's setPeriodicity accepts only periodic boxes (UpdaterGPUScBFM_AB_Type.cu:728-753)
    for name in ["Ring_N6_n6.bfm", "Hydrogel_NStar_117_NoPerXYZ64.bfm"]:
        with pytest.raises(P.InvalidArgument):
            P.UpdaterGPUScBFM_AB_Type().loadBfm(GOLDEN / name)


def test_native_bfm_reader_errors(tmp_path):
    bad = tmp_path / "bad.bfm"
    bad.write_bytes(b"!number_of_monomers=2\n!box_x=16\n!box_y=16\n!box_z=16\n")
    with pytest.raises(RuntimeError):
        P.UpdaterGPUScBFM_AB_Type().loadBfm(bad)
    with pytest.raises(RuntimeError):
        P.UpdaterGPUScBFM_AB_Type().loadBfm(tmp_path / "missing.bfm")


def test_cli_prints_help():
    import subprocess
    exe = Path(P.__file__).parent / "pscbfm_run"
    res = subprocess.run([str(exe), "--help"], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0 and "--max-mcs" in res.stdout and "--save-interval" in res.stdout


def test_lemonade_bond_ids_match_the_reference_fixture():
    from pscbfm_amd.bfm import lemonade_bond_ids
    ref = read_bfm(GOLDEN / "LinearChain_l1024_n64.bfm.gz").bondset
    assert {k: tuple(v) for k, v in ref.items()} == lemonade_bond_ids()


# ---- R250 mode (host side; no kernel runs) -----------------------------------------
def test_library_r250_matches_known_answers_and_oracle():
    seeds = GOLDEN / "resultNormPscBFM.seeds"
    u = P.UpdaterGPUScBFM_AB_Type()
    u.loadR250SeedFile(seeds, 0)
    first = u.r250Next(2)
    # mainCUDASimulatorGPUScBFM_AB_Type.cpp:138-139
    assert int(first[0]) == 4223731124 and f"{first[1] / 2**32:.6f}" == "0.803876"
    o = oracle.R250.from_seed_file(seeds)
    o.rand32(), o.rand32()
    assert u.r250Next(1000).tolist() == [o.rand32() for _ in range(1000)]


def test_library_r250_state_round_trip_and_discard():
    seeds = GOLDEN / "resultNormPscBFM.seeds"
    a = P.UpdaterGPUScBFM_AB_Type()
    a.loadR250SeedFile(seeds, 2)
    a.r250Next(377)                      # leave the ring mid-way
    b = P.UpdaterGPUScBFM_AB_Type()
    b.setR250State(a.r250State())        # continues the same stream
    assert a.r250Next(600).tolist() == b.r250Next(600).tolist()
    c = P.UpdaterGPUScBFM_AB_Type()
    c.loadR250SeedFile(seeds, 2)
    d = P.UpdaterGPUScBFM_AB_Type()
    d.loadR250SeedFile(seeds, 0)
    d.r250Next(2)
    assert c.r250Next(50).tolist() == d.r250Next(50).tolist()


def test_library_r250_errors(tmp_path):
    u = P.UpdaterGPUScBFM_AB_Type()
    with pytest.raises(Exception, match="R250"):
        u.r250Next(1)                    # no state yet
    short = tmp_path / "short.seeds"
    short.write_bytes(b"\0" * (4 * 256))
    with pytest.raises(Exception, match="fewer than 257"):
        u.loadR250SeedFile(short)
    with pytest.raises(Exception, match="cannot open"):
        u.loadR250SeedFile(tmp_path / "missing.seeds")
    with pytest.raises(ValueError):
        u.setR250State(np.zeros(10, dtype=np.uint32))


def _cli(*args):
    exe = Path(P.__file__).parent / "pscbfm_run"
    return subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=120)


def test_cli_reports_input_errors_before_the_gpu(tmp_path):
    out = tmp_path / "out.bfm"
    r = _cli("-i", tmp_path / "missing.bfm", "-m", 1, "-o", out)
    assert r.returncode == 1 and "cannot open BFM file" in r.stderr
    r = _cli("-i", GOLDEN / "Ring_N6_n6.bfm", "-m", 1, "-o", out)  # non-periodic fixture
    assert r.returncode == 1 and "periodic" in r.stderr
    short = tmp_path / "short.seeds"
    short.write_bytes(b"\0" * 40)
    r = _cli("-i", GOLDEN / "LinearChain_l1024_n64.bfm.gz", "-m", 1, "-e", short, "-o", out)
    assert r.returncode == 1 and "fewer than 257" in r.stderr
    r = _cli("-i", GOLDEN / "LinearChain_l1024_n64.bfm.gz", "-m", 1, "-e", short, "--seed", 3, "-o", out)
    assert r.returncode == 1 and "give one of them" in r.stderr
    r = _cli("-m", 1)
    assert r.returncode == 1 and "required" in r.stderr


def test_cli_seed_file_prints_the_reference_known_answers(tmp_path):
    # mainCUDASimulatorGPUScBFM_AB_Type.cpp:138-139: the two draws before the run
    r = _cli("-i", GOLDEN / "LinearChain_l1024_n64.bfm.gz", "-m", 1, "-e", GOLDEN / "resultNormPscBFM.seeds",
             "-o", tmp_path / "out.bfm")
    assert "RNG rand: 4223731124" in r.stderr and "RNG rand: 0.803876" in r.stderr
    # without a GPU the run stops at device selection; with one it completes
    assert r.returncode == 0 or "setGpu" in r.stderr


# ---- run setup without a GPU -------------------------------------------------------
def test_graph_chunk_is_bounded():
    u = P.UpdaterGPUScBFM_AB_Type()
    u.setGraphChunk(4096)
    u.setGraphChunk(0)
    for bad in (4097, 2**31, 2**32 - 1):
        with pytest.raises(P.InvalidArgument, match="4096"):
            u.setGraphChunk(bad)


def test_prepare_graphs_before_initialize_is_runtime_error():
    u = _filled(1)
    with pytest.raises(RuntimeError, match="initialize"):
        u.prepareGraphs(10)


def _staged_without_gpu(s: Sys):
    from pscbfm_amd import workloads as W
    sd = W.SystemData(s.box, s.xyz, s.bonds, 1)
    u = P.UpdaterGPUScBFM_AB_Type()
    u.setPeriodicity(True, True, True)
    u.setNrOfAllMonomers(sd.n)
    u.setMonomerCoordinatesBulk(sd.xyz)
    u.setConnectivityBulk(W.directed_pairs(sd.n, sd.bonds))
    u.setLatticeSize(*sd.box)
    forb = W.standard_forbidden()
    for dx, dy, dz in itertools.product(range(-4, 4), repeat=3):
        u.copyBondSet(dx, dy, dz, bool(forb[(dx & 7) | ((dy & 7) << 3) | ((dz & 7) << 6)]))
    return u


def test_partner_codes_of_chains_need_no_escape():
    # ordered and grown chains keep every partner near its own rank in the
    # partner's species, so all partners have a 16-bit code
    for s in (ordered_chains((64, 64, 64), 256, 32), chains((64, 64, 64), 40, 64, seed=3)):
        partners, escapes = _staged_without_gpu(s).partnerEscapes()
        assert partners == 2 * s.bonds.shape[0] and escapes == 0


def test_crosslinked_network_has_partner_escapes():
    # the network of the GPU test of scattered partners: crosslinks join
    # monomers whose slots lie far apart, or in another species than the
    # row's common base, which the 16-bit codes cannot span; in either slot
    # order
    s = stars((128, 128, 128), 1000, 4, 29, seed=7)
    for spatial in (False, True):
        u = _staged_without_gpu(s)
        u.setSlotOrder(spatial)
        partners, escapes = u.partnerEscapes()
        assert partners == 2 * s.bonds.shape[0]
        assert 0 < escapes < partners


def test_perform_walk_is_validated():
    u = P.UpdaterGPUScBFM_AB_Type()
    u.setPerformWalk(0, 0)
    u.setPerformWalk(128, 4)
    u.setPerformWalk(P.LEGACY_WALK, 0)
    for group, depth in ((32, 4), (320, 4), (64, 2), (64, 3), (64, 8)):
        with pytest.raises(P.InvalidArgument):
            u.setPerformWalk(group, depth)


def test_slot_order_is_validated():
    u = P.UpdaterGPUScBFM_AB_Type()
    for order in (None, "auto", True, "spatial", False, "reference"):
        u.setSlotOrder(order)
    with pytest.raises(P.InvalidArgument):
        u.setSlotOrder("hilbert")
    # the C ABI takes 0, 1 or 2 only
    assert u._lib.pscbfm_set_slot_order(u._h, 3) == 1  # PSCBFM_INVALID_ARGUMENT


def test_clear_pass_is_validated():
    u = P.UpdaterGPUScBFM_AB_Type()
    for p in (0, 1, 2):
        u.setClearPass(p)
    for p in (3, 7):
        with pytest.raises(P.InvalidArgument):
            u.setClearPass(p)


def test_device_code_has_no_flat_memory_instructions(tmp_path):
    # Every device buffer is reached through a global-address-space pointer
    # (descriptor pointers are DeviceBuffer, readable on the device only as
    # global()); no kernel may fall back to generic FLAT loads or stores.
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    csrc = ROOT / "pscbfm_amd" / "csrc"
    asm = tmp_path / "kernels.s"
    for src in sorted(csrc.glob("*.hip")):
        r = subprocess.run([hipcc, "-O3", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "-Xclang",
                            "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S",
                            "-I", str(csrc), "-I", str(ROOT / "include"), str(src), "-o", str(asm)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        text = asm.read_text()
        assert re.search(r"^\s+global_load", text, re.M) or src.name != "kernels.hip"
        flat = re.findall(r"^\s+(flat_(?:load|store|atomic)\w*)", text, re.M)
        assert not flat, f"{src.name}: {sorted(set(flat))}"


def test_read_back_switch_before_initialize():
    u = _filled(1)
    u.setReadBack(False)
    u.setReadBack(True)
    with pytest.raises(RuntimeError, match="initialize"):
        u.readPositionsBack()
