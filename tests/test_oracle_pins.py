"""Pins of the CPU oracle (CPU only).

Each function of oracle/oracle.c is checked against something that does not
restate its own formula: golden vectors evaluated from the reference's text,
published known answers, the reference's fixture data, brute force, or the
naive O(n^2) definition of the two-pass scheme.
"""
import itertools
import json

import numpy as np
import pytest

import oracle
import pscbfm_amd as P
from helpers import GOLDEN, Sys, chains, oracle_for, ordered_chains

G = json.loads((GOLDEN / "golden_vectors.json").read_text())


# ---- move hash (CRITICAL.cu:2-11) --------------------------------------------
def test_hash_matches_vectors_evaluated_from_reference_text():
    for a, h in G["hash"]:
        assert oracle.hash32(a) == h
        assert P.move_hash(a) == h


def test_kernel_draw_composition_matches_reference_text():
    # rn = H(H(j) ^ seed), direction = rn % 6 (CRITICAL.cu:321-323)
    for j, seed, rn, d in G["draws"]:
        assert oracle.hash32(oracle.hash32(j) ^ seed) == rn
        assert rn % 6 == d


# ---- counter RNG ----------------------------------------------------------------
def test_splitmix64_published_known_answers():
    # First outputs of SplitMix64 seeded with 1234567 (reference implementation
    # by S. Vigna, as used to seed xoshiro/xoroshiro generators).
    expected = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                4593380528125082431, 16408922859458223821]
    assert [oracle.splitmix64(1234567, t) for t in range(5)] == expected


def test_substep_draw_oracle_equals_library():
    for seed in (0, 42, 2 ** 63 + 5):
        for n_colors in (1, 2, 3, 7, 254):
            for t in list(range(50)) + [2 ** 32 + 3, 2 ** 40]:
                assert oracle.substep_draw(seed, t, n_colors) == P.substep_draw(seed, t, n_colors)


def test_substep_species_roughly_uniform():
    counts = np.bincount([oracle.substep_draw(42, t, 4)[0] for t in range(40000)], minlength=4)
    assert counts.min() > 9500 and counts.max() < 10500


# ---- bond set and bond index (CRITICAL.cu:243-261) -------------------------------
def test_bondset_matches_reference_fixture():
    forb = oracle.standard_bondset()
    fixture = {tuple(v[:3]) for v in G["bondset"]}
    assert len(fixture) == 108
    allowed = {(dx, dy, dz) for dx, dy, dz in itertools.product(range(-4, 4), repeat=3)
               if not forb[oracle.bond_index(dx, dy, dz)]}
    assert allowed == fixture
    assert {tuple(int(c) for c in v) for v in P.standard_bond_vectors()} == fixture


def test_bond_index_is_component_mod_8_and_aliases_plus_minus_4():
    for dx, dy, dz in itertools.product(range(-8, 9), repeat=3):
        i = oracle.bond_index(dx, dy, dz)
        assert i == (dx % 8) + 8 * (dy % 8) + 64 * (dz % 8)
    assert oracle.bond_index(4, 0, 0) == oracle.bond_index(-4, 0, 0)


def test_direction_table_matches_reference_text():
    import ctypes as C
    for d, vec in enumerate(G["directions"]):
        x, y, z = C.c_int32(), C.c_int32(), C.c_int32()
        oracle.lib().oracle_direction(d, C.byref(x), C.byref(y), C.byref(z))
        assert [x.value, y.value, z.value] == vec


# ---- Z-curve index (UpdaterGPUScBFM_AB_Type.cu:150-176, 210-220) -----------------
def _interleave(x, y, z, bits):
    out = 0
    for i in range(bits):
        out |= ((x >> i) & 1) << (3 * i) | ((y >> i) & 1) << (3 * i + 1) | ((z >> i) & 1) << (3 * i + 2)
    return out


def test_zcurve_bruteforce_cubic():
    rng = np.random.default_rng(1)
    for L in (8, 64, 256, 1024):  # 1024^3 = 2^30 cells; the codes are 32-bit
        for x, y, z in rng.integers(0, L, size=(300, 3)):
            ref = _interleave(int(x), int(y), int(z), 11)
            assert oracle.zcurve_index(int(x), int(y), int(z), L, L, L) == ref
            assert P.zcurve_encode((L, L, L), int(x), int(y), int(z)) == ref


def test_zcurve_noncubic_is_dense_bijection():
    for box in [(8, 16, 32), (32, 8, 8), (16, 64, 8), (128, 32, 32)]:
        V = box[0] * box[1] * box[2]
        codes = {P.zcurve_encode(box, x, y, z)
                 for x in range(box[0]) for y in range(box[1]) for z in range(box[2])}
        assert len(codes) == V and max(codes) == V - 1


def test_zcurve_low_nine_bits_interleave_for_every_box():
    # the bond key uses the low 9 bits as xyz xyz xyz: true whenever each edge >= 8
    for box in [(8, 8, 8), (8, 128, 32), (64, 8, 16)]:
        for x, y, z in itertools.product(range(8), repeat=3):
            assert P.zcurve_encode(box, x, y, z) & 0x1FF == _interleave(x, y, z, 3)


# ---- front plane (CRITICAL.cu:62-176) ----------------------------------------------
def _cubes_overlap(a, b, L):
    return all(((a[k] - b[k]) % L) in (0, 1, L - 1) for k in range(3))


def test_front_plane_equals_cube_overlap_bruteforce():
    L = 8
    box = (L, L, L)
    p = (3, 3, 3)
    dirs = G["directions"]
    for other in itertools.product(range(L), repeat=3):
        if _cubes_overlap(p, other, L):
            continue  # not a valid start state
        occ = np.zeros(L ** 3, np.uint8)
        occ[other[0] + L * (other[1] + L * other[2])] = 1
        for d in range(6):
            q = tuple(p[k] + dirs[d][k] for k in range(3))
            assert oracle.front_blocked(occ, box, p, d) == _cubes_overlap(q, other, L)


def test_dimer_move_legality_enumerated():
    # monomer at the origin, partner at bond vector b: a move by e is legal iff
    # b - e is a bond vector (every bond vector keeps the cubes apart)
    L = 16
    box = (L, L, L)
    forb = oracle.standard_bondset()
    dirs = G["directions"]
    bonds = [tuple(v[:3]) for v in G["bondset"]]
    legal = 0
    for b in bonds:
        p = (5, 5, 5)
        other = tuple(p[k] + b[k] for k in range(3))
        occ = np.zeros(L ** 3, np.uint8)
        occ[(other[0] % L) + L * ((other[1] % L) + L * (other[2] % L))] = 1
        for d in range(6):
            nb = tuple(b[k] - dirs[d][k] for k in range(3))
            bond_ok = not forb[oracle.bond_index(*nb)]
            ok = bond_ok and not oracle.front_blocked(occ, box, p, d)
            assert ok == (nb in bonds)
            legal += ok
    assert 0 < legal < 108 * 6


# ---- the two-pass scheme ------------------------------------------------------------
def _colors_of(s: Sys):
    start, ids = s.csr()
    return P.color_graph(s.n, start, ids)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_two_pass_equals_naive_definition(seed):
    s = chains((16, 16, 16), 8, 12, seed=seed)
    col = _colors_of(s)
    a = oracle_for(s, col, seed=seed * 17)
    b = oracle_for(s, col, seed=seed * 17)
    ta, la = a.run(0, 300)
    tb, lb = b.run_naive(0, 300)
    assert ta == tb and np.array_equal(la, lb) and np.array_equal(a.xyz, b.xyz)
    assert a.check() == 0


def test_two_pass_equals_naive_dense():
    s = ordered_chains((16, 16, 16), 16, 16)   # occupied-site fraction 0.5
    col = _colors_of(s)
    a, b = oracle_for(s, col, 9), oracle_for(s, col, 9)
    _, la = a.run(0, 400)
    _, lb = b.run_naive(0, 400)
    assert np.array_equal(la, lb) and np.array_equal(a.xyz, b.xyz)
    assert la.sum() > 0


def test_processing_order_does_not_matter():
    s = chains((32, 32, 32), 30, 20, seed=5)
    col = _colors_of(s)
    ref = oracle_for(s, col, 77)
    ref.run(0, 200, threads=1)
    for shuffle in (1, 2, 99):
        o = oracle_for(s, col, 77)
        o.run(0, 200, threads=1, shuffle_seed=shuffle)
        assert np.array_equal(o.xyz, ref.xyz)
    par = oracle_for(s, col, 77)
    par.run(0, 200, threads=0)
    assert np.array_equal(par.xyz, ref.xyz)


def test_split_runs_equal_one_run():
    s = chains((32, 32, 32), 20, 20, seed=6)
    col = _colors_of(s)
    a = oracle_for(s, col, 5)
    _, la = a.run(0, 100)
    b = oracle_for(s, col, 5)
    _, l1 = b.run(0, 37)
    _, l2 = b.run(37, 63)
    assert np.array_equal(np.concatenate([l1, l2]), la) and np.array_equal(a.xyz, b.xyz)


def test_single_monomer_accepts_every_move():
    s = Sys((16, 16, 16), np.array([[1, 2, 3]], np.int32), np.zeros((0, 2), np.uint32))
    o = oracle_for(s, np.zeros(1, np.uint8), 3)
    total, log = o.run(0, 500)
    assert total == 500 and np.all(log == 1)


def test_invariants_hold_along_a_run():
    s = ordered_chains((32, 32, 32), 64, 32)
    col = _colors_of(s)
    o = oracle_for(s, col, 11)
    for k in range(10):
        o.run(k * 20, 20)
        assert o.check() == 0


def test_check_detects_overlap_and_bad_bond():
    s = Sys((16, 16, 16), np.array([[0, 0, 0], [1, 0, 0]], np.int32), np.zeros((0, 2), np.uint32))
    assert oracle_for(s, np.array([0, 0], np.uint8), 0).check() == 1
    s = Sys((16, 16, 16), np.array([[0, 0, 0], [4, 0, 0]], np.int32), np.array([[0, 1]], np.uint32))
    assert oracle_for(s, np.array([0, 1], np.uint8), 0).check() == 2


def test_sequential_baseline_statistics_match_colored_scheme():
    # Not bit-comparable (different update order); the mean squared bond
    # length after relaxation must agree within a few percent.
    s = chains((32, 32, 32), 40, 32, seed=8)
    col = _colors_of(s)
    a = oracle_for(s, col, 1)
    a.run(0, 2 * 300)
    b = oracle_for(s, col, 1)
    b.sequential(5, 300)

    def mean_b2(xyz):
        v = xyz[s.bonds[:, 1]] - xyz[s.bonds[:, 0]]
        return float((v.astype(np.int64) ** 2).sum(1).mean())
    assert abs(mean_b2(a.xyz) - mean_b2(b.xyz)) / mean_b2(b.xyz) < 0.05
    assert a.check() == 0 and b.check() == 0


# ---- R250 (the reference's per-substep draws) ---------------------------------------
def test_r250_hits_the_reference_known_answers():
    # mainCUDASimulatorGPUScBFM_AB_Type.cpp:137-139, seeded from tests/resultNormPscBFM.seeds
    # (testCorrectness.sh:208): std::rand() 764080779, r250_rand32() 4223731124, r250_drand() 0.803876
    import ctypes
    words = np.frombuffer((GOLDEN / "resultNormPscBFM.seeds").read_bytes(), dtype="<u4")
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(int(words[0])))
    assert libc.rand() == 764080779
    r = oracle.R250.from_seed_file(GOLDEN / "resultNormPscBFM.seeds")
    assert r.rand32() == 4223731124
    assert f"{r.rand32() / 2**32:.6f}" == "0.803876"


def test_r250_recurrence_closed_form():
    # the buffer form draws x[n] = x[n-250] ^ x[n-147] (x^250 + x^147 + 1, the reciprocal of the
    # 103-lag trinomial), with the 250 state words as x[0..249]; checked on 2000 outputs
    state = np.random.default_rng(3).integers(0, 2**32, 250, dtype=np.uint64).astype(np.uint32)
    r = oracle.R250(state)
    x = list(int(v) for v in state) + [r.rand32() for _ in range(2000)]
    for n in range(250, len(x)):
        assert x[n] == x[n - 250] ^ x[n - 147]


def test_oracle_explicit_draws_equal_counter_draws():
    s = chains((32, 32, 32), 20, 16, seed=7)
    start, ids = s.csr()
    col = P.color_graph(s.n, start, ids)
    nc = int(col.max()) + 1
    a = oracle_for(s, col, 42)
    ta, loga = a.run(0, 40)
    draws = np.array([oracle.substep_draw(42, t, nc) for t in range(40)], dtype=np.uint32).ravel()
    b = oracle_for(s, col, 999)  # the key is ignored when draws are given
    b.draws, b.draw_t0 = draws, 0
    tb, logb = b.run(0, 40)
    assert ta == tb and np.array_equal(loga, logb) and np.array_equal(a.xyz, b.xyz)


def test_oracle_r250_draws_are_order_independent_and_match_naive():
    s = chains((16, 16, 16), 6, 10, seed=3)
    start, ids = s.csr()
    col = P.color_graph(s.n, start, ids)
    draws = oracle.R250.from_seed_file(GOLDEN / "resultNormPscBFM.seeds").substep_draws(30, int(col.max()) + 1)
    res = []
    for mode in ("fast", "shuffled", "naive"):
        o = oracle_for(s, col, 0)
        o.draws = draws
        if mode == "fast":
            t, log = o.run(0, 30)
        elif mode == "shuffled":
            t, log = o.run(0, 30, shuffle_seed=11)
        else:
            t, log = o.run_naive(0, 30)
        res.append((t, log.tolist(), o.xyz.tolist()))
    assert res[0] == res[1] == res[2]
