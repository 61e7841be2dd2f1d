"""GPU parity of R250 mode: the substep draws come from the reference's R250
stream (CRITICAL.cu:509-510) seeded from a LeMonADE seed file, on the GPU and
in the oracle, and the two must agree bit for bit."""
import numpy as np
import pytest

import oracle
import pscbfm_amd as P
from helpers import GOLDEN, Sys, chains, make_updater, oracle_for, stage_updater

pytestmark = pytest.mark.gpu

SEEDS = GOLDEN / "resultNormPscBFM.seeds"
KINDS = [P.LATTICE_BIT, P.LATTICE_U8]


def oracle_r250_draws(n_substeps: int, n_colors: int, discard: int = 2) -> np.ndarray:
    r = oracle.R250.from_seed_file(SEEDS)
    for _ in range(discard):
        r.rand32()
    return r.substep_draws(n_substeps, n_colors)


def run_r250(s, kind, plan, seed=42, colors=None):
    """R250-mode runs of `plan` = [(n_mcs, use_graph, profile), ...]; returns
    (positions, log, stats list, updater colors, updater)."""
    u = make_updater(s, seed, kind, colors=colors)
    u.loadR250SeedFile(SEEDS, 2)
    logs, stats = [], []
    for m, graph, prof in plan:
        st, log = u.run(m, log=True, use_graph=graph, profile=prof)
        logs.append(log)
        stats.append(st)
    return u.getPositions(s.n), np.concatenate(logs), stats, u.colors(s.n), u


@pytest.mark.parametrize("kind", KINDS)
def test_r250_mode_bit_exact_over_split_runs(kind):
    s = chains((32, 32, 32), 20, 16, seed=7)
    plan = [(7, True, False), (1, False, False), (12, True, False), (2, False, True)]
    gx, glog, stats, colors, u = run_r250(s, kind, plan)
    n_mcs = sum(p[0] for p in plan)
    nc = int(colors.max()) + 1
    draws = oracle_r250_draws(n_mcs * nc, nc)
    o = oracle_for(s, colors, 0)
    o.draws = draws
    _, olog = o.run(0, n_mcs * nc)
    print(f"R250 kind {kind}: {int(glog.sum())} accepted over {glog.size} substeps", flush=True)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, o.xyz)
    # attempts: every member of each drawn species, as counted by the host
    counts = np.bincount(colors, minlength=nc)
    assert sum(st.attempted for st in stats) == int(counts[draws[0::2]].sum())
    # the generator advanced by exactly two draws per substep
    r = oracle.R250.from_seed_file(SEEDS)
    for _ in range(2 + 2 * n_mcs * nc):
        r.rand32()
    assert u.r250Next(5).tolist() == [r.rand32() for _ in range(5)]
    u.cleanup()


def test_r250_three_species_and_counter_mode_differs():
    # three species (position along the chain mod 3, a proper coloring): the
    # species draw is r250 % 3, not a power of two
    s = chains((64, 64, 64), 30, 24, seed=3)
    three = (np.arange(s.n) % 24 % 3).astype(np.uint8)
    gx, glog, _, colors, u = run_r250(s, P.LATTICE_BIT, [(30, True, False)], colors=three)
    nc = int(colors.max()) + 1
    assert nc == 3
    o = oracle_for(s, colors, 0)
    o.draws = oracle_r250_draws(30 * nc, nc)
    _, olog = o.run(0, 30 * nc)
    assert np.array_equal(glog, olog) and np.array_equal(gx, o.xyz)
    u.cleanup()
    c = make_updater(s, 42, colors=three)
    _, clog = c.run(30, log=True)
    assert not np.array_equal(clog, glog)  # the counter draw is a different stream
    c.cleanup()


def test_r250_state_handoff_between_updaters():
    # run 10 MCS, hand (positions, R250 state) to a fresh updater, run 10 more:
    # equals 20 MCS in one updater
    s = chains((32, 32, 32), 16, 20, seed=11)
    gx, glog, _, colors, u = run_r250(s, P.LATTICE_BIT, [(20, True, False)])
    u.cleanup()
    a = make_updater(s, 42)
    a.loadR250SeedFile(SEEDS, 2)
    a.run(10)
    mid, state = a.getPositions(s.n), a.r250State()
    a.cleanup()
    b = make_updater(Sys(s.box, mid, s.bonds), 42, colors=colors)
    b.setR250State(state)
    b.run(10)
    assert np.array_equal(b.getPositions(s.n), gx)
    b.cleanup()


def test_batch_and_server_reject_r250_updaters():
    s = chains((32, 32, 32), 4, 8, seed=1)
    u = stage_updater(s, 5)
    u.loadR250SeedFile(SEEDS, 2)
    batch = P.Batch(0, 2)
    with pytest.raises(Exception, match="R250"):
        batch.add(u)
    server = P.Server(hbm_budget=1 << 32, tile_capacity=64, replica_capacity=4)
    with pytest.raises(Exception, match="R250"):
        server.submit(u, 10)
    server.close()


@pytest.mark.parametrize("chunk", [2, 5, 14])
def test_r250_draw_chunks_equal_one_table(chunk):
    # a run longer than the device draw table goes through in chunks of whole
    # sweeps; positions, per-substep log and statistics are those of one table
    s = chains((32, 32, 32), 20, 16, seed=7)
    gx, glog, stats, colors, u = run_r250(s, P.LATTICE_BIT, [(23, True, False)])
    u.cleanup()
    v = make_updater(s, 42)
    v.loadR250SeedFile(SEEDS, 2)
    v.setDrawChunk(chunk)
    st, log = v.run(23, log=True)
    assert np.array_equal(log, glog)
    assert np.array_equal(v.getPositions(s.n), gx)
    assert (st.attempted, st.accepted, st.substeps) == (stats[0].attempted, stats[0].accepted, stats[0].substeps)
    assert v.r250Next(3).tolist() == u_next_after(23 * (int(colors.max()) + 1))
    v.cleanup()


def u_next_after(n_substeps: int):
    r = oracle.R250.from_seed_file(SEEDS)
    for _ in range(2 + 2 * n_substeps):
        r.rand32()
    return [r.rand32() for _ in range(3)]
