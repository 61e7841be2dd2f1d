"""GPU parity of the batched sweep: every replica of a batch must end exactly
where the same system ends when run alone in an updater (which the
single-system tests pin to the oracle), whatever else shares the launch, and
whenever replicas join or leave."""
import numpy as np
import pytest

import pscbfm_amd as P
from helpers import chains, gpu_vs_oracle, make_updater, ordered_chains, stage_updater, stars

pytestmark = pytest.mark.gpu


def _alone(s, seed, plan, kind=P.LATTICE_BIT):
    """Positions and accepted count after running `plan` (MCS per call) alone."""
    u = make_updater(s, seed, kind)
    acc = 0
    for m in plan:
        st, _ = u.run(m)
        acc += st.accepted
    x = u.getPositions(s.n)
    u.cleanup()
    return x, acc


def _mix():
    return [
        (chains((32, 32, 32), 20, 16, seed=7), 42),
        (chains((64, 64, 64), 40, 64, seed=11), 5),
        (chains((32, 64, 128), 30, 32, seed=3), 9),
        (stars((64, 64, 64), 40, 4, 29, seed=9), 77),
        (ordered_chains((32, 32, 32), 64, 32), 1),
        (chains((16, 16, 16), 1, 1, seed=2), 8),
    ]


@pytest.mark.parametrize("kind", [P.LATTICE_BIT, P.LATTICE_U8])
def test_batch_equals_each_system_alone(kind):
    systems = _mix()
    print("batch mix", kind, [s.box for s, _ in systems], flush=True)
    b = P.Batch(tile_capacity=256, replica_capacity=16, lattice_kind=kind)
    ids = [b.add(stage_updater(s, seed, kind)) for s, seed in systems]
    b.run(20)
    for rid, (s, seed) in zip(ids, systems):
        x_alone, acc_alone = _alone(s, seed, [20], kind)
        info = b.info(rid)
        assert info.substeps == 20 * info.n_colors
        assert info.accepted == acc_alone
        assert np.array_equal(b.positions(rid, s.n), x_alone)


@pytest.mark.parametrize("kind", [P.LATTICE_BIT, P.LATTICE_U8])
def test_positions_many_equals_one_by_one(kind):
    systems = _mix()
    b = P.Batch(tile_capacity=256, replica_capacity=16, lattice_kind=kind)
    ids = [b.add(stage_updater(s, seed, kind)) for s, seed in systems]
    b.run(7)
    ns = [s.n for s, _ in systems]
    # any order and a repeated id
    order = [ids[3], ids[0], ids[5], ids[3], ids[1], ids[2], ids[4]]
    many = b.positions_many(order, [ns[ids.index(i)] for i in order])
    for rid, x in zip(order, many):
        assert np.array_equal(x, b.positions(rid, ns[ids.index(rid)]))
    assert b.positions_many([], []) == []
    with pytest.raises(P.InvalidArgument):
        b.positions_many([ids[0]], [ns[0] + 1])
    with pytest.raises(P.InvalidArgument):
        b.positions_many([99], [1])


def test_batch_replica_matches_oracle():
    s, seed = chains((32, 32, 32), 20, 16, seed=7), 42
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=seed, n_mcs=30)
    assert np.array_equal(gx, ox)
    b = P.Batch(tile_capacity=64, replica_capacity=4)
    others = [b.add(stage_updater(o, sd)) for o, sd in _mix()[1:3]]
    rid = b.add(stage_updater(s, seed))
    b.run(30)
    assert np.array_equal(b.positions(rid, s.n), ox)
    assert b.info(rid).accepted == int(olog.sum())
    assert all(b.info(o).live for o in others)


def test_alone_in_batch_equals_in_mix():
    s, seed = chains((64, 64, 64), 40, 64, seed=11), 5
    solo = P.Batch(tile_capacity=64, replica_capacity=2)
    r0 = solo.add(stage_updater(s, seed))
    solo.run(25)
    x_solo = solo.positions(r0, s.n)
    mix = P.Batch(tile_capacity=256, replica_capacity=16)
    for o, sd in _mix():
        mix.add(stage_updater(o, sd))
    r1 = mix.add(stage_updater(s, seed))
    mix.run(25)
    assert np.array_equal(mix.positions(r1, s.n), x_solo)


def test_staggered_targets_and_churn():
    """Replicas stop at their own targets; others join and leave between
    launches (continuous batching); each ends where it ends alone."""
    sys_a = (chains((32, 32, 32), 20, 16, seed=7), 42)
    sys_b = (chains((64, 64, 64), 40, 64, seed=11), 5)
    sys_c = (stars((64, 64, 64), 40, 4, 29, seed=9), 77)
    b = P.Batch(tile_capacity=128, replica_capacity=4)
    a = b.add(stage_updater(*sys_a))
    bb = b.add(stage_updater(*sys_b))
    b.set_target(a, 10)
    b.set_target(bb, 30)
    b.launch(7)
    b.synchronize()
    b.launch(400)  # more than anyone needs; replicas idle at their stops
    b.synchronize()
    ia, ib = b.info(a), b.info(bb)
    assert ia.substeps == 10 * ia.n_colors and ib.substeps == 30 * ib.n_colors
    xa = b.positions(a, sys_a[0].n)
    b.remove(a)
    c = b.add(stage_updater(*sys_c))
    b.set_target(c, 15)
    b.set_target(bb, 5)  # five more for b: 35 in total
    b.launch(1000)
    b.synchronize()
    ic = b.info(c)
    assert ic.substeps == 15 * ic.n_colors
    assert b.info(bb).substeps == 35 * ib.n_colors
    assert np.array_equal(xa, _alone(*sys_a, [10])[0])
    assert np.array_equal(b.positions(bb, sys_b[0].n), _alone(*sys_b, [35])[0])
    assert np.array_equal(b.positions(c, sys_c[0].n), _alone(*sys_c, [15])[0])
    live, used, cap, _ = b.usage()
    assert live == 2 and used == b.info(bb).tiles + ic.tiles and cap == 128


def test_large_replica_spread_over_all_xcds():
    # 64 chains of 1024 on 256^3: 64+ workgroups of 512 monomers, spread round-robin.
    s, seed = chains((256, 256, 256), 64, 1024, seed=42), 42
    b = P.Batch(tile_capacity=512, replica_capacity=4)
    small = b.add(stage_updater(chains((32, 32, 32), 20, 16, seed=7), 42))
    rid = b.add(stage_updater(s, seed))
    assert b.info(rid).tiles >= 64
    b.run(10)
    assert np.array_equal(b.positions(rid, s.n), _alone(s, seed, [10])[0])
    assert b.info(small).live


def test_batch_mixes_slot_orders():
    # Replicas staged in Z-curve slot order and in monomer order share the
    # launches; each ends where it ends alone.
    systems = _mix()
    b = P.Batch(tile_capacity=256, replica_capacity=16)
    ids = [b.add(stage_updater(s, seed, spatial=(k % 2 == 0))) for k, (s, seed) in enumerate(systems)]
    b.run(10)
    for rid, (s, seed) in zip(ids, systems):
        assert np.array_equal(b.positions(rid, s.n), _alone(s, seed, [10])[0])


def test_read_back_in_rounds_equals_alone():
    # The batch decodes coordinates on the device into a buffer of at most
    # 64 MiB of records (5,592,405 monomers) and copies them back in rounds:
    # the 8.4M-monomer melt spans two rounds, so its ranges are split, and a
    # request that repeats it starts a range in the middle of a round.
    melt, mseed = ordered_chains((512, 512, 512), 32768, 256), 42
    small, sseed = chains((32, 32, 32), 20, 16, seed=7), 3
    print("read-back rounds", melt.n, small.n, flush=True)
    b = P.Batch(tile_capacity=0, replica_capacity=4)
    rs = b.add(stage_updater(small, sseed))
    rm = b.add(stage_updater(melt, mseed))
    b.run(2)
    want_m = _alone(melt, mseed, [2])[0]
    want_s = _alone(small, sseed, [2])[0]
    got = b.positions_many([rs, rm, rs, rm], [small.n, melt.n, small.n, melt.n])
    assert np.array_equal(got[0], want_s) and np.array_equal(got[2], want_s)
    assert np.array_equal(got[1], want_m) and np.array_equal(got[3], want_m)
    assert np.array_equal(b.positions(rm, melt.n), want_m)


def test_capacity_and_id_errors():
    s = chains((32, 32, 32), 20, 16, seed=7)
    b = P.Batch(tile_capacity=8, replica_capacity=2)
    _, tiles = b.footprint(stage_updater(s, 1))
    assert tiles >= 1
    ids = [b.add(stage_updater(s, 1)), b.add(stage_updater(s, 2))]
    with pytest.raises(RuntimeError):
        b.add(stage_updater(s, 3))  # replica capacity
    b.remove(ids[0])
    with pytest.raises(P.InvalidArgument):
        b.remove(ids[0])
    with pytest.raises(P.InvalidArgument):
        b.info(99)
    big = chains((128, 128, 128), 64, 256, seed=3)
    with pytest.raises(RuntimeError):
        b.add(stage_updater(big, 1))  # tile capacity
    assert b.usage()[0] == 1


def test_growable_table_equals_alone():
    systems = _mix()
    b = P.Batch(0, len(systems))  # table starts at 8 workgroups and grows
    ids = [b.add(stage_updater(s, seed)) for s, seed in systems]
    live, used, cap, _ = b.usage()
    assert live == len(systems) and used <= cap and cap % 8 == 0
    b.run(12)
    for rid, (s, seed) in zip(ids, systems):
        assert np.array_equal(b.positions(rid, s.n), _alone(s, seed, [12])[0])


def test_autotune_leaves_the_trajectory_alone():
    s, seed = chains((64, 64, 64), 40, 64, seed=11), 5
    u = make_updater(s, seed)
    u.run(6)
    block, slots, trials = u.autotune(3)
    assert (block, slots) in trials and len(trials) == len(P.TUNE_CANDIDATES)
    assert min(trials.values()) == trials[(block, slots)]
    u.run(9)
    assert np.array_equal(u.getPositions(s.n), _alone(s, seed, [15])[0])


def test_batch_autotune_picks_a_candidate():
    staged = [stage_updater(s, seed) for s, seed in _mix()]
    block, slots, trials = P.autotune_batch(staged, mcs=2)
    assert (block, slots) in P.TUNE_CANDIDATES
    assert trials[(block, slots)] == min(trials.values())


@pytest.mark.parametrize("kind", [P.LATTICE_BIT, P.LATTICE_U8])
def test_batch_mixing_degree_buckets_matches_alone_and_oracle(kind):
    # the batch kernel runs at the widest neighbour table of its replicas (8
    # here) over replicas whose own tables are narrower (0, 2, 4)
    from helpers import GOLDEN, Sys, oracle_for
    from pscbfm_amd.bfm import read_bfm
    gel = read_bfm(GOLDEN / "Hydrogel_NStar_117_NoPerXYZ64.bfm")
    systems = [
        (stars((64, 64, 64), 20, 8, 10, seed=3), 5),
        (Sys(gel.box, gel.xyz, gel.bonds), 6),
        (chains((32, 32, 32), 20, 16, seed=7), 42),
        (stars((64, 64, 64), 40, 4, 29, seed=9), 77),
        (chains((16, 16, 16), 1, 1, seed=2), 8),
    ]
    b = P.Batch(tile_capacity=128, replica_capacity=8, lattice_kind=kind)
    ids = [b.add(stage_updater(s, seed, kind)) for s, seed in systems]
    b.run(25)
    for rid, (s, seed) in zip(ids, systems):
        u = make_updater(s, seed, kind)
        colors = u.colors(s.n)
        u.run(25)
        x_alone = u.getPositions(s.n)
        u.cleanup()
        o = oracle_for(s, colors, seed)
        total, _ = o.run(0, 25 * (int(colors.max()) + 1))
        got = b.positions(rid, s.n)
        assert np.array_equal(got, x_alone) and np.array_equal(got, o.xyz)
        assert b.info(rid).accepted == total
