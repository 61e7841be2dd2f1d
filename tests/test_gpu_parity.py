"""GPU parity: the HIP sweep against the CPU oracle, bit for bit.

Every case compares final unfolded coordinates and the accepted count of
every substep.  Sizes are chosen so the oracle finishes in seconds.
"""
import numpy as np
import pytest

import pscbfm_amd as P
from helpers import chains, gpu_vs_oracle, make_updater, oracle_for, ordered_chains, stars, Sys

pytestmark = pytest.mark.gpu

KINDS = [P.LATTICE_BIT, P.LATTICE_U8]


@pytest.mark.parametrize("kind", KINDS)
def test_small_chains_bit_exact(kind):
    s = chains((32, 32, 32), 20, 16, seed=7)
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=42, n_mcs=50, kind=kind)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)
    assert glog.sum() > 0


@pytest.mark.parametrize("kind", KINDS)
def test_c1_shape_reduced_mcs(kind):
    # Config "plain single system": 64 chains of 1024 on 256^3, seed 42, at 20 MCS.
    s = chains((256, 256, 256), 64, 1024, seed=42)
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=42, n_mcs=20, kind=kind)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)


@pytest.mark.parametrize("kind", KINDS)
def test_dense_ordered_melt(kind):
    # phi = 0.5 ordered packing, the start state of the dense-melt workload, small box.
    box = (64, 64, 64)
    n_chains = 64 * 64 * 64 // 16 // 32
    s = ordered_chains(box, n_chains, 32)
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=3, n_mcs=30, kind=kind)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)


def test_split_calls_and_graph_off_equal_one_call():
    s = chains((64, 64, 64), 40, 64, seed=11)
    a = gpu_vs_oracle(s, seed=5, n_mcs=24, chunks=(24,))
    b = gpu_vs_oracle(s, seed=5, n_mcs=24, chunks=(1, 7, 16), use_graph=False)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(a[0], a[2]) and np.array_equal(a[1], a[3])


def test_star_network_multicolor():
    s = stars((64, 64, 64), 40, 4, 29, seed=9)
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=77, n_mcs=20)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)


@pytest.mark.parametrize("block,slots", [(64, 1), (64, 4), (128, 2), (256, 1), (192, 4)])
def test_launch_shape_does_not_change_results(block, slots):
    s = chains((64, 64, 64), 30, 40, seed=2)
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=8, n_mcs=10, block=block, slots=slots)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)


@pytest.mark.parametrize("kind", [P.LATTICE_BIT, P.LATTICE_U8])
def test_invariants_after_run(kind):
    s = chains((64, 64, 64), 50, 50, seed=4)
    u = make_updater(s, seed=1, kind=kind)
    u.runSimulationOnGPU(25)
    u.checkSystem()
    # the occupancy lattice holds exactly the wrapped corners of the
    # positions read back, one cell per monomer
    lat = u.downloadLattice(64 ** 3)
    corners = u.getPositions(s.n) % 64
    want = np.zeros(64 ** 3, np.uint8)
    want[[P.zcurve_encode(s.box, *c) for c in corners]] = 1
    assert int(lat.sum()) == s.n and np.array_equal(lat, want)
    u.cleanup()


@pytest.mark.parametrize("clear_pass", [1, 2])
@pytest.mark.parametrize("kind", [P.LATTICE_BIT, P.LATTICE_U8])
def test_proposal_lattice_after_run(kind, clear_pass):
    # The reference clears its proposal lattice before every substep
    # (CRITICAL.cu:561-573) and its legacy updater checks that it is empty
    # after a run.  Here substep t marks buffer t & 1 and its pass 1 or pass 2
    # clears the other buffer, so after substeps 0..T-1 the buffer the next
    # substep uses is empty and the other holds the last substep's marks: at
    # least one per accepted move (their targets are distinct), at most one
    # per member of the drawn species.
    s = chains((64, 64, 64), 50, 50, seed=4)
    u = make_updater(s, seed=1, kind=kind)
    u.setClearPass(clear_pass)
    _, log = u.run(25, log=True)
    steps = len(log)
    bits = u.proposalBits()
    assert bits[steps & 1] == 0
    assert 0 < log[-1] <= bits[(steps - 1) & 1] <= int(max(u.colorSizes()))
    u.cleanup()


@pytest.mark.parametrize("walk", [0, 128])
@pytest.mark.parametrize("kind", [P.LATTICE_BIT, P.LATTICE_U8])
def test_clear_pass_does_not_change_results(kind, walk):
    # Either pass of substep t may zero the other parity's proposal buffer:
    # pass 2 of substep t-1 was its last reader and pass 1 of substep t+1 is
    # its next writer.  Switching the pass between runs keeps that order, so
    # runs that alternate equal the oracle's uninterrupted run.
    s = chains((64, 64, 64), 40, 64, seed=13)
    u = make_updater(s, 9, kind)
    if walk:
        u.setPerformWalk(walk, 1)
    logs = []
    for clear_pass, m in [(1, 4), (2, 3), (1, 5), (0, 3)]:
        u.setClearPass(clear_pass)
        _, log = u.run(m, log=True)
        logs.append(log)
    x = u.getPositions(s.n)
    colors = u.colors(s.n)
    u.cleanup()
    o = oracle_for(s, colors, 9)
    _, olog = o.run(0, 15 * (int(colors.max()) + 1))
    assert np.array_equal(np.concatenate(logs), olog)
    assert np.array_equal(x, o.xyz)


@pytest.mark.parametrize("kind", [P.LATTICE_BIT, P.LATTICE_U8])
@pytest.mark.parametrize("system", ["network", "melt"])
def test_slot_order_does_not_change_results(kind, system):
    # Slots of a species in Z-curve order of the starting corners (the
    # default) or in monomer order: every move is drawn from the member's
    # index in the reference's order either way (CRITICAL.cu:322-323), so
    # both runs equal the oracle's, with the same species.
    s = stars((64, 64, 64), 40, 4, 29, seed=9) if system == "network" else ordered_chains((64, 64, 64), 64, 32)
    runs = []
    for spatial in (True, False):
        u = make_updater(s, 21, kind, spatial=spatial)
        _, log = u.run(12, log=True)
        runs.append((u.getPositions(s.n), log, u.colors(s.n)))
        u.cleanup()
    colors = runs[0][2]
    o = oracle_for(s, colors, 21)
    _, olog = o.run(0, 12 * (int(colors.max()) + 1))
    for x, log, c in runs:
        assert np.array_equal(c, colors)
        assert np.array_equal(log, olog)
        assert np.array_equal(x, o.xyz)


def test_single_monomer_always_moves():
    s = Sys((16, 16, 16), np.array([[3, 4, 5]], np.int32), np.zeros((0, 2), np.uint32))
    u = make_updater(s, seed=0)
    stats, log = u.run(100, log=True)
    assert stats.accepted == 100 and np.all(log == 1)
    u.cleanup()


@pytest.mark.parametrize("pblock", [64, 128, 192])
def test_perform_block_does_not_change_results(pblock):
    s = chains((64, 64, 64), 40, 64, seed=11)
    u = make_updater(s, 5)
    u.setPerformBlock(pblock)
    _, log = u.run(15, log=True)
    x = u.getPositions(s.n)
    u.cleanup()
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=5, n_mcs=15)
    assert np.array_equal(x, ox) and np.array_equal(log, olog)


def test_network_config_bit_exact():
    # Star network on 256^3 at the densest the seeded growth reaches: 7000
    # 4-arm stars of arm length 29 (819000 monomers, occupied fraction 0.39),
    # arm ends crosslinked; 3 species from the coloring.
    s = stars((256, 256, 256), 7000, 4, 29, seed=3)
    print("network", s.n, flush=True)
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=17, n_mcs=6)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)


@pytest.mark.parametrize("kind", KINDS)
def test_plain_config_full_length_bit_exact(kind):
    # The plain single-system case in full: 64 SAW chains of 1024 on 256^3,
    # seed 42, 10^4 MCS, compared with the oracle at the end, on both lattice
    # kinds (uint8 is the reference's default, UpdaterGPUScBFM_AB_Type.cu:12-13).
    s = chains((256, 256, 256), 64, 1024, seed=42)
    print("plain full length", s.n, "kind", kind, flush=True)
    u = make_updater(s, 42, kind=kind)
    u.runSimulationOnGPU(10000)
    gx = u.getPositions(s.n)
    _, acc = u.counters()
    colors = u.colors(s.n)
    u.cleanup()
    from helpers import oracle_for
    o = oracle_for(s, colors, 42)
    _, olog = o.run(0, 10000 * (int(colors.max()) + 1))
    assert np.array_equal(gx, o.xyz)
    assert acc == int(olog.sum())


MELT_BOX = (512, 512, 512)


def _melt():
    # The dense-melt workload at full size: 32768 chains of N=256 (8,388,608
    # monomers) on 512^3, ordered packing at occupied-site fraction 0.5.
    return ordered_chains(MELT_BOX, 32768, 256)


@pytest.mark.parametrize("kind", KINDS)
def test_dense_melt_config_full_size_bit_exact(kind):
    s = _melt()
    print("dense melt full size", s.n, "kind", kind, flush=True)
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=42, n_mcs=3, kind=kind)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)
    assert glog.sum() > 0


def test_dense_melt_relaxed_on_both_lattices_bit_exact():
    # 200 MCS of the full-size melt on the 1-bit and the uint8 lattice end in
    # the same state (the lattice kind does not change a trajectory); 3 more
    # MCS from there on the uint8 lattice match the oracle started from it.
    from helpers import oracle_for
    s = _melt()
    print("dense melt relaxed", s.n, flush=True)
    xs = {}
    for kind in KINDS:
        u = make_updater(s, 42, kind=kind)
        u.runSimulationOnGPU(200)
        xs[kind] = u.getPositions(s.n)
        if kind == P.LATTICE_U8:
            colors = u.colors(s.n)
            t0, _ = u.counters()
            _, glog = u.run(3, log=True)
            gx = u.getPositions(s.n)
        u.cleanup()
    assert np.array_equal(xs[P.LATTICE_BIT], xs[P.LATTICE_U8])
    o = oracle_for(Sys(s.box, xs[P.LATTICE_U8], s.bonds), colors, 42)
    _, olog = o.run(t0, 3 * (int(colors.max()) + 1))
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, o.xyz)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("group,depth,block,slots", [
    (64, 1, 0, 0), (64, 4, 0, 0), (128, 1, 0, 0), (128, 4, 0, 0), (256, 1, 0, 0), (256, 4, 0, 0),
    (P.LEGACY_WALK, 0, 0, 0),
    (64, 4, 64, 1),     # pass-1 lists of 64 entries: the speculative reads are clamped to the list
    (128, 4, 128, 4),
])
def test_perform_walk_does_not_change_results(kind, group, depth, block, slots):
    # a dilute system (occupied fraction 0.04): lists fill to about half their
    # capacity, so a list group walks several rounds of entries
    s = chains((64, 64, 64), 40, 64, seed=11)
    u = make_updater(s, 5, kind=kind, block=block, slots=slots)
    u.setPerformWalk(group, depth)
    _, log = u.run(15, log=True)
    x = u.getPositions(s.n)
    colors = u.colors(s.n)
    u.cleanup()
    from helpers import oracle_for
    o = oracle_for(s, colors, 5)
    _, olog = o.run(0, 15 * (int(colors.max()) + 1))
    assert np.array_equal(log, olog) and np.array_equal(x, o.xyz)


def test_runs_without_read_back_keep_the_last_coordinates():
    # with the read-back off the getters keep the coordinates of the last
    # read-back; readPositionsBack() then returns those of the run alone
    s = chains((64, 64, 64), 40, 64, seed=11)
    u = make_updater(s, 5)
    before = u.getPositions(s.n)
    u.setReadBack(False)
    _, log1 = u.run(7, log=True)
    assert np.array_equal(u.getPositions(s.n), before)
    _, log2 = u.run(8, log=True)
    u.readPositionsBack()
    x = u.getPositions(s.n)
    u.cleanup()
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=5, n_mcs=15)
    assert np.array_equal(np.concatenate([log1, log2]), olog)
    assert np.array_equal(x, ox)


@pytest.mark.parametrize("kind", KINDS)
def test_network_workload_full_size_bit_exact(kind):
    # The star-network workload of the benchmark at full size: 12800 4-arm
    # stars of arm 29 (1,497,600 monomers) with arm ends crosslinked, in a
    # 512x256x256 box; 2 MCS against the oracle.
    s = stars((512, 256, 256), 12800, 4, 29, seed=42)
    print("network full size", s.n, "kind", kind, flush=True)
    from helpers import stage_updater
    partners, escapes = stage_updater(s, 42, kind).partnerEscapes()
    print("partners", partners, "escapes", escapes, flush=True)
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=42, n_mcs=2, kind=kind)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)


@pytest.mark.parametrize("kind", KINDS)
def test_dense_network_full_size_bit_exact(kind):
    # The dense star network of the benchmark (12800 4-arm stars of arm 29,
    # 1,497,600 monomers on 256^3, occupied fraction 0.714): the 100-sweep
    # run of the stars that builds it, on the GPU and on the oracle, then
    # 3 MCS of the crosslinked network (four species) against the oracle.
    import pscbfm_amd.workloads as W
    from helpers import oracle_relax
    stars0 = W.dense_stars(42)
    print("dense stars", stars0.n, "kind", kind, flush=True)
    moved = W.relax_on_gpu(kind)(stars0, W.DENSE_NETWORK["relax_mcs"])
    assert np.array_equal(moved, oracle_relax(stars0, W.DENSE_NETWORK["relax_mcs"]))
    sd = W.crosslinked(stars0, moved, 12800, 29)
    s = Sys(sd.box, sd.xyz, sd.bonds)
    print("dense network bonds", s.bonds.shape[0], flush=True)
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=42, n_mcs=3, kind=kind)
    assert glog.size == 3 * 4   # four species
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)
