"""The allocation-free run entry point (pscbfm_run_ex_noalloc) against
pscbfm_run_ex, bit for bit, through the C ABI.

Two calls in a row share the GPU's scratch array, the second with a smaller
system and fewer sweeps, so the second must clear what the first left there;
consecutive calls on one handle continue its trajectory; R250 mode reads its
draws from the scratch array; a call beyond the scratch maximum is refused
before any device work.
"""
import numpy as np
import pytest

import oracle
import pscbfm_amd as P
from helpers import GOLDEN, chains, make_updater, oracle_for

pytestmark = pytest.mark.gpu


def _pair(s, seed, kind):
    return make_updater(s, seed, kind=kind), make_updater(s, seed, kind=kind)


@pytest.mark.parametrize("kind", [P.LATTICE_BIT, P.LATTICE_U8])
def test_noalloc_equals_run_ex_on_two_calls_in_a_row(kind):
    big = chains((64, 64, 64), 40, 64, seed=11)
    small = chains((32, 32, 32), 12, 16, seed=4)
    for s, seed, n_mcs in ((big, 5, 30), (small, 9, 7)):
        print(f"noalloc kind {kind}: {s.n} monomers, {n_mcs} MCS", flush=True)
        a, b = _pair(s, seed, kind)
        sa, la = a.runNoAlloc(n_mcs, log=True)
        sb, lb = b.run(n_mcs, log=True)
        xa, xb = a.getPositions(s.n), b.getPositions(s.n)
        assert np.array_equal(la, lb) and la.sum() > 0
        assert np.array_equal(xa, xb)
        assert (sa.accepted, sa.attempted, sa.substeps, sa.bond_terms) == \
            (sb.accepted, sb.attempted, sb.substeps, sb.bond_terms)
        assert a.counters() == b.counters()
        colors = a.colors(s.n)
        o = oracle_for(s, colors, seed)
        _, olog = o.run(0, n_mcs * (int(colors.max()) + 1))
        assert np.array_equal(la, olog) and np.array_equal(xa, o.xyz)
        a.cleanup()
        b.cleanup()


def test_noalloc_calls_continue_the_trajectory():
    s = chains((64, 64, 64), 30, 40, seed=2)
    a, b = _pair(s, 8, P.LATTICE_BIT)
    logs_a, logs_b = [], []
    for n_mcs, with_log in ((20, True), (3, False), (7, True)):
        sa, la = a.runNoAlloc(n_mcs, log=with_log)
        sb, lb = b.run(n_mcs, log=with_log)
        assert sa.accepted == sb.accepted
        if with_log:
            assert np.array_equal(la, lb)
    assert np.array_equal(a.getPositions(s.n), b.getPositions(s.n))
    a.cleanup()
    b.cleanup()


def test_noalloc_r250_mode_equals_run_ex():
    s = chains((32, 32, 32), 24, 16, seed=1)
    a, b = _pair(s, 2024, P.LATTICE_BIT)
    for u in (a, b):
        u.loadR250SeedFile(GOLDEN / "resultNormPscBFM.seeds", 2)
    for n_mcs in (12, 5):
        _, la = a.runNoAlloc(n_mcs, log=True)
        _, lb = b.run(n_mcs, log=True, use_graph=False)
        assert np.array_equal(la, lb)
    assert np.array_equal(a.getPositions(s.n), b.getPositions(s.n))
    assert np.array_equal(a.r250State(), b.r250State())
    a.cleanup()
    b.cleanup()


def test_noalloc_refuses_runs_beyond_the_scratch():
    s = chains((32, 32, 32), 12, 16, seed=4)
    u = make_updater(s, 3)
    nc = u.nColors()
    before = u.counters()
    too_many = (1 << 18) // nc + 1
    with pytest.raises(P.InvalidArgument, match="scratch maximum"):
        u.runNoAlloc(too_many, log=True)
    assert u.counters() == before
    # the largest call that fits runs
    st, log = u.runNoAlloc((1 << 18) // nc // 64, log=True)
    assert log.size == st.substeps and int(log.sum()) == st.accepted
    u.cleanup()
