"""GPU tests of checkpoint/resume: with the counter draw keyed by the global
substep, a run resumed from a state of age m (setStartMcs, a BFM frame's
!mcs) continues the uninterrupted trajectory bit for bit, in an updater,
through BFM files, through the command-line simulator, and in a batch or the
request loop."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import pscbfm_amd as P
from helpers import Sys, chains, make_updater, oracle_for, stage_updater, stars
from pscbfm_amd.bfm import read_bfm, write_bfm

pytestmark = pytest.mark.gpu
EXE = Path(P.__file__).parent / "pscbfm_run"


def _resumed(s, seed, start, mcs, colors=None, kind=P.LATTICE_BIT, spatial=None):
    u = stage_updater(s, seed, kind, colors=colors, spatial=spatial)
    u.setStartMcs(start)
    u.initialize()
    st, log = u.run(mcs, log=True)
    x = u.getPositions(s.n)
    t, _ = u.counters()
    u.cleanup()
    return x, log, t


@pytest.mark.parametrize("kind", [P.LATTICE_BIT, P.LATTICE_U8])
def test_resumed_updater_continues_the_trajectory(kind):
    s = stars((64, 64, 64), 20, 4, 10, seed=4)
    u = make_updater(s, 31, kind)
    colors = u.colors(s.n)
    nc = int(colors.max()) + 1
    _, log_a = u.run(12, log=True)
    mid = u.getPositions(s.n)
    _, log_b = u.run(9, log=True)
    end = u.getPositions(s.n)
    u.cleanup()
    x, log, t = _resumed(Sys(s.box, mid, s.bonds), 31, 12, 9, colors, kind)
    assert t == 21 * nc
    assert np.array_equal(log, log_b) and np.array_equal(x, end)
    # and the oracle, started at substep 12 * nColors from the same state
    o = oracle_for(Sys(s.box, mid, s.bonds), colors, 31)
    _, olog = o.run(12 * nc, 9 * nc)
    assert np.array_equal(olog, log_b) and np.array_equal(o.xyz, end)


@pytest.mark.parametrize("first_spatial", [True, False])
def test_resume_across_slot_orders(first_spatial):
    # The slot order is invisible in results: a run in one order, resumed
    # from its coordinates in the other, continues the uninterrupted run.
    s = stars((64, 64, 64), 20, 4, 10, seed=5)
    u = make_updater(s, 13, spatial=first_spatial)
    colors = u.colors(s.n)
    u.run(7)
    mid = u.getPositions(s.n)
    _, log_b = u.run(8, log=True)
    end = u.getPositions(s.n)
    u.cleanup()
    x, log, _ = _resumed(Sys(s.box, mid, s.bonds), 13, 7, 8, colors, spatial=not first_spatial)
    assert np.array_equal(log, log_b) and np.array_equal(x, end)


def test_bfm_checkpoint_round_trip(tmp_path):
    s = chains((64, 64, 64), 40, 32, seed=6)
    u = make_updater(s, 8)
    u.run(15)
    u.writeBfm(tmp_path / "ck.bfm", 15)
    u.run(10)
    end = u.getPositions(s.n)
    u.cleanup()
    v = P.UpdaterGPUScBFM_AB_Type()
    v.setGpu(0)
    age = v.loadBfm(tmp_path / "ck.bfm")
    assert age == 15
    v.setSeed(8)
    v.setStartMcs(age)
    v.initialize()
    v.run(10)
    assert np.array_equal(v.getPositions(s.n), end)
    v.cleanup()


def test_cli_resumed_from_its_own_output_equals_one_run(tmp_path):
    s = chains((64, 64, 64), 40, 64, seed=11)
    src = tmp_path / "in.bfm"
    write_bfm(src, s.box, s.xyz, s.bonds, mcs=100)
    one, part = tmp_path / "one.bfm", tmp_path / "part.bfm"
    run = lambda *a: subprocess.run([str(EXE), *map(str, a)], capture_output=True, text=True, timeout=120)
    print("cli: one run of 30", flush=True)
    r = run("-i", src, "-m", 30, "-s", 10, "-o", one, "--seed", 5)
    assert r.returncode == 0, r.stderr
    print("cli: 20, then 10 resumed from the last frame", flush=True)
    r = run("-i", src, "-m", 20, "-s", 10, "-o", part, "--seed", 5)
    assert r.returncode == 0, r.stderr
    r = run("-i", part, "-m", 10, "-s", 10, "-o", part, "--seed", 5)  # READ_LAST_CONFIG_SAVE, append
    assert r.returncode == 0, r.stderr
    a, b = read_bfm(one), read_bfm(part)
    assert [f[0] for f in a.frames] == [100, 110, 120, 130]
    assert [f[0] for f in b.frames][-1] == 130
    assert np.array_equal(a.frames[-1][1], b.frames[-1][1])


def test_batch_and_server_start_replicas_at_their_age():
    s = chains((32, 32, 32), 20, 16, seed=7)
    x_ref, log_ref, _ = _resumed(s, 42, 40, 12)
    u = stage_updater(s, 42)
    u.setStartMcs(40)
    b = P.Batch(tile_capacity=64, replica_capacity=4)
    rid = b.add(u)
    other = b.add(stage_updater(chains((64, 64, 64), 30, 32, seed=2), 3))
    b.run(12)
    assert np.array_equal(b.positions(rid, s.n), x_ref)
    assert b.info(rid).accepted == int(log_ref.sum())
    srv = P.Server(hbm_budget=1 << 32, tile_capacity=64, replica_capacity=4, quantum=8)
    job = srv.submit(u, 12)
    srv.drain()
    assert np.array_equal(srv.result(job, s.n), x_ref)
    srv.close()
