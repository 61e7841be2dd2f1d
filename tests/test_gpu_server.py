"""GPU tests of the request loop: jobs submitted while others run are
admitted into the shared batch as the budget allows, and each job's result
is the one its system has alone."""
import threading

import numpy as np
import pytest

import pscbfm_amd as P
from helpers import chains, make_updater, oracle_for, ordered_chains, stage_updater, stars

pytestmark = pytest.mark.gpu


def _alone(s, seed, mcs, kind=P.LATTICE_BIT):
    u = make_updater(s, seed, kind)
    colors = u.colors(s.n)
    u.run(mcs)
    x = u.getPositions(s.n)
    acc = u.counters()[1]
    u.cleanup()
    return x, acc, colors


def _cpu(s, seed, mcs, colors):
    """The same system's single-system run on the CPU oracle."""
    o = oracle_for(s, colors, seed)
    total, _ = o.run(0, mcs * (int(colors.max()) + 1))
    return o.xyz, total


def _jobs():
    return [
        (chains((32, 32, 32), 20, 16, seed=7), 42, 30),
        (chains((64, 64, 64), 40, 64, seed=11), 5, 12),
        (stars((64, 64, 64), 40, 4, 29, seed=9), 77, 20),
        (ordered_chains((32, 32, 32), 64, 32), 1, 25),
        (chains((32, 64, 128), 30, 32, seed=3), 9, 40),
        (chains((16, 16, 16), 1, 1, seed=2), 8, 7),
    ]


@pytest.mark.parametrize("kind", [P.LATTICE_BIT, P.LATTICE_U8])
def test_each_served_job_equals_its_run_alone(kind):
    jobs = _jobs()
    print(f"served jobs, lattice kind {kind}", flush=True)
    srv = P.Server(hbm_budget=1 << 34, tile_capacity=256, replica_capacity=16, lattice_kind=kind, quantum=8)
    ids = []
    # submit from two threads while the loop is already running
    def submit(part):
        for s, seed, mcs in part:
            ids.append((srv.submit(stage_updater(s, seed), mcs), s, seed, mcs))
    t = threading.Thread(target=submit, args=(jobs[3:],))
    submit(jobs[:3])
    t.start()
    t.join()
    srv.drain()
    st = srv.stats()
    assert st.done == len(jobs) and st.failed == 0 and st.running == 0 and st.resident_bytes == 0
    for job, s, seed, mcs in ids:
        info = srv.job(job)
        assert info.state == P.JOB_DONE and info.t_submit <= info.t_admit <= info.t_done
        x_alone, acc_alone, colors = _alone(s, seed, mcs, kind)
        x_cpu, acc_cpu = _cpu(s, seed, mcs, colors)
        served = srv.result(job, s.n)
        assert np.array_equal(served, x_alone) and np.array_equal(served, x_cpu)
        assert info.accepted == acc_alone == acc_cpu
        srv.release(job)
    srv.close()


def test_budget_limits_residency_and_queues_the_rest():
    jobs = _jobs()[:4]
    staged = [stage_updater(s, seed) for s, seed, _ in jobs]
    probe = P.Batch(0, 1)
    sizes = [probe.footprint(u)[0] for u in staged]
    del probe
    budget = max(sizes) + min(sizes)  # at most two resident at a time
    srv = P.Server(hbm_budget=budget, tile_capacity=256, replica_capacity=16, quantum=4)
    ids = [srv.submit(u, mcs) for u, (_, _, mcs) in zip(staged, jobs)]
    srv.drain()
    st = srv.stats()
    assert st.done == len(jobs) and st.peak_resident_bytes <= budget
    for job, (s, seed, mcs) in zip(ids, jobs):
        assert np.array_equal(srv.result(job, s.n), _alone(s, seed, mcs)[0])
    srv.close()


def test_job_larger_than_the_budget_fails_alone():
    s, seed, mcs = _jobs()[1]
    srv = P.Server(hbm_budget=1 << 16, tile_capacity=64, replica_capacity=4)
    job = srv.submit(stage_updater(s, seed), mcs)
    assert srv.wait(job, 30.0)
    with pytest.raises(RuntimeError):
        srv.job(job)
    srv.close()


def test_long_job_beside_short_ones_equals_its_run_alone_and_the_oracle():
    # a job of 20,000 sweeps shares the batch with short jobs that arrive,
    # run and retire while it runs; its result is the one it has alone and
    # on the CPU oracle
    long_job = (chains((32, 32, 64), 16, 24, seed=13), 2024, 20000)
    short = [(chains((32, 32, 32), 10, 16, seed=k), 100 + k, 50 + 10 * k) for k in range(6)]
    srv = P.Server(hbm_budget=1 << 34, tile_capacity=256, replica_capacity=16, quantum=64)
    s, seed, mcs = long_job
    lid = srv.submit(stage_updater(s, seed), mcs)
    sids = [(srv.submit(stage_updater(x, sd), m), x, sd, m) for x, sd, m in short]
    srv.drain()
    for jid, x, sd, m in sids:
        assert np.array_equal(srv.result(jid, x.n), _alone(x, sd, m)[0])
    got = srv.result(lid, s.n)
    x, acc, colors = _alone(s, seed, mcs)
    assert np.array_equal(got, x)
    assert srv.job(lid).accepted == acc
    ox, oacc = _cpu(s, seed, mcs, colors)
    assert np.array_equal(got, ox) and acc == oacc
    srv.close()
