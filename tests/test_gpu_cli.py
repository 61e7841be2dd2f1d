"""GPU test of the command-line simulator (the reference's main with -i -m -s
-o -g -e): every frame it appends equals the updater's state after the same
number of sweeps with the same seed."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
import pscbfm_amd as P
from helpers import GOLDEN, chains
from pscbfm_amd.bfm import read_bfm, write_bfm

pytestmark = pytest.mark.gpu
EXE = Path(P.__file__).parent / "pscbfm_run"


def _updater_from(path, seed, kind=P.LATTICE_BIT):
    u = P.UpdaterGPUScBFM_AB_Type()
    u.setGpu(0)
    u.setLatticeKind(kind)
    mcs = u.loadBfm(path)
    u.setSeed(seed)
    u.setStartMcs(mcs)  # the simulator continues the draws from the frame's age
    u.initialize()
    return u, mcs


@pytest.mark.parametrize("lattice,kind", [("bit", P.LATTICE_BIT), ("u8", P.LATTICE_U8)])
def test_cli_frames_equal_the_updater(tmp_path, lattice, kind):
    s = chains((64, 64, 64), 40, 64, seed=11)
    src = tmp_path / "in.bfm"
    write_bfm(src, s.box, s.xyz, s.bonds, mcs=100)
    out = tmp_path / "out.bfm"
    print(f"cli run, lattice {lattice}", flush=True)
    res = subprocess.run([str(EXE), "-i", str(src), "-m", "20", "-s", "10", "-o", str(out), "-g", "0",
                          "--seed", "42", "--lattice", lattice], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    got = read_bfm(out)
    assert [f[0] for f in got.frames] == [100, 110, 120]
    u, mcs0 = _updater_from(src, 42, kind)
    assert mcs0 == 100 and np.array_equal(got.frames[0][1], s.xyz)
    for k in (1, 2):
        u.run(10)
        assert np.array_equal(got.frames[k][1], u.getPositions(s.n))
    u.cleanup()


def test_cli_seed_file_and_reference_fixture(tmp_path):
    # the reference's 64 chains of 1024 with its own seed file (-e): the substep draws are the
    # R250 stream of the file after the two draws the reference's main prints
    src = GOLDEN / "LinearChain_l1024_n64.bfm.gz"
    seeds = GOLDEN / "resultNormPscBFM.seeds"
    out = tmp_path / "lc.bfm"
    print("cli run with a seed file", flush=True)
    res = subprocess.run([str(EXE), "-e", str(seeds), "-i", str(src), "-m", "4", "-s", "2", "-o", str(out)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    # mainCUDASimulatorGPUScBFM_AB_Type.cpp:138-139
    assert "RNG rand: 4223731124" in res.stderr and "RNG rand: 0.803876" in res.stderr
    got = read_bfm(out)
    assert len(got.frames) == 3
    u = P.UpdaterGPUScBFM_AB_Type()
    u.setGpu(0)
    u.setStartMcs(u.loadBfm(src))
    u.initialize()
    u.loadR250SeedFile(seeds, 2)
    for k in (1, 2):
        u.run(2)
        assert np.array_equal(got.frames[k][1], u.getPositions(got.n))
    colors = u.colors(got.n)
    u.cleanup()
    # and the oracle, fed the same R250 draws from the same start
    start, ids = _csr(got.n, got.bonds)
    o = oracle.System(got.box, got.frames[0][1], start, ids, colors, oracle.standard_bondset(), 0)
    r = oracle.R250.from_seed_file(seeds)
    r.rand32(), r.rand32()
    nc = int(colors.max()) + 1
    o.draws = r.substep_draws(4 * nc, nc)
    o.run(0, 4 * nc)
    assert np.array_equal(o.xyz, got.xyz)


def _csr(n, bonds):
    pairs = np.concatenate([bonds, bonds[:, ::-1]]) if len(bonds) else np.zeros((0, 2), np.int64)
    pairs = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]
    start = np.zeros(n + 1, dtype=np.uint32)
    np.add.at(start, pairs[:, 0] + 1, 1)
    return np.cumsum(start).astype(np.uint32), pairs[:, 1].astype(np.uint32)
