"""GPU test of the command-line simulator (the reference's main with -i -m -s
-o -g -e): every frame it appends equals the updater's state after the same
number of sweeps with the same seed."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import pscbfm_amd as P
from helpers import GOLDEN, chains
from pscbfm_amd.bfm import read_bfm, write_bfm

pytestmark = pytest.mark.gpu
EXE = Path(P.__file__).parent / "pscbfm_run"


def _updater_from(path, seed):
    u = P.UpdaterGPUScBFM_AB_Type()
    u.setGpu(0)
    mcs = u.loadBfm(path)
    u.setSeed(seed)
    u.initialize()
    return u, mcs


def test_cli_frames_equal_the_updater(tmp_path):
    s = chains((64, 64, 64), 40, 64, seed=11)
    src = tmp_path / "in.bfm"
    write_bfm(src, s.box, s.xyz, s.bonds, mcs=100)
    out = tmp_path / "out.bfm"
    print("cli run", flush=True)
    res = subprocess.run([str(EXE), "-i", str(src), "-m", "20", "-s", "10", "-o", str(out), "-g", "0",
                          "--seed", "42"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    got = read_bfm(out)
    assert [f[0] for f in got.frames] == [100, 110, 120]
    u, mcs0 = _updater_from(src, 42)
    assert mcs0 == 100 and np.array_equal(got.frames[0][1], s.xyz)
    for k in (1, 2):
        u.run(10)
        assert np.array_equal(got.frames[k][1], u.getPositions(s.n))
    u.cleanup()


def test_cli_seed_file_and_reference_fixture(tmp_path):
    # the reference's 64 chains of 1024 with its own seed file (-e): first two words key the draws
    src = GOLDEN / "LinearChain_l1024_n64.bfm.gz"
    seeds = GOLDEN / "resultNormPscBFM.seeds"
    out = tmp_path / "lc.bfm"
    res = subprocess.run([str(EXE), "-e", str(seeds), "-i", str(src), "-m", "4", "-s", "2", "-o", str(out)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    key = int.from_bytes(seeds.read_bytes()[:8], "little")
    u, _ = _updater_from(src, key)
    u.run(4)
    got = read_bfm(out)
    assert len(got.frames) == 3 and np.array_equal(got.xyz, u.getPositions(got.n))
    u.cleanup()
