"""GPU parity on edge shapes, bit for bit against the oracle: the
reference's own ring and hydrogel fixtures (run in periodic boxes of their
size), degree-8 networks (the widest neighbour table), the smallest, the largest
(2^32 cells) and strongly non-cubic boxes, and coordinates unfolded far from
the box."""
import numpy as np
import pytest

import pscbfm_amd as P
from helpers import GOLDEN, Sys, chains, gpu_vs_oracle, stars
from pscbfm_amd.bfm import read_bfm

pytestmark = pytest.mark.gpu

KINDS = [P.LATTICE_BIT, P.LATTICE_U8]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("name", ["Ring_N6_n6.bfm", "Hydrogel_NStar_117_NoPerXYZ64.bfm"])
def test_reference_fixture_bit_exact(name, kind):
    # tests/Ring_N6_n6.bfm (rings of six) and the degree-6 hydrogel of
    # tests/coloring-test-Hydrogel_*.bfm; both are non-periodic files whose
    # monomers stay clear of the faces, so they are valid periodic states too
    b = read_bfm(GOLDEN / name)
    s = Sys(b.box, b.xyz, b.bonds)
    print(f"fixture {name} kind {kind}: {s.n} monomers", flush=True)
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=11, n_mcs=300, kind=kind)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)
    assert glog.sum() > 0


def test_degree_eight_stars_bit_exact():
    s = stars((64, 64, 64), 20, 8, 10, seed=3)
    assert np.bincount(s.bonds.ravel(), minlength=s.n).max() == 8
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=5, n_mcs=100)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("box,n_chains,length", [((8, 8, 8), 2, 4), ((8, 16, 64), 6, 16), ((128, 8, 32), 10, 20)])
def test_small_and_non_cubic_boxes(box, n_chains, length, kind):
    s = chains(box, n_chains, length, seed=5)
    print(f"box {box} kind {kind}: {s.n} monomers", flush=True)
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=9, n_mcs=300, kind=kind)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)
    # monomers crossed the periodic faces, so the image counters were used
    assert np.any(gx.min(axis=0) < 0) or np.any(gx.max(axis=0) >= np.array(box))


def test_far_unfolded_coordinates_shift_the_result():
    s = chains((32, 32, 32), 12, 16, seed=4)
    shift = np.array([1000 * 32, -2000 * 32, 7 * 32], dtype=np.int32)
    far = Sys(s.box, s.xyz + shift, s.bonds)
    gx, glog, _, _ = gpu_vs_oracle(s, seed=21, n_mcs=100)
    fx, flog, ox, olog = gpu_vs_oracle(far, seed=21, n_mcs=100)
    assert np.array_equal(flog, glog) and np.array_equal(flog, olog)
    assert np.array_equal(fx, gx + shift) and np.array_equal(fx, ox)


def test_largest_box_across_the_wrapping_corner():
    # 2048 x 2048 x 1024 = 2^32 cells, the most a 32-bit cell code spans; the
    # chains straddle the corner where all three axes wrap, so the top code
    # bits and the periodic images are in play on every move
    small = chains((32, 32, 32), 12, 16, seed=4)
    box = (2048, 2048, 1024)
    shift = np.array([box[0] - 16, box[1] - 16, box[2] - 16], dtype=np.int32)
    s = Sys(box, small.xyz + shift, small.bonds)
    print(f"largest box {box}: {s.n} monomers", flush=True)
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=13, n_mcs=40)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)
    assert glog.sum() > 0


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("spatial", [True, False])
def test_crosslinked_network_with_scattered_partners(kind, spatial):
    # 1000 four-arm stars grown into a connected network: a large share of the
    # bonds join monomers whose slots lie far from the row's common partner
    # base, so the uint8 kernel's 16-bit partner codes fall back to the full
    # neighbour matrix for them (in either slot order; most in monomer order)
    s = stars((128, 128, 128), 1000, 4, 29, seed=7)
    print(f"network kind {kind} spatial {spatial}: {s.n} monomers", flush=True)
    from helpers import stage_updater
    partners, escapes = stage_updater(s, 17, kind, spatial=spatial).partnerEscapes()
    assert 0 < escapes < partners  # the escape path is taken
    gx, glog, ox, olog = gpu_vs_oracle(s, seed=17, n_mcs=30, kind=kind, spatial=spatial)
    assert np.array_equal(glog, olog)
    assert np.array_equal(gx, ox)
    assert glog.sum() > 0
