"""The refined region the deck driver checks particles against: the finest
level's box under bmx.tag_region tagging (src/setup/bmx_tagging.cpp), the
error buffer, the refinement ratio and the blocking factor, worked by hand
for the shipped decks.

exec/fungi/input_fungi: n_cell 16 16 8 over 0.128 x 0.128 x 0.064, two
levels of ratio 2, n_error_buf 2, blocking_factor 2, tagged slab
z in [0.04, 0.052].  Level 0 (dz 0.008): centres 0.044, 0.052 tagged (cells
5, 6); grown to 3..8, within 0..7: 3..7; level 1 cells 6..15.  Level 1
(dz 0.004): centres 0.042 .. 0.05 tagged (10..12); grown to 8..14; level 2
cells 16..29, z in [0.032, 0.060).  x and y: the whole domain.
exec/nlev_large_test/input_real_nlev: n_cell 32 32 16 over 0.0128 x 0.0128 x
0.0064, one level of ratio 2, n_error_buf 2, blocking_factor 2, tagged slab
z in [0.0044, 0.0052].  Level 0 (dz 0.0004): centres 0.0046, 0.005 tagged
(cells 11, 12); grown to 9..14; level 1 cells 18..29, z in [0.0036, 0.006)."""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

CSRC = os.path.join(ROOT, "boltzmannmfx_amd", "csrc")


@pytest.fixture(scope="module")
def region(bmx, tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.fail("g++ is needed")
    bmx.load_library()  # built, and loadable without a device
    libdir = os.path.dirname(bmx.LIB_PATH)
    exe = str(tmp_path_factory.mktemp("region") / "region")
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", CSRC, "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "deck", "region.cpp"), "-o", exe,
                    "-L", libdir, "-l:" + os.path.basename(bmx.LIB_PATH),
                    "-Wl,-rpath," + libdir], check=True)

    def run(deck, levels=False):
        args = [exe, os.path.join(GOLDEN, "ref_decks", deck)] + (["levels"] if levels else [])
        out = subprocess.run(args, check=True, capture_output=True, text=True).stdout
        if levels:
            rows = [line.split() for line in out.strip().splitlines()]
            return [([float(v) for v in r[:6]], [int(v) for v in r[6:9]],
                     [int(v) for v in r[9:]]) for r in rows]
        out = out.split()
        return [float(v) for v in out[:6]], [int(v) for v in out[6:]]
    return run


def test_fungi_deck_region(region):
    box, n_cell = region("fungi/input_fungi")
    assert n_cell == [64, 64, 32]
    assert box == pytest.approx([0.0, 0.0, 0.032, 0.128, 0.128, 0.060], abs=1e-15)


def test_nlev_deck_region(region):
    box, n_cell = region("nlev_large_test/input_real_nlev")
    assert n_cell == [64, 64, 32]
    assert box == pytest.approx([0.0, 0.0, 0.0036, 0.0128, 0.0128, 0.006], abs=1e-15)


def test_nlev_deck_levels(region):
    """The two levels the driver runs particles on when some lie outside the
    refined slab: level 0 is the domain on amr.n_cell (periodic in x and y);
    level 1 is the slab on the refined cells, periodic in x and y because it
    spans the domain there, not in z."""
    levels = region("nlev_large_test/input_real_nlev", levels=True)
    assert len(levels) == 2
    box0, n0, per0 = levels[0]
    assert box0 == pytest.approx([0.0, 0.0, 0.0, 0.0128, 0.0128, 0.0064], abs=1e-15)
    assert n0 == [32, 32, 16] and per0 == [1, 1, 0]
    box1, n1, per1 = levels[1]
    assert box1 == pytest.approx([0.0, 0.0, 0.0036, 0.0128, 0.0128, 0.006], abs=1e-15)
    # (0.006 - 0.0036) / 2e-4 = 12 refined cells in z
    assert n1 == [64, 64, 12] and per1 == [1, 1, 0]
    # the last level is the finest region
    fin, _ = region("nlev_large_test/input_real_nlev")
    assert fin == pytest.approx(box1, abs=1e-15)


def test_fungi_deck_levels(region):
    levels = region("fungi/input_fungi", levels=True)
    fin, n_fin = region("fungi/input_fungi")
    assert levels[0][0][:3] == [0.0, 0.0, 0.0]
    assert levels[-1][0] == pytest.approx(fin, abs=1e-15)
    # the finest level's cells are those of the finest mesh
    dz = (fin[5] - fin[2]) / levels[-1][1][2]
    assert dz == pytest.approx(levels[0][0][5] / n_fin[2], rel=1e-12)
